// CDDTCast / PCDDT (RangeLib.h:971-1770) built on the GPU.
//
// Included at the end of rl_lib.hip (uses its types and helpers).  The query
// is rl::cddt_calc_range in rl_device.h; this file builds and prunes the
// tables:
//   1. per-angle constants (one thread per angle, the library's sinf/cosf);
//   2. edge cells (occupied with a free 8-neighbour, OMap::make_edge_map);
//   3. every (edge cell, angle < td/2) pair projects the cell centre into the
//      bins its footprint covers: a counting pass (atomics), an exclusive scan
//      into bin offsets, and a filling pass;
//   4. each bin sorted (segmented bitonic sort, rl_scan_sort.h) and its
//      duplicates dropped;
//   5. prune: every free cell marks the entries a query from it can return,
//      unmarked entries are dropped.
// The result equals the reference's compressed_lut bin for bin: the order in
// which entries arrive does not matter after sorting and deduplication.

#define RL_CDDT_EPS 0.00001

// 1. Constants of angle a (RangeLib.h:991-1058).
__global__ void k_cddt_consts(TrigDev trig, int td, int W, int H, float m2pi_div_td, int* __restrict__ width,
                              float* __restrict__ trans, float* __restrict__ cosv, float* __restrict__ sinv) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= td) return;
  const float angle = (float)a * m2pi_div_td;
  float sn, cs;
  rl::sincos_libc(angle, trig, &sn, &cs);
  cosv[a] = cs;
  sinv[a] = sn;
  const float rotated_height = fabsf((float)W * sn) + fabsf((float)H * cs);
  width[a] = (int)(unsigned)ceil((double)rotated_height - RL_CDDT_EPS);
  const float lt = (float)H * cs, rt = (float)W * sn + (float)H * cs, rb = (float)W * sn;
  const float m2 = (rb < rt) ? rb : rt;  // std::min(rt, rb)
  const float mn = (m2 < lt) ? m2 : lt;  // std::min(lt, ...)
  const double t = -1.0 * (double)mn - RL_CDDT_EPS;
  trans[a] = (float)(t > 0.0 ? t : 0.0);  // std::max(0.0, t)
}

// 2. Edge cells, compacted into `list` (one atomic per wave).
__global__ void k_cddt_edges(const uint8_t* __restrict__ occ, int W, int H, int* __restrict__ list,
                             unsigned* __restrict__ n_out) {
  const unsigned i = blockIdx.x * blockDim.x + threadIdx.x;  // W * H < 2^31 (checked by the caller)
  bool e = false;
  if (i < (unsigned)W * (unsigned)H) {
    const int x = (int)(i / (unsigned)H), y = (int)(i - (unsigned)x * (unsigned)H);
    if (occ[i]) {
      const int dx[8] = {1, -1, 0, 0, 1, -1, 1, -1}, dy[8] = {0, 0, 1, -1, 1, 1, -1, -1};
      for (int k = 0; k < 8; ++k) {
        const int cx = x + dx[k], cy = y + dy[k];
        if (0 <= cx && 0 <= cy && cx < W && cy < H && !occ[(size_t)cx * H + cy]) {
          e = true;
          break;
        }
      }
    }
  }
  const unsigned long long mask = __ballot(e);
  const int lane = threadIdx.x & 63;
  unsigned base = 0;
  if (lane == 0 && mask) base = atomicAdd(n_out, (unsigned)__popcll(mask));
  base = __shfl(base, 0);
  if (e) list[base + __popcll(mask & ((1ull << lane) - 1ull))] = (int)i;
}

// Footprint of edge cell `cell` (x * H + y) at angle a (RangeLib.h:1096-1115):
// the projected centre and the bins [lower, upper] it is stored in.
struct CddtProj {
  float cx;
  int lower, upper;
};
__device__ __forceinline__ CddtProj cddt_project(int cell, int H, float cs, float sn, float tr) {
  const int x = cell / H, y = cell - (cell / H) * H;
  const float px = (float)(x + 0.5), py = (float)(y + 0.5);
  const float half = (float)((double)(fabsf(sn) + fabsf(cs)) / 2.0);
  CddtProj p;
  p.cx = px * cs - py * sn;
  const float cy = (px * sn + py * cs) + tr;
  p.upper = (int)((double)(cy + half) - RL_CDDT_EPS);
  p.lower = (int)((double)(cy - half) + RL_CDDT_EPS);
  return p;
}

// 3. Counting (fill == false) or filling (fill == true) pass; one thread per
// (edge cell, angle): angle a0 + blockIdx.y, edge cells along blockIdx.x (the
// edge list holds runs of neighbouring cells, one run per wave of
// k_cddt_edges), so a wave's lanes mostly project nearby cells under one angle.
// Bins outside [0, width) are skipped (the reference would write outside its
// table; it does not happen for maps whose edge cells project inside the
// rotated bounding box, which the translation and width are built to cover).
template <bool FILL>
__global__ __launch_bounds__(256) void k_cddt_scatter(const int* __restrict__ edges, int n_edges, int a0, int H,
                                                      const int* __restrict__ width, const float* __restrict__ trans,
                                                      const float* __restrict__ cosv, const float* __restrict__ sinv,
                                                      const long long* __restrict__ bin_off,
                                                      unsigned long long* __restrict__ counter,
                                                      float* __restrict__ vals) {
  const int e = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (e >= n_edges) return;
  const int a = a0 + (int)blockIdx.y;
  const CddtProj p = cddt_project(edges[e], H, cosv[a], sinv[a], trans[a]);
  for (int b = p.lower; b <= p.upper; ++b) {
    if (b < 0 || b >= width[a]) continue;
    const unsigned long long pos = atomicAdd(&counter[bin_off[a] + b], 1ull);
    if (FILL) vals[pos] = p.cx;
  }
}

// 4./5. Per bin: count the entries kept (COPY == false) or copy them to their
// new place (COPY == true).  mode 0 keeps the first of each run of equal
// values (after sorting), mode 1 keeps the entries marked used.
template <bool COPY>
__global__ void k_cddt_compact(long long nb, const long long* __restrict__ start, const float* __restrict__ vals,
                               const uint8_t* __restrict__ used, int mode,
                               const long long* __restrict__ new_start, long long* __restrict__ kept,
                               float* __restrict__ out) {
  const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const long long s = start[b], e = start[b + 1];
  long long o = COPY ? new_start[b] : 0, k = 0;
  for (long long j = s; j < e; ++j) {
    const bool keep = mode == 0 ? (j == s || vals[j] != vals[j - 1]) : used[j] != 0;
    if (!keep) continue;
    if (COPY) out[o++] = vals[j];
    ++k;
  }
  if (!COPY) kept[b] = k;
}

// 5. Prune marks (RangeLib.h:1172-1275), one thread per (cell, angle): angle
// a = blockIdx.y, cells x-major along blockIdx.x, so the lanes of a wave are
// neighbouring cells of one column under one angle and look up neighbouring
// bins.  Where the reference's linear search leaves its index unset (no entry
// at or beyond the cell), the lower_bound value is taken, as its binary search.
// W * H < 2^31 (checked when the tables are built).
__global__ __launch_bounds__(256) void k_cddt_prune_mark(const uint8_t* __restrict__ occ, int W, int H, float max_range,
                                                         const int* __restrict__ width, const float* __restrict__ trans,
                                                         const float* __restrict__ cosv, const float* __restrict__ sinv,
                                                         const long long* __restrict__ bin_off,
                                                         const long long* __restrict__ start,
                                                         const float* __restrict__ vals, uint8_t* __restrict__ used,
                                                         int a0) {
  const unsigned cell = blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= (unsigned)W * (unsigned)H) return;
  if (occ[cell]) return;  // (every exit below leaves no mark, so the order of the tests is free)
  const int a = a0 + (int)blockIdx.y;
  const int x = (int)(cell / (unsigned)H), y = (int)(cell - (unsigned)x * (unsigned)H);
  const float fx = (float)(0.5 + x), fy = (float)(0.5 + y);
  const float cs = cosv[a], sn = sinv[a];
  const float lx = fx * cs - fy * sn;
  const float ly = (fx * sn + fy * cs) + trans[a];
  const unsigned li = (unsigned)rl::trunc_x86(ly);
  if (li >= (unsigned)width[a]) return;
  const long long b = bin_off[a] + li, s = start[b];
  const int n = (int)(start[b + 1] - s);  // fewer than 2^31 entries in all (checked at build)
  const float* v = vals + s;
  uint8_t* u = used + s;
  const int high = n - 1;
  if (high == -1) return;
  if (v[high] < lx && lx - v[high] < max_range) {
    u[high] = 1;
    return;
  }
  int lo = 0, hi = n;  // lower_bound
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (v[mid] < lx) lo = mid + 1; else hi = mid;
  }
  if (lo < n) u[lo] = 1;
  if (lo >= 1) u[lo - 1] = 1;
}

// RangeMethod::calc_range_many_radial_optimized (RangeLib.h:616-676), one
// thread per output slot s of particle p.  The reference loops a = 0..mpi
// writing pair(a).r to slot a and pair(a).r_inv to slot a + off, then a =
// mpi+1..off-1 writing calc_range to slot a, the beam angle growing by `step`
// each iteration; a slot keeps its last write.  So slot s is: the single ray
// if mpi < s < off; else pair(s - off).r_inv if that iteration exists and
// either off <= 0 or s > mpi; else pair(s).r if s <= mpi; else untouched.
// Pairs exist for CDDTCast only (-1, -1 for the other methods, RangeLib.h:419).
template <int METHOD>
__global__ __launch_bounds__(256) void k_radial(MapDev m, TrigDev trig, float max_range, const float* __restrict__ ins,
                                                float* __restrict__ outs, int N, int nr, float min_angle, float step,
                                                int mpi, int off) {
  __shared__ float lut[RL_LUT_SIZE];
  stage_lut<METHOD>(m, lut);
  const unsigned r = blockIdx.x * blockDim.x + threadIdx.x;  // N * nr < 2^31 (checked by the caller)
  if (r >= (unsigned)N * (unsigned)nr) return;
  const int p = (int)(r / (unsigned)nr), s = (int)(r - (unsigned)p * (unsigned)nr);
  int kind = 0, a = 0;  // 1: pair r, 2: pair r_inv, 3: single ray
  if (s > mpi && s < off) {
    kind = 3;
    a = s;
  } else {
    const bool first = s <= mpi;
    const int a2 = s - off;
    const bool second = a2 >= 0 && a2 <= mpi;
    if (second && (off <= 0 || !first)) {
      kind = 2;
      a = a2;
    } else if (first) {
      kind = 1;
      a = s;
    }
  }
  if (kind == 0) return;
  float angle = min_angle;
  for (int k = 0; k < a; ++k) angle += step;
  float gx, gy;
  rl::to_grid(m.world, ins[3 * (size_t)p], ins[3 * (size_t)p + 1], &gx, &gy);
  const float h = (-ins[3 * (size_t)p + 2] + m.world.rotation_const) - angle;
  float v;
  if (kind == 3) {
    int unused;
    v = cast_one<METHOD, false>(m, lut, rl::Box{nullptr, 0, 0}, trig, max_range, gy, gx, h, &unused);
  } else {
    float r0 = -1.0f, r1 = -1.0f;
    if (METHOD == RL_METHOD_CDDT) rl::cddt_calc_range_pair(m, max_range, gy, gx, h, &r0, &r1);
    v = kind == 1 ? r0 : r1;
  }
  outs[r] = v * m.world.scale;
}

namespace {

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes) {
    if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return fail(RL_ERR_NOMEM, "CDDT: device allocation of %zu bytes failed", bytes);
    }
    return RL_OK;
  }
  template <class T>
  T* as() const {
    return (T*)p;
  }
};

#include "rl_scan_sort.h"

void cddt_free(rl_map* m) {
  void* ps[] = {m->cddt_width, m->cddt_trans, m->cddt_cos, m->cddt_sin, m->cddt_bin_off, m->cddt_start, m->cddt_vals};
  for (void* p : ps)
    if (p) (void)hipFree(p);
  m->cddt_width = nullptr;
  m->cddt_trans = m->cddt_cos = m->cddt_sin = m->cddt_vals = nullptr;
  m->cddt_bin_off = m->cddt_start = nullptr;
  m->cddt_td = 0;
  m->cddt_nb = m->cddt_nv = 0;
}

// Exclusive scan of counts[0..n) into out[0..n] (out[n] = total), on the stream.
int scan_counts(hipStream_t st, const unsigned long long* counts, long long* out, long long n) {
  // the scan writes n + 1 values: counts is read with a trailing zero
  return exclusive_sum_u64(st, counts, (unsigned long long*)out, n + 1);
}

// Keeps the entries selected by `mode` (see k_cddt_compact): new start and values.
int cddt_compact(rl_map* m, int mode, const uint8_t* used) {
  hipStream_t st = m->ctx->stream;
  const long long nb = m->cddt_nb;
  DevBuf kept, nstart, nvals;
  TRY(kept.alloc((size_t)(nb + 1) * sizeof(unsigned long long)));
  TRY(nstart.alloc((size_t)(nb + 1) * sizeof(long long)));
  HIP_TRY(hipMemsetAsync(kept.p, 0, (size_t)(nb + 1) * sizeof(unsigned long long), st));
  const unsigned g = (unsigned)grid_for(nb, 256);
  hipLaunchKernelGGL(k_cddt_compact<false>, dim3(g), dim3(256), 0, st, nb, (const long long*)m->cddt_start,
                     (const float*)m->cddt_vals, used, mode, (const long long*)nullptr, kept.as<long long>(),
                     (float*)nullptr);
  HIP_TRY(hipGetLastError());
  TRY(scan_counts(st, kept.as<unsigned long long>(), nstart.as<long long>(), nb));
  long long total = 0;
  HIP_TRY(hipMemcpy(&total, nstart.as<long long>() + nb, sizeof(long long), hipMemcpyDeviceToHost));
  TRY(nvals.alloc((size_t)total * sizeof(float)));
  hipLaunchKernelGGL(k_cddt_compact<true>, dim3(g), dim3(256), 0, st, nb, (const long long*)m->cddt_start,
                     (const float*)m->cddt_vals, used, mode, (const long long*)nstart.p, (long long*)nullptr,
                     nvals.as<float>());
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  (void)hipFree(m->cddt_start);
  (void)hipFree(m->cddt_vals);
  m->cddt_start = nstart.as<long long>();
  m->cddt_vals = nvals.as<float>();
  nstart.p = nvals.p = nullptr;
  m->cddt_nv = total;
  return RL_OK;
}

int cddt_build_locked(rl_map* m, int td) {
  rl_ctx* c = m->ctx;
  hipStream_t st = c->stream;
  const int W = m->width, H = m->height;
  if ((long long)W * H > 0x7fffffffLL) return fail(RL_ERR_ARG, "CDDT: maps above 2^31 cells are not supported");
  cddt_free(m);
  m->cddt_td = td;
  int nang = 0;
  while (nang < td && nang < td / 2.0) ++nang;  // a < theta_discretization / 2.0
  m->cddt_nang = nang;
  m->cddt_td_div_2pi = (float)((double)td / 6.28318530718);          // RangeLib.h:976
  const float m2pi_div_td = (float)(6.28318530718 / (double)(float)td);  // :977
  HIP_TRY(hipMalloc(&m->cddt_width, td * sizeof(int)));
  HIP_TRY(hipMalloc(&m->cddt_trans, td * sizeof(float)));
  HIP_TRY(hipMalloc(&m->cddt_cos, td * sizeof(float)));
  HIP_TRY(hipMalloc(&m->cddt_sin, td * sizeof(float)));
  HIP_TRY(hipMalloc(&m->cddt_bin_off, (td + 1) * sizeof(long long)));
  hipLaunchKernelGGL(k_cddt_consts, dim3(grid_for(td, 64)), dim3(64), 0, st, c->trig, td, W, H, m2pi_div_td,
                     m->cddt_width, m->cddt_trans, m->cddt_cos, m->cddt_sin);
  HIP_TRY(hipGetLastError());
  std::vector<int> widths(td);
  HIP_TRY(hipMemcpyAsync(widths.data(), m->cddt_width, td * sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  std::vector<long long> off(td + 1, 0);
  for (int a = 0; a < td; ++a) off[a + 1] = off[a] + widths[a];
  const long long nb = off[td];
  m->cddt_nb = nb;
  HIP_TRY(hipMemcpy(m->cddt_bin_off, off.data(), (td + 1) * sizeof(long long), hipMemcpyHostToDevice));
  // edge cells
  const long long N = (long long)W * H;
  DevBuf edges, n_edges;
  TRY(edges.alloc((size_t)N * sizeof(int)));
  TRY(n_edges.alloc(sizeof(unsigned)));
  HIP_TRY(hipMemsetAsync(n_edges.p, 0, sizeof(unsigned), st));
  hipLaunchKernelGGL(k_cddt_edges, dim3(grid_for(N, 256)), dim3(256), 0, st, (const uint8_t*)m->d_occ, W, H,
                     edges.as<int>(), n_edges.as<unsigned>());
  HIP_TRY(hipGetLastError());
  unsigned ne = 0;
  HIP_TRY(hipMemcpyAsync(&ne, n_edges.p, sizeof(unsigned), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  // count, scan, fill
  DevBuf counter;
  TRY(counter.alloc((size_t)(nb + 1) * sizeof(unsigned long long)));
  HIP_TRY(hipMemsetAsync(counter.p, 0, (size_t)(nb + 1) * sizeof(unsigned long long), st));
  const long long pairs = (long long)ne * nang;
  if (pairs > 0) {
    for (int a0 = 0; a0 < nang; a0 += 65535) {  // grid rows: at most 65535 angles per launch
      hipLaunchKernelGGL(k_cddt_scatter<false>, dim3(grid_for(ne, 256), (unsigned)std::min(65535, nang - a0)),
                         dim3(256), 0, st, (const int*)edges.p, (int)ne, a0, H, (const int*)m->cddt_width,
                         (const float*)m->cddt_trans, (const float*)m->cddt_cos, (const float*)m->cddt_sin,
                         (const long long*)m->cddt_bin_off, counter.as<unsigned long long>(), (float*)nullptr);
      HIP_TRY(hipGetLastError());
    }
  }
  HIP_TRY(hipMalloc(&m->cddt_start, (size_t)(nb + 1) * sizeof(long long)));
  TRY(scan_counts(st, counter.as<unsigned long long>(), m->cddt_start, nb));
  long long total = 0;
  HIP_TRY(hipMemcpy(&total, m->cddt_start + nb, sizeof(long long), hipMemcpyDeviceToHost));
  if (total > 0x7fffffffLL) return fail(RL_ERR_ARG, "CDDT: %lld table entries exceed 2^31", total);
  HIP_TRY(hipMemcpy(counter.p, m->cddt_start, (size_t)(nb + 1) * sizeof(long long), hipMemcpyDeviceToDevice));
  DevBuf raw;
  TRY(raw.alloc((size_t)total * sizeof(float)));
  if (pairs > 0) {
    for (int a0 = 0; a0 < nang; a0 += 65535) {  // grid rows: at most 65535 angles per launch
      hipLaunchKernelGGL(k_cddt_scatter<true>, dim3(grid_for(ne, 256), (unsigned)std::min(65535, nang - a0)),
                         dim3(256), 0, st, (const int*)edges.p, (int)ne, a0, H, (const int*)m->cddt_width,
                         (const float*)m->cddt_trans, (const float*)m->cddt_cos, (const float*)m->cddt_sin,
                         (const long long*)m->cddt_bin_off, counter.as<unsigned long long>(), raw.as<float>());
      HIP_TRY(hipGetLastError());
    }
  }
  // sort each bin
  HIP_TRY(hipMalloc(&m->cddt_vals, (size_t)(total ? total : 1) * sizeof(float)));
  if (total > 0) TRY(segmented_sort_f32(st, raw.as<float>(), m->cddt_vals, m->cddt_start, nb));
  m->cddt_nv = total;
  // drop duplicates
  return cddt_compact(m, 0, nullptr);
}

int cddt_prune_locked(rl_map* m, float max_range) {
  hipStream_t st = m->ctx->stream;
  const int W = m->width, H = m->height;
  DevBuf used;
  TRY(used.alloc((size_t)m->cddt_nv + 1));
  HIP_TRY(hipMemsetAsync(used.p, 0, (size_t)m->cddt_nv + 1, st));
  const long long cells = (long long)W * H;
  for (int a0 = 0; cells > 0 && a0 < m->cddt_nang; a0 += 65535) {  // grid rows: at most 65535 angles per launch
    const int na = std::min(65535, m->cddt_nang - a0);
    hipLaunchKernelGGL(k_cddt_prune_mark, dim3(grid_for(cells, 256), (unsigned)na), dim3(256), 0, st,
                       (const uint8_t*)m->d_occ, W, H, max_range, (const int*)m->cddt_width, (const float*)m->cddt_trans,
                       (const float*)m->cddt_cos, (const float*)m->cddt_sin, (const long long*)m->cddt_bin_off,
                       (const long long*)m->cddt_start, (const float*)m->cddt_vals, used.as<uint8_t>(), a0);
    HIP_TRY(hipGetLastError());
  }
  return cddt_compact(m, 1, used.as<uint8_t>());
}

template <int METHOD>
int launch_radial_t(rl_map* m, float max_range, const float* ins, float* outs, int N, int nr, float min_angle,
                    float step, int mpi, int off) {
  const long long n = (long long)N * nr;
  if (n == 0) return RL_OK;
  hipLaunchKernelGGL(k_radial<METHOD>, dim3(grid_for(n, 256)), dim3(256), 0, m->ctx->stream, m->dev(),
                     m->ctx->trig, max_range, ins, outs, N, nr, min_angle, step, mpi, off);
  HIP_TRY(hipGetLastError());
  return RL_OK;
}

int launch_radial(rl_map* m, int method, float max_range, const float* ins, float* outs, int N, int nr,
                  float min_angle, float max_angle) {
  TRY(check_map(m, method, max_range));
  if ((long long)N * nr > 0x7fffffffLL) return fail(RL_ERR_ARG, "batch too large");
  // RangeLib.h:631-641
  const float step = (max_angle - min_angle) / (float)(nr - 1);
  const int mpi = (int)((double)(float)nr / 3.0);
  const float iof = (float)(((double)nr - 1.0) * 3.14159265358979323846 / (double)(max_angle - min_angle));
  const int off = (int)roundf(iof);
  switch (method) {
    case RL_METHOD_BL: return launch_radial_t<RL_METHOD_BL>(m, max_range, ins, outs, N, nr, min_angle, step, mpi, off);
    case RL_METHOD_GLT: return launch_radial_t<RL_METHOD_GLT>(m, max_range, ins, outs, N, nr, min_angle, step, mpi, off);
    case RL_METHOD_CDDT: return launch_radial_t<RL_METHOD_CDDT>(m, max_range, ins, outs, N, nr, min_angle, step, mpi, off);
    default: return launch_radial_t<RL_METHOD_RM>(m, max_range, ins, outs, N, nr, min_angle, step, mpi, off);
  }
}

}  // namespace

extern "C" {

int rl_calc_range_many_radial_optimized_dev(rl_map* m, int method, float max_range, const float* d_ins,
                                            float* d_outs, int num_particles, int num_rays, float min_angle,
                                            float max_angle) {
  CHECK_ARG(m && d_ins && d_outs && num_particles >= 0 && num_rays >= 2, "bad argument");
  std::lock_guard<std::mutex> lk(m->ctx->mu);
  TRY(use_device(m->ctx));
  return launch_radial(m, method, max_range, d_ins, d_outs, num_particles, num_rays, min_angle, max_angle);
}

int rl_calc_range_many_radial_optimized(rl_map* m, int method, float max_range, const float* ins, float* outs,
                                        int num_particles, int num_rays, float min_angle, float max_angle) {
  CHECK_ARG(m && ins && outs && num_particles >= 0 && num_rays >= 2, "bad argument");
  rl_ctx* c = m->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(use_device(c));
  const long long n = (long long)num_particles * num_rays;
  if (n > 0x7fffffffLL) return fail(RL_ERR_ARG, "batch too large");
  if (n == 0) return RL_OK;
  void *d_ins, *d_out;
  TRY(c->staging(0, (size_t)num_particles * 3 * sizeof(float), &d_ins));
  TRY(c->staging(1, (size_t)n * sizeof(float), &d_out));
  HIP_TRY(hipMemcpyAsync(d_ins, ins, (size_t)num_particles * 3 * sizeof(float), hipMemcpyHostToDevice, c->stream));
  // slots the reference leaves untouched keep the caller's values
  HIP_TRY(hipMemcpyAsync(d_out, outs, (size_t)n * sizeof(float), hipMemcpyHostToDevice, c->stream));
  TRY(launch_radial(m, method, max_range, (const float*)d_ins, (float*)d_out, num_particles, num_rays, min_angle,
                    max_angle));
  HIP_TRY(hipMemcpyAsync(outs, d_out, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RL_OK;
}

int rl_map_build_cddt(rl_map* m, int theta_disc) {
  CHECK_ARG(m, "map is NULL");
  CHECK_ARG(theta_disc >= 1 && theta_disc <= 65536, "theta_discretization outside 1..65536");
  rl_ctx* c = m->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(use_device(c));
  HIP_TRY(hipStreamSynchronize(c->stream));
  m->gen.fetch_add(1);
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  (void)hipEventRecord(e0, c->stream);
  int rc = cddt_build_locked(m, theta_disc);
  (void)hipEventRecord(e1, c->stream);
  (void)hipEventSynchronize(e1);
  (void)hipEventElapsedTime(&m->cddt_ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != RL_OK) {
    std::string msg = g_err;
    cddt_free(m);
    return fail(rc, "%s", msg.c_str());
  }
  return RL_OK;
}

int rl_map_prune_cddt(rl_map* m, float max_range) {
  CHECK_ARG(m, "map is NULL");
  rl_ctx* c = m->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(use_device(c));
  if (!m->cddt_vals) return fail(RL_ERR_STATE, "no CDDT tables (call rl_map_build_cddt first)");
  HIP_TRY(hipStreamSynchronize(c->stream));
  m->gen.fetch_add(1);
  return cddt_prune_locked(m, max_range);
}

int rl_map_get_cddt(rl_map* m, int* theta_disc, long long* nbins, long long* nvals, int* widths, float* trans,
                    long long* start, float* vals, float* build_ms) {
  CHECK_ARG(m, "map is NULL");
  rl_ctx* c = m->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(use_device(c));
  if (!m->cddt_vals) return fail(RL_ERR_STATE, "no CDDT tables (call rl_map_build_cddt first)");
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (theta_disc) *theta_disc = m->cddt_td;
  if (nbins) *nbins = m->cddt_nb;
  if (nvals) *nvals = m->cddt_nv;
  if (build_ms) *build_ms = m->cddt_ms;
  if (widths) HIP_TRY(hipMemcpy(widths, m->cddt_width, m->cddt_td * sizeof(int), hipMemcpyDeviceToHost));
  if (trans) HIP_TRY(hipMemcpy(trans, m->cddt_trans, m->cddt_td * sizeof(float), hipMemcpyDeviceToHost));
  if (start)
    HIP_TRY(hipMemcpy(start, m->cddt_start, (size_t)(m->cddt_nb + 1) * sizeof(long long), hipMemcpyDeviceToHost));
  if (vals && m->cddt_nv)
    HIP_TRY(hipMemcpy(vals, m->cddt_vals, (size_t)m->cddt_nv * sizeof(float), hipMemcpyDeviceToHost));
  return RL_OK;
}

int rl_debug_exclusive_sum_dev(rl_ctx* c, const unsigned long long* in, unsigned long long* out, long long n) {
  CHECK_ARG(c && (n <= 0 || (in && out)), "bad argument");
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(use_device(c));
  return exclusive_sum_u64(c->stream, in, out, n);
}

int rl_debug_segmented_sort_dev(rl_ctx* c, const float* in, float* out, const long long* start, long long nseg) {
  CHECK_ARG(c && (nseg <= 0 || (in && out && start)), "bad argument");
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(use_device(c));
  return segmented_sort_f32(c->stream, in, out, start, nseg);
}

}  // extern "C"
