// librangelib_hip: ray casting on MI355X (gfx950) behind the C ABI of
// include/rangelib_hip.h.
//
// Kernels
//   k_edt_cols      distance-transform pass 1 (closed form along y, per column)
//   k_edt_init      f = 0 (occupied) / FLT_MAX (free) for the envelope pass
//   k_edt_envelope  distance-transform lower envelope, one line per lane,
//                   replicating vendor/distance_transform.h:1054-1096
//   k_edt_rows      pass 2 as an exact bounded search in LDS on the rows where
//                   it gives the envelope's bits; the others go to k_edt_envelope
//   k_cast          one ray per lane: ray marching or Bresenham, grid / world /
//                   repeated-angles frames
//   k_sensor        eval_sensor_model: one particle per wave segment
//   k_fused         calc_range_repeat_angles_eval_sensor_model
//
// Build: hipcc --offload-arch=gfx950 -O3 -ffp-contract=off (see build.py).
#include <dlfcn.h>
#include <gnu/libc-version.h>
#include <hip/hip_runtime.h>
#include <libgen.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/rangelib_hip.h"
#include "rl_device.h"
#include "rl_png.h"

#pragma clang fp contract(off)

using rl::MapDev;
using rl::TrigDev;
using rl::WorldDev;

// Staging area of rl_calc_range_repeat_angles_eval_sensor_model_noalloc: one
// copy of this array per device, held until the process ends.  Host code must
// not pass the array's name to a kernel or a copy (on the host it is a stand-in,
// not a device address): rl_create looks up its address on its device with
// hipGetSymbolAddress and every use goes through that pointer (rl_ctx::d_scratch).
static __device__ __align__(256) unsigned char scratch[RL_NOALLOC_SCRATCH_BYTES];

// ---------------------------------------------------------------------------
// Kernels

// Pass 1 along y for column x = blockIdx.x.  For a 0 / FLT_MAX input the
// reference's lower envelope yields (y - nearest occupied y)^2, or FLT_MAX for
// a column without obstacles; every intermediate of the envelope is an exact
// integer while y^2 < 2^24, i.e. for heights up to 4096, where this closed form
// is used.
__global__ __launch_bounds__(256) void k_edt_cols(const uint8_t* __restrict__ occ,
                                                  float* __restrict__ g, int H) {
  __shared__ int s_last[256];
  __shared__ int s_first[256];
  const int x = blockIdx.x;
  const int t = threadIdx.x;
  const uint8_t* col = occ + (size_t)x * H;
  float* gc = g + (size_t)x * H;
  const int chunk = (H + 255) / 256;
  const int y0 = min(H, t * chunk), y1 = min(H, y0 + chunk);
  int last = -1, first = INT_MAX;
  for (int y = y0; y < y1; ++y)
    if (col[y]) {
      if (first == INT_MAX) first = y;
      last = y;
    }
  s_last[t] = last;
  s_first[t] = first;
  __syncthreads();
  if (t == 0) {  // exclusive prefix max / suffix min over the 256 chunks
    int run = -1;
    for (int i = 0; i < 256; ++i) {
      int v = s_last[i];
      s_last[i] = run;
      run = v > run ? v : run;
    }
    run = INT_MAX;
    for (int i = 255; i >= 0; --i) {
      int v = s_first[i];
      s_first[i] = run;
      run = v < run ? v : run;
    }
  }
  __syncthreads();
  int prev = s_last[t];
  for (int y = y0; y < y1; ++y) {
    if (col[y]) prev = y;
    gc[y] = __int_as_float(prev);  // parked here until the backward sweep
  }
  int next = s_first[t];
  for (int y = y1 - 1; y >= y0; --y) {
    if (col[y]) next = y;
    int p = __float_as_int(gc[y]);
    int dmin = INT_MAX;
    if (p >= 0) dmin = y - p;
    if (next != INT_MAX && next - y < dmin) dmin = next - y;
    gc[y] = (dmin == INT_MAX) ? FLT_MAX : (float)((unsigned)dmin * (unsigned)dmin);
  }
}

__global__ void k_edt_init(const uint8_t* __restrict__ occ, float* __restrict__ f, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = occ[i] ? 0.0f : FLT_MAX;
}

// Lower envelope of parabolas along one line per lane
// (vendor/distance_transform.h:1054-1096, types as there: float values, double
// intersections).  Line `l` starts at f + l*ls and steps by `es`; the stacks
// v[] and z[] of line l live at v[k*n_lines + l] so that neighbouring lanes
// touch neighbouring words.
__global__ __launch_bounds__(64) void k_edt_envelope(const float* __restrict__ f,
                                                     float* __restrict__ D, int n_lines, int n,
                                                     long long es, long long ls,
                                                     int* __restrict__ vbuf,
                                                     double* __restrict__ zbuf, int do_sqrt,
                                                     const int* __restrict__ only) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= n_lines) return;
  if (only && !only[l]) return;
  const float* fl = f + (long long)l * ls;
  float* Dl = D + (long long)l * ls;
  int* v = vbuf + l;
  double* z = zbuf + l;
  const long long L = n_lines;
  if (n == 1) {
    float val = fl[0];
    Dl[0] = do_sqrt ? __builtin_sqrtf(val) : val;
    return;
  }
  unsigned k = 0;
  double s = 0.0;
  v[0] = 0;
  z[0] = -DBL_MAX;
  z[L] = DBL_MAX;
  for (unsigned q = 1; q < (unsigned)n; ++q) {
    float fq = fl[(long long)q * es] + (float)(q * q);
    ++k;
    do {
      --k;
      unsigned vk = (unsigned)v[(long long)k * L];
      float rhs = fl[(long long)vk * es] + (float)(vk * vk);
      s = ((double)fq - (double)rhs) / ((double)(2u * q) - (double)(2u * vk));
    } while (s <= z[(long long)k * L]);
    ++k;
    v[(long long)k * L] = (int)q;
    z[(long long)k * L] = s;
    z[(long long)(k + 1) * L] = DBL_MAX;
  }
  k = 0;
  for (unsigned q = 0; q < (unsigned)n; ++q) {
    while (z[(long long)(k + 1) * L] < (double)q) ++k;
    unsigned vk = (unsigned)v[(long long)k * L];
    float dq = (float)q - (float)vk;
    float val = fl[(long long)vk * es] + dq * dq;
    Dl[(long long)q * es] = do_sqrt ? __builtin_sqrtf(val) : val;
  }
}

// Pass 2 along x for EDT_RB consecutive rows per workgroup, as a bounded
// search in LDS instead of the sequential lower envelope, on rows where the two
// provably give the same bits.
//
// Why they agree: let every finite input of a row be a non-negative integer
// with g[x'] + x'^2 < 2^24, and W <= 4096.  Then the envelope's float sums
// f[q] + q^2 are exact, every intersection it forms is a correctly rounded
// quotient of two exact integers with a denominator of at most 2W, and two
// distinct such quotients differ by far more than their rounding, so each
// comparison it makes has the exact outcome: the parabola it assigns to q is
// an exact argmin of g[x'] + (q - x')^2.  Its output float(g[v] + (q - v)^2)
// is therefore the minimum over x' of the same float expression (rounding is
// monotone and (q - x')^2 is exact below 4096).  FLT_MAX inputs (columns
// without obstacles) only ever intersect the finite ones at magnitudes beyond
// 1e34, so they are never assigned while a finite input exists, and give
// FLT_MAX when none does -- as the minimum does.
//
// The search from x stops at the first distance d with d^2 >= best: every
// candidate beyond it is at least d^2.  Rows whose inputs break the condition
// are marked in slow[y] (every row's mark is written) and left to
// k_edt_envelope.  LDS holds the rows x-major, s[x * EDT_RB + r], the order
// they sit in global memory, so loads, searches and stores are unit-stride
// across the lanes.
constexpr int EDT_RB = 4;
__global__ __launch_bounds__(256) void k_edt_rows(const float* __restrict__ g,
                                                  float* __restrict__ D, int W, int H,
                                                  int* __restrict__ slow, int do_sqrt) {
  extern __shared__ float s_g[];
  __shared__ int s_bad[EDT_RB], s_fin[EDT_RB];
  const int y0 = blockIdx.x * EDT_RB;
  const int t = threadIdx.x;
  const int n = EDT_RB * W;
  if (t < EDT_RB) {
    s_bad[t] = 0;
    s_fin[t] = 0;
  }
  __syncthreads();
  for (int c = t; c < n; c += 256) {
    const int r = c % EDT_RB, x = c / EDT_RB, y = y0 + r;
    float v = FLT_MAX;
    if (y < H) {
      v = rl::glb(g)[(size_t)x * H + y];
      if (v != FLT_MAX) {
        const bool exact = v >= 0.0f && v == floorf(v) &&
                           (double)v + (double)x * (double)x < 16777216.0;
        if (exact)
          s_fin[r] = 1;
        else
          s_bad[r] = 1;
      }
    }
    s_g[c] = v;
  }
  __syncthreads();
  if (t < EDT_RB && y0 + t < H) slow[y0 + t] = s_bad[t];
  for (int c = t; c < n; c += 256) {
    const int r = c % EDT_RB, x = c / EDT_RB, y = y0 + r;
    if (y >= H || s_bad[r]) continue;
    float best = FLT_MAX;
    if (s_fin[r]) {
      best = s_g[c];
      const int dmax = max(x, W - 1 - x);
      for (int d = 1; d <= dmax; ++d) {
        const float dd = (float)(d * d);
        if (dd >= best) break;
        if (x >= d) {
          const float v = s_g[c - d * EDT_RB] + dd;
          best = v < best ? v : best;
        }
        if (x + d < W) {
          const float v = s_g[c + d * EDT_RB] + dd;
          best = v < best ? v : best;
        }
      }
    }
    D[(size_t)x * H + y] = do_sqrt ? __builtin_sqrtf(best) : best;
  }
}

// Pose of ray r in the frame the method consumes.  FRAME 0 = grid (no
// conversion), 1 = world (RangeLib.h:463-475), 2 = repeated angles
// (RangeLib.h:506-518).  Returns the factor the range is multiplied by.
template <int FRAME>
__device__ __forceinline__ float ray_pose(const MapDev& m, const float* __restrict__ ins,
                                          const float* __restrict__ angles, int r, int num_angles,
                                          float* x, float* y, float* heading) {
  if (FRAME == 0) {
    *x = ins[3 * (size_t)r];
    *y = ins[3 * (size_t)r + 1];
    *heading = ins[3 * (size_t)r + 2];
    return 1.0f;
  }
  int p = r, a = 0;
  if (FRAME == 2) {
    p = r / num_angles;
    a = r - p * num_angles;
  }
  float gx, gy;
  rl::to_grid(m.world, ins[3 * (size_t)p], ins[3 * (size_t)p + 1], &gx, &gy);
  float theta = -ins[3 * (size_t)p + 2] + m.world.rotation_const;
  if (FRAME == 2) theta = theta - angles[a];
  *x = gy;  // calc_range(y, x, theta): the grid axes swap (RangeLib.h:475)
  *y = gx;
  *heading = theta;
  return m.world.scale;
}

template <int METHOD, bool COUNT>
__device__ __forceinline__ float cast_one(const MapDev& m, const float* lut, const rl::Box& box,
                                          const TrigDev& trig, float max_range, float x, float y,
                                          float heading, int* cnt) {
  if (METHOD == RL_METHOD_GLT) return rl::glt_calc_range<COUNT>(m, x, y, heading, cnt);
  if (METHOD == RL_METHOD_CDDT) return rl::cddt_calc_range<COUNT>(m, max_range, x, y, heading, cnt);
  if (METHOD == RL_METHOD_RM && m.idx8 && rl::fast_march_ok(x, y, max_range))
    return rl::rm_cast_idx8<COUNT>(m, lut, box, trig, max_range, x, y, heading, cnt);
  float s, c;
  rl::sincos_libc(heading, trig, &s, &c);
  if (METHOD == RL_METHOD_BL) return rl::bl_trace<COUNT>(m, max_range, x, y, s, c, cnt);
  return rl::rm_march<COUNT>(m.edt, m.width, m.height, max_range, x, y, c, s, cnt);
}

// Copies the distance table into LDS (ray marching over the compact map only).
// Must be reached by every thread of the block.
template <int METHOD>
__device__ __forceinline__ void stage_lut(const MapDev& m, float* sh) {
  if (METHOD == RL_METHOD_RM && m.idx8) {
    for (int i = threadIdx.x; i < RL_LUT_SIZE; i += blockDim.x) sh[i] = m.steps[i];
    __syncthreads();
  }
}

template <int METHOD, int FRAME, bool COUNT>
__global__ __launch_bounds__(256) void k_cast(MapDev m, TrigDev trig, float max_range,
                                              const float* __restrict__ ins,
                                              const float* __restrict__ angles,
                                              float* __restrict__ outs, int* __restrict__ counts,
                                              int n_rays, int num_angles) {
  __shared__ float lut[RL_LUT_SIZE];
  stage_lut<METHOD>(m, lut);
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = r < n_rays;  // the whole wave stays to the end: rl::solo_tail needs every lane
  float x = 0.0f, y = 0.0f, h = 0.0f, scale = 1.0f;
  if (act) scale = ray_pose<FRAME>(m, ins, angles, r, num_angles, &x, &y, &h);
  // ray marching over the compact map: a wave's last marching ray is finished
  // by the whole wave (uniform across the wave)
  const bool solo = RL_SOLO_TAIL && !COUNT && METHOD == RL_METHOD_RM && m.idx8;
  rl::MarchTail tail{false, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0u};
  int cnt = 0;
  float d = 0.0f;
  if (METHOD == RL_METHOD_BL && m.bits) {  // Bresenham over the occupancy bitmap, a window slot per thread
    __shared__ uint4 win[2 * 256];
    if (act) {
      float s, c;
      rl::sincos_libc(h, trig, &s, &c);
      d = rl::bl_trace_bits<COUNT>(m, win + 2 * threadIdx.x, max_range, x, y, s, c, &cnt);
    }
  } else if (act) {
    d = (solo && rl::fast_march_ok(x, y, max_range))
            ? rl::rm_cast_idx8_t<false, false, true>(m, lut, rl::Box{nullptr, 0, 0}, trig, max_range, x, y, h, &cnt,
                                                     &tail)
            : cast_one<METHOD, COUNT>(m, lut, rl::Box{nullptr, 0, 0}, trig, max_range, x, y, h, &cnt);
  }
  if (solo) d = rl::solo_tail(m, lut, max_range, tail, d);
  if (!act) return;
  if (COUNT)
    counts[r] = cnt;
  else
    outs[r] = (FRAME == 0) ? d : d * scale;
}

// Regrouping, measured on 1 M Bresenham queries of a 2048^2 corridor map
// (R = 1000): 256 threads with 1, 2, 4, 6 or 8 rounds between regroupings
// 149, 134, 127, 129, 129 us; 128 or 512 threads (4 and 2 rounds) 133, 140 us;
// one walk per lane to its end 133 us.  Each regrouping costs three barriers
// and the parked state's LDS traffic, and waits for the workgroup's slowest wave.
#ifndef RL_BL_WG
#define RL_BL_WG 256  // threads of a k_cast_bl workgroup, whose walks are regrouped together
#endif
#ifndef RL_BL_PHASE
#define RL_BL_PHASE 4  // rounds of every walk between regroupings (k_cast_bl)
#endif
#define RL_BL_PARK 18  // words of a parked walk: rl::BlWalk and its ray

// A walk and its ray parked in slot i of LDS (one array per word, so that
// consecutive slots are in consecutive banks).
__device__ __forceinline__ void park_walk(unsigned (*park)[RL_BL_WG], int i, const rl::BlWalk& w, int ray) {
  park[0][i] = __float_as_uint(w.x), park[1][i] = __float_as_uint(w.y), park[2][i] = __float_as_uint(w.err);
  park[3][i] = __float_as_uint(w.deltax), park[4][i] = __float_as_uint(w.deltay);
  park[5][i] = __float_as_uint(w.xstep), park[6][i] = __float_as_uint(w.ystep);
  park[7][i] = __float_as_uint(w.limx), park[8][i] = w.limy;
  park[9][i] = __float_as_uint(w.x0), park[10][i] = __float_as_uint(w.y0);
  park[11][i] = w.sx, park[12][i] = w.sy, park[13][i] = (unsigned)w.xend, park[14][i] = (unsigned)w.left;
  park[15][i] = w.bbase, park[16][i] = w.brow, park[17][i] = (unsigned)ray;
}

__device__ __forceinline__ void unpark_walk(const unsigned (*park)[RL_BL_WG], int i, rl::BlWalk* w, int* ray) {
  w->x = __uint_as_float(park[0][i]), w->y = __uint_as_float(park[1][i]), w->err = __uint_as_float(park[2][i]);
  w->deltax = __uint_as_float(park[3][i]), w->deltay = __uint_as_float(park[4][i]);
  w->xstep = __uint_as_float(park[5][i]), w->ystep = __uint_as_float(park[6][i]);
  w->limx = __uint_as_float(park[7][i]), w->limy = park[8][i];
  w->x0 = __uint_as_float(park[9][i]), w->y0 = __uint_as_float(park[10][i]);
  w->sx = park[11][i], w->sy = park[12][i], w->xend = (int)park[13][i], w->left = (int)park[14][i];
  w->bbase = park[15][i], w->brow = park[16][i], *ray = (int)park[17][i];
}

// Bresenham rays (BresenhamsLine::calc_range, RangeLib.h:696-769, one per
// thread as numpy_calc_range / calc_range_many cast them) over the occupancy
// bitmap, with the walks of a workgroup regrouped every RL_BL_PHASE rounds:
// the walks still running move (through LDS) to the workgroup's first
// threads, so they fill whole waves and the other waves stop issuing.  Walk
// lengths vary widely (on corridor maps the longest is ~20x the mean), and
// with one walk per lane to the end most lanes sat idle while their wave's
// longest walk ran.  A walk's state moves unchanged, so each ray's result is
// that of one lane's walk.
template <int FRAME>
__global__ __launch_bounds__(RL_BL_WG) void k_cast_bl(MapDev m, TrigDev trig, float max_range,
                                                      const float* __restrict__ ins,
                                                      const float* __restrict__ angles,
                                                      float* __restrict__ outs, int n_rays, int num_angles) {
  __shared__ unsigned park[RL_BL_PARK][RL_BL_WG];
  __shared__ uint4 win[2 * RL_BL_WG];
  __shared__ int n_live;
  const int tid = threadIdx.x;
  const int r = blockIdx.x * RL_BL_WG + tid;
  const __amdgpu_buffer_rsrc_t bits =
      __builtin_amdgcn_make_buffer_rsrc((void*)m.bits, (short)0, (int)(m.bits_bytes + 16u), 0x00020000);
  const __amdgpu_buffer_rsrc_t occ = rl::occ_rsrc(m.occ, m.width, m.height);
  const float scale = (FRAME == 0) ? 1.0f : m.world.scale;
  rl::BlWalk w{};
  int ray = r;
  bool live = false;
  if (r < n_rays) {
    float x, y, h, s, c, d;
    ray_pose<FRAME>(m, ins, angles, r, num_angles, &x, &y, &h);
    rl::sincos_libc(h, trig, &s, &c);
    live = rl::bl_begin(m, max_range, x, y, s, c, &w, &d);
    if (!live) outs[r] = (FRAME == 0) ? d : d * scale;  // ended before its loop
  }
  for (;;) {  // ends when no walk of the workgroup is left (the same in every thread)
#pragma unroll 1
    for (int j = 0; j < RL_BL_PHASE; ++j) {
      if (live) {
        bool hit;
        int unused = 0;
        if (rl::bl_round_bits<false>(bits, occ, w, win + 2 * tid, &hit, &unused)) {
          const float d = rl::bl_range(w, hit, max_range);
          outs[ray] = (FRAME == 0) ? d : d * scale;
          live = false;
        }
      }
    }
    // regroup: each wave's running walks take consecutive slots
    if (tid == 0) n_live = 0;
    __syncthreads();
    const unsigned long long lm = __builtin_amdgcn_ballot_w64(live);
    int base = 0;
    if ((tid & 63) == 0 && lm) base = atomicAdd(&n_live, (int)__popcll(lm));
    base = __builtin_amdgcn_readfirstlane(base);
    if (live)
      park_walk(park, base + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(lm >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((unsigned)lm, 0u)),
                w, ray);
    __syncthreads();
    const int n = n_live;
    if (n == 0) break;
    live = tid < n;
    if (live) unpark_walk(park, tid, &w, &ray);
    __syncthreads();  // every slot read before the next regrouping writes
  }
}

// Particles are laid out in wave segments of S lanes (S a power of two,
// <= 64); lane j of a segment handles beams j, j + S, ...  The weight is the
// product over beams in beam order, as the reference's loop forms it
// (RangeLib.h:542-554), so it is bit-identical: lane l stores its term in
// slot[l] (a 64-double LDS slot per wave), then every lane reads its
// segment's terms slot[base], slot[base + 1], ... back in beam order (one
// address per segment and read) and multiplies them into w in that order.
// The terms stay in vector registers (reading them out lane by lane into
// scalar registers took 0.8 us more per particle-filter step, and cross-lane
// shuffles put two permutes on the chain per term).
__device__ __forceinline__ double slot_product(double w, const RL_LDS double* t, int cnt) {
  int jj = 0;
  for (; jj + 8 <= cnt; jj += 8) {
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = t[jj + q];
#pragma unroll
    for (int q = 0; q < 8; ++q) w *= v[q];
  }
  for (; jj < cnt; ++jj) w *= t[jj];
  return w;
}

__device__ __forceinline__ double segment_product(double w, double term, int base, int cnt, int S, double* slot) {
  RL_LDS double* t = (RL_LDS double*)slot;
  // the previous chunk's reads of the slot come before this chunk's writes
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  t[threadIdx.x & 63] = term;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // one particle per wave: the same addresses in every lane (a per-lane
  // segment offset there cost 0.2 us per particle-filter step)
  if (S == 64) return slot_product(w, t, cnt);
  return slot_product(w, t + base, cnt);
}

__global__ __launch_bounds__(256) void k_sensor(const float* __restrict__ obs,
                                                const float* __restrict__ ranges,
                                                double* __restrict__ outs,
                                                const double* __restrict__ table, int K,
                                                float inv_scale, int M, int N, int S) {
  const int lane = threadIdx.x & 63;
  const int wave = (int)(((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int per_wave = 64 / S;
  const int seg = lane / S, j0 = lane - seg * S, base = seg * S;
  const int p = wave * per_wave + seg;
  const bool valid = p < N;
  const float hi = (float)((double)(float)K - 1.0);
  __shared__ double terms[4 * 64];
  double w = 1.0;
  for (int c0 = 0; c0 < M; c0 += S) {
    int j = c0 + j0;
    double term = 1.0;
    if (valid && j < M) {
      float r = rl::clamp_table(obs[j] * inv_scale, hi);
      float d = rl::clamp_table(ranges[(size_t)p * M + j] * inv_scale, hi);
      term = rl::glb(table)[(size_t)(int)r * (size_t)K + (size_t)(int)d];
    }
    w = segment_product(w, term, base, min(S, M - c0), S, terms + (threadIdx.x >> 6) * 64);
  }
  if (valid && j0 == 0) outs[p] = w;
}

template <int METHOD>
__global__ __launch_bounds__(256) void k_fused(MapDev m, TrigDev trig, float max_range,
                                               const float* __restrict__ ins,
                                               const float* __restrict__ angles,
                                               const float* __restrict__ obs,
                                               const double* __restrict__ table, int K,
                                               double* __restrict__ weights, int N, int M, int S) {
  const int lane = threadIdx.x & 63;
  const int wave = (int)(((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int per_wave = 64 / S;
  const int seg = lane / S, j0 = lane - seg * S, base = seg * S;
  const int p = wave * per_wave + seg;
  const bool valid = p < N;
  const float hi = (float)((double)(float)K - 1.0);
  __shared__ float lut[RL_LUT_SIZE];
  __shared__ double terms[4 * 64];
  stage_lut<METHOD>(m, lut);
  float gx = 0.0f, gy = 0.0f, theta = 0.0f;
  if (valid) {  // RangeLib.h:589-597
    rl::to_grid(m.world, ins[3 * (size_t)p], ins[3 * (size_t)p + 1], &gx, &gy);
    theta = -ins[3 * (size_t)p + 2] + m.world.rotation_const;
  }
  double w = 1.0;
  for (int c0 = 0; c0 < M; c0 += S) {
    int j = c0 + j0;
    double term = 1.0;
    if (valid && j < M) {
      int unused;
      // the range stays in grid units here: no world_scale (RangeLib.h:602-603)
      float d = cast_one<METHOD, false>(m, lut, rl::Box{nullptr, 0, 0}, trig, max_range, gy, gx,
                                        theta - angles[j], &unused);
      d = rl::clamp_table(d, hi);
      float r = rl::clamp_table(obs[j] * m.world.inv_scale, hi);
      term = rl::glb(table)[(size_t)(int)r * (size_t)K + (size_t)(int)d];
    }
    w = segment_product(w, term, base, min(S, M - c0), S, terms + (threadIdx.x >> 6) * 64);
  }
  if (valid && j0 == 0) weights[p] = w;
}

// XCD-aware scheduling of particles.  The map is cut into RL_REGIONS
// rectangles (4 x 2 or 2 x 4, the longer side cut in four); k_region buckets
// the particles by the region their origin cell lies in.  The particle kernel
// then runs RL_REGIONS * ceil(N / per_block) workgroups: workgroup b takes the
// (b / RL_REGIONS)-th chunk of region b % RL_REGIONS and exits at once when
// that chunk is empty.  Workgroups are dispatched to the XCDs round-robin
// (b -> XCD b % 8 on MI355X), so each XCD's L2 serves about one eighth of the
// map; correctness does not depend on that placement.  No atomics.
#define RL_REGIONS 8
#define RL_XCD_SCHED_MAX 65536  // larger batches fill the GPU many times over: plain order

struct Sched {
  int* perm;    // particle indices grouped by region
  int* count;   // [RL_REGIONS]
  int* offset;  // [RL_REGIONS]
};

__device__ __forceinline__ int region_of(const MapDev& m, float x0, float y0) {
  // x0, y0: grid coordinates of the ray origin (first index, second index)
  int cx = (x0 >= 0.0f && x0 < (float)m.width) ? (int)x0 : 0;
  int cy = (y0 >= 0.0f && y0 < (float)m.height) ? (int)y0 : 0;
  int rx = m.width >= m.height ? 4 : 2;
  int ry = RL_REGIONS / rx;
  int ix = (int)(((long long)cx * rx) / m.width);
  int iy = (int)(((long long)cy * ry) / m.height);
  return ix * ry + iy;
}

__device__ __forceinline__ int particle_region(const MapDev& m, const float* __restrict__ ins, int p) {
  float gx, gy;
  rl::to_grid(m.world, ins[3 * (size_t)p], ins[3 * (size_t)p + 1], &gx, &gy);
  return region_of(m, gy, gx);  // calc_range(y, x, .) marches from (gy, gx)
}

// One workgroup of 1024 threads buckets N particles.
__global__ __launch_bounds__(1024) void k_region(MapDev m, const float* __restrict__ ins, int N, Sched s) {
  __shared__ int cnt[RL_REGIONS], fill[RL_REGIONS];
  if (threadIdx.x < RL_REGIONS) cnt[threadIdx.x] = 0;
  __syncthreads();
  for (int p = threadIdx.x; p < N; p += blockDim.x) atomicAdd(&cnt[particle_region(m, ins, p)], 1);
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int r = 0; r < RL_REGIONS; ++r) {
      s.count[r] = cnt[r];
      s.offset[r] = run;
      fill[r] = run;
      run += cnt[r];
    }
  }
  __syncthreads();
  for (int p = threadIdx.x; p < N; p += blockDim.x) {
    int pos = atomicAdd(&fill[particle_region(m, ins, p)], 1);
    s.perm[pos] = p;
  }
}

// A particle's pose and its first beam's angle and observation, loaded before
// the kernel stages its tables so that the loads overlap (ok = false: not
// loaded, particle_beams loads them itself).
struct BeamIn {
  float x, y, th, a0, o0;
  bool ok;
};

template <bool SENSOR>
__device__ __forceinline__ BeamIn load_beam_in(const float* __restrict__ ins, const float* __restrict__ angles,
                                               const float* __restrict__ obs, int p, int M, int j0) {
  BeamIn b{0.0f, 0.0f, 0.0f, 0.0f, 0.0f, true};
  if (p >= 0) {
    b.x = ins[3 * (size_t)p];
    b.y = ins[3 * (size_t)p + 1];
    b.th = ins[3 * (size_t)p + 2];
    if (j0 < M) {
      b.a0 = angles[j0];
      if (SENSOR) b.o0 = obs[j0];
    }
  }
  return b;
}

// All beams of particle p (p < 0: none) on the lanes of one wave segment:
// numpy_calc_range_angles (RangeLib.h:482-520) and, with SENSOR,
// eval_sensor_model (RangeLib.h:533-555) on the ranges just written -- each
// range is read back as the separate call would (range * inv_world_scale), so
// both outputs are bit-identical to the two calls.  Must be reached by the
// whole wave (the product broadcasts across lanes).  SOLO: a wave's last
// marching ray is finished by the whole wave (rl::solo_tail), which shortens
// the launch's longest chain; launches bound by their total work (the multi-map
// batch) do without it.
template <int METHOD, bool SENSOR, bool SOLO = true>
__device__ __forceinline__ void particle_beams(const MapDev& m, const float* lut, const TrigDev& trig,
                                               float max_range, const float* __restrict__ ins,
                                               const float* __restrict__ angles,
                                               const float* __restrict__ obs,
                                               const double* __restrict__ table, int K,
                                               float* __restrict__ ranges,
                                               double* __restrict__ weights, int p, int M, int S,
                                               int j0, int base, uint8_t* box_lds, double* terms_lds,
                                               BeamIn pre) {
  const bool valid = p >= 0;
  const float hi = (float)((double)(float)K - 1.0);
  float gx = 0.0f, gy = 0.0f, theta = 0.0f;
  if (valid) {
    const float wx = pre.ok ? pre.x : ins[3 * (size_t)p];
    const float wy = pre.ok ? pre.y : ins[3 * (size_t)p + 1];
    const float wth = pre.ok ? pre.th : ins[3 * (size_t)p + 2];
    rl::to_grid(m.world, wx, wy, &gx, &gy);
    theta = -wth + m.world.rotation_const;
  }
  // one particle per wave: its beams share the LDS box around the origin
  rl::Box box{nullptr, 0, 0};
  if (METHOD == RL_METHOD_RM && S == 64 && m.idx8 && m.use_box && box_lds)
    box = rl::stage_box(m, box_lds, gy, gx, valid && rl::fast_march_ok(gy, gx, max_range));
  // ray marching over the compact map: a wave's last marching ray is finished
  // by the whole wave (rl::solo_tail); the same for every lane of the wave
  const bool solo = SOLO && RL_SOLO_TAIL && METHOD == RL_METHOD_RM && m.idx8 && !box.lds;
  const bool fast = rl::fast_march_ok(gy, gx, max_range);
  double w = 1.0;
  for (int c0 = 0; c0 < M; c0 += S) {
    int j = c0 + j0;
    double term = 1.0;
    const bool act = valid && j < M;
    const bool first = c0 == 0 && pre.ok;
    rl::MarchTail tail{false, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0u};
    float d = 0.0f;
    if (act) {
      int unused;
      const float aj = first ? pre.a0 : angles[j];
      d = (solo && fast) ? rl::rm_cast_idx8_t<false, false, true>(m, lut, box, trig, max_range, gy, gx, theta - aj,
                                                                  &unused, &tail)
                         : cast_one<METHOD, false>(m, lut, box, trig, max_range, gy, gx, theta - aj, &unused);
    }
    if (solo) d = rl::solo_tail(m, lut, max_range, tail, d);
    if (act) {
      d = d * m.world.scale;
      ranges[(size_t)p * M + j] = d;
      if (SENSOR) {
        float r = rl::clamp_table((first ? pre.o0 : obs[j]) * m.world.inv_scale, hi);
        float dd = rl::clamp_table(d * m.world.inv_scale, hi);
        term = rl::glb(table)[(size_t)(int)r * (size_t)K + (size_t)(int)dd];
      }
    }
    if (SENSOR)
      w = segment_product(w, term, base, min(S, M - c0), S, terms_lds);
  }
  if (SENSOR && valid && j0 == 0) weights[p] = w;
}

// Particles in input order: wave segment -> particle.
template <int METHOD, bool SENSOR>
__global__ __launch_bounds__(256) void k_particles(MapDev m, TrigDev trig, float max_range,
                                                   const float* __restrict__ ins,
                                                   const float* __restrict__ angles,
                                                   const float* __restrict__ obs,
                                                   const double* __restrict__ table, int K,
                                                   float* __restrict__ ranges,
                                                   double* __restrict__ weights, int N, int M, int S) {
  const int lane = threadIdx.x & 63;
  const int wave = (int)(((size_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int per_wave = 64 / S;
  const int seg = lane / S, j0 = lane - seg * S, base = seg * S;
  const int p = wave * per_wave + seg < N ? wave * per_wave + seg : -1;
  const BeamIn pre = load_beam_in<SENSOR>(ins, angles, obs, p, M, j0);
  __shared__ float lut[RL_LUT_SIZE];
  stage_lut<METHOD>(m, lut);
  __shared__ __align__(16) uint8_t boxes[4 * RL_BOX_BYTES];
  __shared__ double terms[4 * 64];
  particle_beams<METHOD, SENSOR>(m, lut, trig, max_range, ins, angles, obs, table, K, ranges, weights, p, M, S,
                                 j0, base, boxes + (threadIdx.x >> 6) * RL_BOX_BYTES, terms + (threadIdx.x >> 6) * 64,
                                 pre);
}

// Particles of one region chunk per workgroup (see Sched).
template <int METHOD, bool SENSOR>
__global__ __launch_bounds__(256) void k_particles_xcd(MapDev m, TrigDev trig, float max_range,
                                                       const float* __restrict__ ins,
                                                       const float* __restrict__ angles,
                                                       const float* __restrict__ obs,
                                                       const double* __restrict__ table, int K,
                                                       float* __restrict__ ranges,
                                                       double* __restrict__ weights, int M, int S,
                                                       Sched sched) {
  const int per_wave = 64 / S;
  const int per_block = (blockDim.x >> 6) * per_wave;
  const int r = blockIdx.x % RL_REGIONS;
  const long long first = (long long)(blockIdx.x / RL_REGIONS) * per_block;
  const int cnt = sched.count[r];
  if (first >= cnt) return;  // uniform across the workgroup
  const int lane = threadIdx.x & 63;
  const int wib = threadIdx.x >> 6;
  const int seg = lane / S, j0 = lane - seg * S, base = seg * S;
  const long long idx = first + wib * per_wave + seg;
  const int p = idx < cnt ? sched.perm[sched.offset[r] + idx] : -1;
  const BeamIn pre = load_beam_in<SENSOR>(ins, angles, obs, p, M, j0);
  __shared__ float lut[RL_LUT_SIZE];
  stage_lut<METHOD>(m, lut);
  __shared__ __align__(16) uint8_t boxes[4 * RL_BOX_BYTES];
  __shared__ double terms[4 * 64];
  particle_beams<METHOD, SENSOR>(m, lut, trig, max_range, ins, angles, obs, table, K, ranges, weights, p, M, S,
                                 j0, base, boxes + wib * RL_BOX_BYTES, terms + wib * 64, pre);
}

// GiantLUTCast's table by tiles of RL_GLT_TILE consecutive cells (x-major
// order, so mostly neighbours along y) and all their bins: a wave marches 4
// bins' directions from 16 neighbouring cells, rays of similar length, instead
// of 64 directions from one cell (a wave runs as long as its longest ray).
// The tile's entries are assembled in LDS in the table's order (cell-major,
// RL_GLT_TILE * td entries, the launch's dynamic LDS) and stored contiguously.
// Same values as k_glt_build, entry by entry.  Measured on the 2048^2
// corridor map with 108 bins: tiles of 4, 8, 16, 32, 64, 128, 256 cells took
// 4.84, 3.95, 3.69, 3.81, 4.03, 5.18, 11.4 ms (one entry per thread: 5.34 ms);
// larger tiles hold more LDS per workgroup and leave fewer of them per CU.
#ifndef RL_GLT_TILE
#define RL_GLT_TILE 16  // cells per tile of k_glt_build_tiles (a power of two)
#endif
#define RL_GLT_TILE_TD 256  // the largest td built by tiles
__global__ __launch_bounds__(256) void k_glt_build_tiles(MapDev m, float max_range, int td,
                                                         const float2* __restrict__ dirs, float limits_div_max,
                                                         uint16_t* __restrict__ out, long long ncells) {
  __shared__ float lut[RL_LUT_SIZE];
  extern __shared__ uint16_t tile[];  // RL_GLT_TILE * td entries
  stage_lut<RL_METHOD_RM>(m, lut);
  const long long ntiles = (ncells + RL_GLT_TILE - 1) / RL_GLT_TILE;
  for (long long t = blockIdx.x; t < ntiles; t += gridDim.x) {  // the same in every thread
    const long long first = t * RL_GLT_TILE;
    const int nc = (int)min((long long)RL_GLT_TILE, ncells - first);
    for (int r = threadIdx.x; r < RL_GLT_TILE * td; r += blockDim.x) {
      const int c = r & (RL_GLT_TILE - 1), bin = r / RL_GLT_TILE;  // a wave: 64 / RL_GLT_TILE bins
      if (c >= nc) continue;
      const long long cell = first + c;
      const int x = (int)(cell / m.height), y = (int)(cell - (long long)x * m.height);
      const float2 d = dirs[bin];
      const float fx = (float)x, fy = (float)y;
      int unused;
      // the plain march: the table is bound by the work of its many rays, not
      // by the longest chain, and speculation's extra reads cost more than
      // they save (2048^2 corridor map, 108 bins, one entry per thread: 5.31 ms
      // plain, 6.10 ms at width 2 after 8 reads, 5.64-5.78 ms after 16 or 32)
      float v = (m.idx8 && rl::fast_march_ok(fx, fy, max_range))
                    ? rl::rm_march_idx8<false, 1>(m, lut, rl::Box{nullptr, 0, 0}, max_range, fx, fy, d.x, d.y,
                                                  &unused)
                    : rl::rm_march<false>(m.edt, m.width, m.height, max_range, fx, fy, d.x, d.y, &unused);
      v = (v < max_range) ? v : max_range;  // std::min(max_range, r)
      tile[c * td + bin] = (uint16_t)(unsigned)(v * limits_div_max);
    }
    __syncthreads();
    uint16_t* dst = out + first * td;
    for (int e = threadIdx.x; e < nc * td; e += blockDim.x) dst[e] = tile[e];
    __syncthreads();  // the tile is stored before the next one is assembled
  }
}

// GiantLUTCast's ray directions: (cosf, sinf) of angle i * step for i < td, the
// values every ray of bin i uses (the reference evaluates them per ray; the same
// function of the same float gives the same bits).
__global__ void k_glt_dirs(TrigDev trig, int td, float step, float2* __restrict__ dirs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= td) return;
  float s, c;
  rl::sincos_libc((float)i * step, trig, &s, &c);
  dirs[i] = make_float2(c, s);
}

// GiantLUTCast table (RangeLib.h:1781-1816): entry k = (cell * td + i), cell =
// x * H + y, is the ray marched from the integer point (x, y) at angle
// i * step, clamped to max_range and quantized to uint16; `step` and
// `limits_div_max` are the reference's float constants.  One ray per thread
// (grid-stride); neighbouring threads share a start cell.  WIDE: 64-bit entry
// indices (tables of 2^32 entries or more), else 32-bit index arithmetic.
template <bool WIDE>
__global__ __launch_bounds__(256) void k_glt_build(MapDev m, float max_range, int td,
                                                   const float2* __restrict__ dirs, float limits_div_max,
                                                   uint16_t* __restrict__ out, long long total) {
  __shared__ float lut[RL_LUT_SIZE];
  stage_lut<RL_METHOD_RM>(m, lut);
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < total; k += stride) {
    int x, y, i;
    if (WIDE) {
      const long long cell = k / td;
      i = (int)(k - cell * td);
      x = (int)(cell / m.height);
      y = (int)(cell - (long long)x * m.height);
    } else {
      const unsigned k32 = (unsigned)k, cell = k32 / (unsigned)td;
      i = (int)(k32 - cell * (unsigned)td);
      x = (int)(cell / (unsigned)m.height);
      y = (int)(cell - (unsigned)x * (unsigned)m.height);
    }
    const float2 d = dirs[i];
    const float fx = (float)x, fy = (float)y;
    int unused;
    // the default speculation: the plain march instantiated here as well (as
    // k_glt_build_tiles marches) changed the compiler's code for the particle
    // kernels (k_particles<RM, sensor> 2073 -> 2557 VALU instructions, about
    // 1% slower); this kernel builds only tables of more than 256 bins
    float r = (m.idx8 && rl::fast_march_ok(fx, fy, max_range))
                  ? rl::rm_march_idx8<false>(m, lut, rl::Box{nullptr, 0, 0}, max_range, fx, fy, d.x, d.y, &unused)
                  : rl::rm_march<false>(m.edt, m.width, m.height, max_range, fx, fy, d.x, d.y, &unused);
    r = (r < max_range) ? r : max_range;  // std::min(max_range, r)
    out[k] = (uint16_t)(unsigned)(r * limits_div_max);
  }
}

// Compact index map from the fp32 distance transform (layout in rl_device.h).
// Occupancy bitmap tiles (layout in MapDev): one thread per 8x8 block of
// cells writes its x-major and its y-major word.  Cells past the map's edge
// are free.
__global__ void k_build_bits(const uint8_t* __restrict__ occ, unsigned long long* __restrict__ bits, int W, int H,
                             int ntx, int nty) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)ntx * nty) return;
  const int tx = (int)(i / nty), ty = (int)(i - (long long)tx * nty);
  unsigned long long xm = 0, ym = 0;
  for (int u = 0; u < 8 && tx * 8 + u < W; ++u)
    for (int v = 0; v < 8 && ty * 8 + v < H; ++v)
      if (occ[(size_t)(tx * 8 + u) * H + (size_t)(ty * 8 + v)]) {
        xm |= 1ull << (8 * u + v);
        ym |= 1ull << (8 * v + u);
      }
  bits[(size_t)tx * nty + ty] = xm;
  bits[(size_t)ntx * nty + (size_t)ty * ntx + tx] = ym;
}

__global__ void k_build_idx8(const float* __restrict__ edt, const float* __restrict__ lut,
                             uint8_t* __restrict__ idx8, int W, int H, int nby, size_t nbytes) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nbytes) return;
  size_t b = i >> 6;
  unsigned o = (unsigned)(i & 63);
  unsigned px = (unsigned)(b / (size_t)nby) * 8u + (o & 7u);
  unsigned py = (unsigned)(b % (size_t)nby) * 8u + (o >> 3);
  unsigned code = RL_IDX8_ESCAPE;
  if (px < (unsigned)W && py < (unsigned)H) {
    float d = edt[(size_t)px * H + py];
    if (d <= 0.0f) {
      code = 0u;  // the march stops at any d <= 0 and does not use its value
    } else {
      int lo = 0, hi = RL_IDX8_ESCAPE - 1;  // lut[0..254] strictly increasing
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (lut[mid] < d) lo = mid + 1; else hi = mid;
      }
      if (lut[lo] == d) code = (unsigned)lo;
    }
  }
  idx8[i] = (uint8_t)code;
}

__global__ void k_debug_sincos(TrigDev trig, const float* __restrict__ x, float* __restrict__ s,
                               float* __restrict__ c, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float a, b;
  rl::sincos_libc(x[i], trig, &a, &b);
  s[i] = a;
  c[i] = b;
}

__global__ void k_debug_sqrt(const float* __restrict__ x, float* __restrict__ out, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = __builtin_sqrtf(x[i]);
}

// Dependent-load latency: one lane follows next[] for `steps` loads (each
// address is the previous load's value) and records the elapsed wall-clock
// ticks and the final index (so the loads cannot be dropped).  The caller
// bounds every next[] entry by the buffer's length.
__global__ void k_debug_chase(const unsigned* __restrict__ next, int steps, long long* __restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  unsigned i = 0;
  const long long t0 = wall_clock64();
  for (int k = 0; k < steps; ++k) i = rl::glb(next)[i];
  // the last load completes before the second clock read
  __asm__ volatile("s_waitcnt vmcnt(0)" ::"v"(i) : "memory");
  const long long t1 = wall_clock64();
  out[0] = t1 - t0;
  out[1] = (long long)i;
}

// ---------------------------------------------------------------------------
// Host side

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(expr)                                                              \
  do {                                                                             \
    hipError_t e_ = (expr);                                                        \
    if (e_ != hipSuccess) return fail(RL_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

#define CHECK_ARG(cond, msg)                       \
  do {                                             \
    if (!(cond)) return fail(RL_ERR_ARG, "%s", msg); \
  } while (0)

struct TrigFile {
  bool loaded = false;
  std::string error;
  double delta_s = 0, delta_c = 0;
  std::vector<uint32_t> off_s, off_c;
  std::vector<uint16_t> keys_s, keys_c;
  int e_min_s = 0, e_min_c = 0;
  std::vector<uint8_t> bits_s, bits_c;  // flat exception bitmaps (rl_trig.h)
  std::string libc_version;
};

struct TrigHeader {
  char magic[8];
  uint32_t nbuckets, abs_limit, nkeys_sin, nkeys_cos;
  double delta_sin, delta_cos;
  char libc_version[32];
};

std::mutex g_trig_mu;
TrigFile g_trig;

std::string library_dir() {
  Dl_info info;
  if (dladdr((void*)&rl_create, &info) && info.dli_fname) {
    std::vector<char> p(info.dli_fname, info.dli_fname + strlen(info.dli_fname) + 1);
    return std::string(dirname(p.data()));
  }
  return ".";
}

// Loads the sinf/cosf exception table once per process.
const TrigFile* trig_table(std::string* err) {
  std::lock_guard<std::mutex> lk(g_trig_mu);
  if (g_trig.loaded) return &g_trig;
  const char* env = getenv("RL_TRIG_TABLE");
  std::string path = env ? std::string(env) : library_dir() + "/_data/trig_exc.bin";
  FILE* fp = fopen(path.c_str(), "rb");
  if (!fp) {
    *err = "trig exception table not found: " + path + " (run the build)";
    return nullptr;
  }
  TrigHeader h;
  bool ok = fread(&h, sizeof(h), 1, fp) == 1 && memcmp(h.magic, "RLTRIG01", 8) == 0 &&
            h.nbuckets == RLT_NBUCKETS && h.abs_limit == RLT_ABS_LIMIT;
  if (ok) {
    g_trig.off_s.resize(h.nbuckets + 1);
    g_trig.off_c.resize(h.nbuckets + 1);
    g_trig.keys_s.resize(h.nkeys_sin);
    g_trig.keys_c.resize(h.nkeys_cos);
    ok = fread(g_trig.off_s.data(), 4, h.nbuckets + 1, fp) == h.nbuckets + 1 &&
         fread(g_trig.keys_s.data(), 2, h.nkeys_sin, fp) == h.nkeys_sin &&
         fread(g_trig.off_c.data(), 4, h.nbuckets + 1, fp) == h.nbuckets + 1 &&
         fread(g_trig.keys_c.data(), 2, h.nkeys_cos, fp) == h.nkeys_cos;
    ok = ok && g_trig.off_s.back() == h.nkeys_sin && g_trig.off_c.back() == h.nkeys_cos;
  }
  fclose(fp);
  if (!ok) {
    *err = "corrupt trig exception table: " + path;
    g_trig = TrigFile();
    return nullptr;
  }
  g_trig.delta_s = h.delta_sin;
  g_trig.delta_c = h.delta_cos;
  g_trig.e_min_s = rltrig::table_min_exponent(g_trig.off_s.data(), g_trig.keys_s.data());
  g_trig.e_min_c = rltrig::table_min_exponent(g_trig.off_c.data(), g_trig.keys_c.data());
  g_trig.bits_s.assign(rltrig::bitmap_bytes(g_trig.e_min_s), 0);
  g_trig.bits_c.assign(rltrig::bitmap_bytes(g_trig.e_min_c), 0);
  rltrig::table_to_bitmap(g_trig.off_s.data(), g_trig.keys_s.data(), g_trig.e_min_s, g_trig.bits_s.data());
  rltrig::table_to_bitmap(g_trig.off_c.data(), g_trig.keys_c.data(), g_trig.e_min_c, g_trig.bits_c.data());
  h.libc_version[31] = 0;
  g_trig.libc_version = h.libc_version;
  g_trig.loaded = true;
  return &g_trig;
}

// The exception table lists where the C library that generated it rounds
// sinf/cosf away from the nearest float; the reference calls the C library of
// the running process, so the two must be the same release for bit parity.
// RL_TRIG_ALLOW_LIBC_MISMATCH=1 accepts another release (results may then
// differ from the reference's in the last place of a ray direction).
int check_trig_libc(const TrigFile* tf) {
  const char* running = gnu_get_libc_version();
  if (tf->libc_version == running) return RL_OK;
  const char* allow = getenv("RL_TRIG_ALLOW_LIBC_MISMATCH");
  if (allow && strcmp(allow, "1") == 0) return RL_OK;
  return fail(RL_ERR_STATE,
              "sinf/cosf exception table was generated with C library %s but this process runs %s: "
              "rebuild the table (python -m range_libc_amd.build --force) or set "
              "RL_TRIG_ALLOW_LIBC_MISMATCH=1",
              tf->libc_version.c_str(), running);
}

int grid_for(long long n, int block) { return (int)((n + block - 1) / block); }

int segment_size(int m) {
  int s = 1;
  while (s < m && s < 64) s <<= 1;
  return s;
}

}  // namespace

#define RL_PINNED_BYTES (512u << 10)  // host-pointer calls whose parts fit run zero-copy (rl_ctx::pinned)

struct rl_ctx {
  int device = 0;
  hipStream_t own_stream = nullptr;
  hipStream_t stream = nullptr;
  void* trig_buf = nullptr;
  TrigDev trig{};
  float* d_lut = nullptr;  // RL_LUT_SIZE tabulated distances, then their march steps (see rl_device.h)
  int xcd_sched = 0;          // RL_OPT_XCD_SCHED
  int lds_box = 0;            // RL_OPT_LDS_BOX
  // the caller's handle plus one per map / batch created on this context: the
  // context is freed when the last of them is destroyed, in any order
  std::atomic<int> refs{1};
  std::mutex mu;
  void* d_scratch = nullptr;  // this device's copy of `scratch` (no-allocation entry point)
  void* stage[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  size_t stage_cap[6] = {0, 0, 0, 0, 0, 0};
  void* pin = nullptr;  // RL_PINNED_BYTES of pinned host memory for small host-pointer calls

  // Small host-pointer calls copy their inputs into this pinned block on the
  // host and the kernel reads and writes it in place (it is mapped into the
  // device's address space): one launch and one synchronisation per call
  // instead of two or more copies through device staging.  Calls hold the
  // context's mutex until they have synchronised, so one block serves all.
  int pinned(void** out) {
    if (!pin && hipHostMalloc(&pin, RL_PINNED_BYTES, hipHostMallocDefault) != hipSuccess) {
      (void)hipGetLastError();
      pin = nullptr;
      return fail(RL_ERR_NOMEM, "hipHostMalloc of %u pinned bytes failed", (unsigned)RL_PINNED_BYTES);
    }
    *out = pin;
    return RL_OK;
  }

  // Device staging buffer `slot` of at least `bytes` (host-pointer entry points).
  int staging(int slot, size_t bytes, void** out) {
    if (bytes == 0) bytes = 16;
    if (stage_cap[slot] < bytes) {
      if (stage[slot]) HIP_TRY(hipFree(stage[slot]));
      stage[slot] = nullptr;
      stage_cap[slot] = 0;
      size_t cap = bytes + bytes / 4;
      if (hipMalloc(&stage[slot], cap) != hipSuccess) {
        (void)hipGetLastError();
        return fail(RL_ERR_NOMEM, "hipMalloc of %zu staging bytes failed", cap);
      }
      stage_cap[slot] = cap;
    }
    *out = stage[slot];
    return RL_OK;
  }
};

struct rl_map {
  rl_ctx* ctx = nullptr;
  // the caller's handle plus one per surrogate and per batch job list holding this map
  std::atomic<int> refs{1};
  // bumped whenever a device buffer or setting that a launch descriptor copies
  // changes (world frame, sensor table, GiantLUTCast and CDDT tables): batches
  // re-upload their descriptors when it moved
  std::atomic<unsigned long long> gen{0};
  int width = 0, height = 0;
  float* d_edt = nullptr;
  uint8_t* d_occ = nullptr;
  uint8_t* d_idx8 = nullptr;  // compact blocked index map (null when disabled)
  uint8_t* d_bits = nullptr;  // occupancy bitmap tiles, both orientations (see MapDev)
  unsigned bits_bytes = 0;
  int* d_sched = nullptr;     // region buckets: count | offset | pad | perm
  int sched_cap = 0;
  double* d_table = nullptr;
  int table_width = 0;
  int nby = 0;
  WorldDev world{};
  float edt_ms = 0.0f;
  uint16_t* d_glt = nullptr;  // GiantLUTCast table (rl_map_build_giant_lut)
  int glt_td = 0;
  float glt_R = 0.0f, glt_scale = 0.0f, glt_td_div_2pi = 0.0f, glt_ms = 0.0f;
  // CDDTCast tables (rl_map_build_cddt; layout in rl_device.h)
  int* cddt_width = nullptr;
  float *cddt_trans = nullptr, *cddt_cos = nullptr, *cddt_sin = nullptr, *cddt_vals = nullptr;
  long long *cddt_bin_off = nullptr, *cddt_start = nullptr;
  int cddt_td = 0, cddt_nang = 0;
  long long cddt_nb = 0, cddt_nv = 0;
  float cddt_td_div_2pi = 0.0f, cddt_ms = 0.0f;

  MapDev dev() const {
    return MapDev{d_edt, d_occ, d_idx8, ctx->d_lut + RL_LUT_SIZE, width, height, nby, ctx->lds_box, world,
                  d_glt, glt_td, glt_R, glt_scale, glt_td_div_2pi,
                  rl::CddtDev{cddt_width, cddt_trans, cddt_cos, cddt_sin, cddt_bin_off, cddt_start, cddt_vals,
                              cddt_td, cddt_td_div_2pi},
                  d_bits, bits_bytes, bits_bytes / 2u, (width + 7) / 8, (height + 7) / 8};
  }
};

namespace {

void set_world(rl_map* m, float scale, float angle, float ox, float oy, float sin_a, float cos_a) {
  // RangeLib.h:442-450 / CudaRangeLib.h:26-34
  m->world.inv_scale = (float)(1.0 / (double)scale);
  m->world.scale = scale;
  m->world.origin_x = ox;
  m->world.origin_y = oy;
  m->world.sin_a = sin_a;
  m->world.cos_a = cos_a;
  m->world.rotation_const = (float)(-1.0 * (double)angle - 3.0 * 3.14159265358979323846 / 2.0);
}

int check_method(int method) {
  if (method < RL_METHOD_RM || method > RL_METHOD_CDDT)
    return fail(RL_ERR_ARG, "unknown range method %d", method);
  return RL_OK;
}

// The method can run on this map at this max_range: ray marching needs the
// distance transform; a GiantLUTCast table answers only for the max_range it
// was built with (its quantization and out-of-map value depend on it).
int check_map(const rl_map* m, int method, float max_range);

#define TRY(expr)              \
  do {                         \
    int rc_ = (expr);          \
    if (rc_ != RL_OK) return rc_; \
  } while (0)

template <int METHOD, int FRAME, bool COUNT>
int launch_cast_t(rl_map* m, float max_range, const float* ins, const float* angles, float* outs,
                  int* counts, int n_rays, int num_angles) {
  if (n_rays == 0) return RL_OK;
  if (METHOD == RL_METHOD_BL && !COUNT && m->d_bits) {
    hipLaunchKernelGGL((k_cast_bl<FRAME>), dim3(grid_for(n_rays, RL_BL_WG)), dim3(RL_BL_WG), 0, m->ctx->stream,
                       m->dev(), m->ctx->trig, max_range, ins, angles, outs, n_rays, num_angles);
    HIP_TRY(hipGetLastError());
    return RL_OK;
  }
  hipLaunchKernelGGL((k_cast<METHOD, FRAME, COUNT>), dim3(grid_for(n_rays, 256)), dim3(256), 0,
                     m->ctx->stream, m->dev(), m->ctx->trig, max_range, ins, angles, outs, counts,
                     n_rays, num_angles);
  HIP_TRY(hipGetLastError());
  return RL_OK;
}

template <int FRAME, bool COUNT>
int launch_cast_f(rl_map* m, int method, float max_range, const float* ins, const float* angles,
                  float* outs, int* counts, int n_rays, int num_angles) {
  if (method == RL_METHOD_BL)
    return launch_cast_t<RL_METHOD_BL, FRAME, COUNT>(m, max_range, ins, angles, outs, counts, n_rays, num_angles);
  if (method == RL_METHOD_GLT)
    return launch_cast_t<RL_METHOD_GLT, FRAME, COUNT>(m, max_range, ins, angles, outs, counts, n_rays, num_angles);
  if (method == RL_METHOD_CDDT)
    return launch_cast_t<RL_METHOD_CDDT, FRAME, COUNT>(m, max_range, ins, angles, outs, counts, n_rays, num_angles);
  return launch_cast_t<RL_METHOD_RM, FRAME, COUNT>(m, max_range, ins, angles, outs, counts, n_rays, num_angles);
}

int launch_particles(rl_map* m, int method, bool sensor, float max_range, const float* ins,
                     const float* angles, const float* obs, float* ranges, double* weights, int N, int M);

int launch_cast(rl_map* m, int method, int frame, bool count, float max_range, const float* ins,
                const float* angles, float* outs, int* counts, int n_rays, int num_angles) {
  TRY(check_map(m, method, max_range));
  if (frame == RL_FRAME_GRID)
    return count ? launch_cast_f<0, true>(m, method, max_range, ins, angles, outs, counts, n_rays, num_angles)
                 : launch_cast_f<0, false>(m, method, max_range, ins, angles, outs, counts, n_rays, num_angles);
  if (frame == RL_FRAME_WORLD)
    return count ? launch_cast_f<1, true>(m, method, max_range, ins, angles, outs, counts, n_rays, num_angles)
                 : launch_cast_f<1, false>(m, method, max_range, ins, angles, outs, counts, n_rays, num_angles);
  if (frame == RL_FRAME_ANGLES && !count && num_angles > 0)
    return launch_particles(m, method, false, max_range, ins, angles, nullptr, outs, nullptr,
                            n_rays / num_angles, num_angles);
  if (frame == RL_FRAME_ANGLES)
    return count ? launch_cast_f<2, true>(m, method, max_range, ins, angles, outs, counts, n_rays, num_angles)
                 : launch_cast_f<2, false>(m, method, max_range, ins, angles, outs, counts, n_rays, num_angles);
  return fail(RL_ERR_ARG, "unknown frame %d", frame);
}

int launch_sensor(rl_map* m, const float* obs, const float* ranges, double* outs, int M, int N) {
  if (!m->d_table) return fail(RL_ERR_STATE, "no sensor model set (call set_sensor_model first)");
  if (N == 0) return RL_OK;
  int S = segment_size(M);
  long long waves = ((long long)N + 64 / S - 1) / (64 / S);
  hipLaunchKernelGGL(k_sensor, dim3(grid_for(waves * 64, 256)), dim3(256), 0, m->ctx->stream, obs,
                     ranges, outs, m->d_table, m->table_width, m->world.inv_scale, M, N, S);
  HIP_TRY(hipGetLastError());
  return RL_OK;
}

int launch_fused(rl_map* m, int method, float max_range, const float* ins, const float* angles,
                 const float* obs, double* weights, int N, int M) {
  TRY(check_map(m, method, max_range));
  if (!m->d_table) return fail(RL_ERR_STATE, "no sensor model set (call set_sensor_model first)");
  if (N == 0) return RL_OK;
  int S = segment_size(M);
  long long waves = ((long long)N + 64 / S - 1) / (64 / S);
  dim3 grid(grid_for(waves * 64, 256)), block(256);
  if (method == RL_METHOD_BL)
    hipLaunchKernelGGL(k_fused<RL_METHOD_BL>, grid, block, 0, m->ctx->stream, m->dev(), m->ctx->trig,
                       max_range, ins, angles, obs, m->d_table, m->table_width, weights, N, M, S);
  else if (method == RL_METHOD_GLT)
    hipLaunchKernelGGL(k_fused<RL_METHOD_GLT>, grid, block, 0, m->ctx->stream, m->dev(), m->ctx->trig,
                       max_range, ins, angles, obs, m->d_table, m->table_width, weights, N, M, S);
  else if (method == RL_METHOD_CDDT)
    hipLaunchKernelGGL(k_fused<RL_METHOD_CDDT>, grid, block, 0, m->ctx->stream, m->dev(), m->ctx->trig,
                       max_range, ins, angles, obs, m->d_table, m->table_width, weights, N, M, S);
  else
    hipLaunchKernelGGL(k_fused<RL_METHOD_RM>, grid, block, 0, m->ctx->stream, m->dev(), m->ctx->trig,
                       max_range, ins, angles, obs, m->d_table, m->table_width, weights, N, M, S);
  HIP_TRY(hipGetLastError());
  return RL_OK;
}

// Region buckets for XCD-aware scheduling, grown to hold n particles.
int ensure_sched(rl_map* m, int n) {
  if (m->sched_cap >= n && m->d_sched) return RL_OK;
  if (m->d_sched) HIP_TRY(hipFree(m->d_sched));
  m->d_sched = nullptr;
  m->sched_cap = 0;
  size_t cap = (size_t)n + (size_t)n / 4 + 64;
  if (hipMalloc(&m->d_sched, (cap + 4 * RL_REGIONS) * sizeof(int)) != hipSuccess) {
    (void)hipGetLastError();
    m->d_sched = nullptr;
    return fail(RL_ERR_NOMEM, "hipMalloc of the scheduling buffer failed (inside graph capture, "
                              "make one uncaptured call of this size first)");
  }
  m->sched_cap = (int)cap;
  return RL_OK;
}

template <int METHOD, bool SENSOR>
int launch_particles_t(rl_map* m, float max_range, const float* ins, const float* angles,
                       const float* obs, float* ranges, double* weights, int N, int M) {
  rl_ctx* c = m->ctx;
  const int S = segment_size(M);
  const int K = SENSOR ? m->table_width : 1;
  const double* table = SENSOR ? m->d_table : nullptr;
  if (c->xcd_sched && N <= RL_XCD_SCHED_MAX) {
    TRY(ensure_sched(m, N));
    Sched sc{m->d_sched + 4 * RL_REGIONS, m->d_sched, m->d_sched + RL_REGIONS};
    hipLaunchKernelGGL(k_region, dim3(1), dim3(1024), 0, c->stream, m->dev(), ins, N, sc);
    HIP_TRY(hipGetLastError());
    // enough chunks for all particles to fall in one region; empty chunks exit
    long long per_block = 4LL * (64 / S);
    long long blocks = RL_REGIONS * ((N + per_block - 1) / per_block);
    hipLaunchKernelGGL((k_particles_xcd<METHOD, SENSOR>), dim3((unsigned)blocks), dim3(256), 0, c->stream,
                       m->dev(), c->trig, max_range, ins, angles, obs, table, K, ranges, weights, M, S, sc);
  } else {
    long long waves = ((long long)N + 64 / S - 1) / (64 / S);
    hipLaunchKernelGGL((k_particles<METHOD, SENSOR>), dim3(grid_for(waves * 64, 256)), dim3(256), 0,
                       c->stream, m->dev(), c->trig, max_range, ins, angles, obs, table, K, ranges, weights,
                       N, M, S);
  }
  HIP_TRY(hipGetLastError());
  return RL_OK;
}

int launch_particles(rl_map* m, int method, bool sensor, float max_range, const float* ins,
                     const float* angles, const float* obs, float* ranges, double* weights, int N, int M) {
  TRY(check_map(m, method, max_range));
  if (sensor && !m->d_table) return fail(RL_ERR_STATE, "no sensor model set (call set_sensor_model first)");
  if (N == 0) return RL_OK;
  if (method == RL_METHOD_BL)
    return sensor ? launch_particles_t<RL_METHOD_BL, true>(m, max_range, ins, angles, obs, ranges, weights, N, M)
                  : launch_particles_t<RL_METHOD_BL, false>(m, max_range, ins, angles, obs, ranges, weights, N, M);
  if (method == RL_METHOD_GLT)
    return sensor ? launch_particles_t<RL_METHOD_GLT, true>(m, max_range, ins, angles, obs, ranges, weights, N, M)
                  : launch_particles_t<RL_METHOD_GLT, false>(m, max_range, ins, angles, obs, ranges, weights, N, M);
  if (method == RL_METHOD_CDDT)
    return sensor ? launch_particles_t<RL_METHOD_CDDT, true>(m, max_range, ins, angles, obs, ranges, weights, N, M)
                  : launch_particles_t<RL_METHOD_CDDT, false>(m, max_range, ins, angles, obs, ranges, weights, N, M);
  return sensor ? launch_particles_t<RL_METHOD_RM, true>(m, max_range, ins, angles, obs, ranges, weights, N, M)
                : launch_particles_t<RL_METHOD_RM, false>(m, max_range, ins, angles, obs, ranges, weights, N, M);
}

// Builds the distance transform of m->d_occ into m->d_edt.
// Compact map from m->d_edt (unless RL_MAP_NO_COMPACT).  Every value <= 0
// gets code 0 (a hit, as in the fp32 march); any other value that is not a
// tabulated distance (NaN, not sqrtf of a small integer) gets the escape code,
// so the compact march reads it from the fp32 map and behaves as the fp32
// march.
int build_compact(rl_map* m, unsigned flags) {
  if (flags & RL_MAP_NO_COMPACT) return RL_OK;
  rl_ctx* c = m->ctx;
  const int W = m->width, H = m->height;
  m->nby = (H + 7) / 8;
  const size_t nbytes = (size_t)((W + 7) / 8) * (size_t)m->nby * 64;
  if (nbytes > 0xFFFFFFFFull) return RL_OK;  // 32-bit offsets only: such maps march the fp32 map
  if (hipMalloc(&m->d_idx8, nbytes) != hipSuccess) {
    (void)hipGetLastError();
    m->d_idx8 = nullptr;
    return fail(RL_ERR_NOMEM, "device allocation of the %zu-byte compact map failed", nbytes);
  }
  hipLaunchKernelGGL(k_build_idx8, dim3(grid_for((long long)nbytes, 256)), dim3(256), 0, c->stream,
                     (const float*)m->d_edt, (const float*)c->d_lut, m->d_idx8, W, H, m->nby, nbytes);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(RL_ERR_HIP, "compact map build: %s", hipGetErrorString(e));
  return RL_OK;
}

// Occupancy bitmap from m->d_occ (both orientations, 16 bytes of padding so
// that a window read at the last tile stays in the buffer).  Maps whose bitmap
// would pass 4 GiB keep walking the byte grid.
int build_bits(rl_map* m) {
  rl_ctx* c = m->ctx;
  const int ntx = (m->width + 7) / 8, nty = (m->height + 7) / 8;
  const size_t nbytes = (size_t)ntx * nty * 16;
  if (nbytes + 16 > 0xFFFFFFFFull) return RL_OK;
  if (hipMalloc(&m->d_bits, nbytes + 16) != hipSuccess) {
    (void)hipGetLastError();
    m->d_bits = nullptr;
    return fail(RL_ERR_NOMEM, "device allocation of the %zu-byte occupancy bitmap failed", nbytes + 16);
  }
  m->bits_bytes = (unsigned)nbytes;
  hipError_t e = hipMemsetAsync(m->d_bits + nbytes, 0, 16, c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_build_bits, dim3(grid_for((long long)ntx * nty, 256)), dim3(256), 0, c->stream,
                       (const uint8_t*)m->d_occ, (unsigned long long*)m->d_bits, m->width, m->height, ntx, nty);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(RL_ERR_HIP, "occupancy bitmap build: %s", hipGetErrorString(e));
  return RL_OK;
}

int build_edt(rl_map* m, unsigned flags) {
  rl_ctx* c = m->ctx;
  const int W = m->width, H = m->height;
  const size_t N = (size_t)W * H;
  float* g = nullptr;
  int* v = nullptr;
  double* z = nullptr;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  int rc = RL_OK;
  do {
    if (hipMalloc(&g, N * sizeof(float)) != hipSuccess ||
        hipMalloc(&v, (N + (size_t)H) * sizeof(int)) != hipSuccess ||
        hipMalloc(&z, (N + (size_t)W + (size_t)H) * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      rc = fail(RL_ERR_NOMEM, "device allocation for the %dx%d distance transform failed", W, H);
      break;
    }
    if (hipEventRecord(e0, c->stream) != hipSuccess) {
      rc = fail(RL_ERR_HIP, "hipEventRecord failed");
      break;
    }
    if (H <= 4096 && !(flags & RL_MAP_EDT_ENVELOPE_ONLY)) {
      hipLaunchKernelGGL(k_edt_cols, dim3(W), dim3(256), 0, c->stream, m->d_occ, g, H);
    } else {
      hipLaunchKernelGGL(k_edt_init, dim3(grid_for((long long)N, 256)), dim3(256), 0, c->stream,
                         m->d_occ, m->d_edt, N);
      hipLaunchKernelGGL(k_edt_envelope, dim3(grid_for(W, 64)), dim3(64), 0, c->stream,
                         (const float*)m->d_edt, g, W, H, 1LL, (long long)H, v, z, 0,
                         (const int*)nullptr);
    }
    // Pass 2: the bounded search where it is exact, the lower envelope on the
    // rows it marks (or on every row when the envelope is asked for).
    const int* rows = nullptr;
    // the row search needs EDT_RB rows of the map in LDS (plus its flags) within 64 KiB
    if ((size_t)EDT_RB * W * sizeof(float) + 64 <= 65536 && !(flags & RL_MAP_EDT_ENVELOPE_ONLY)) {
      int* slow = v + N;
      hipLaunchKernelGGL(k_edt_rows, dim3((H + EDT_RB - 1) / EDT_RB), dim3(256),
                         (size_t)EDT_RB * W * sizeof(float), c->stream, (const float*)g,
                         m->d_edt, W, H, slow, 1);
      rows = slow;
    }
    hipLaunchKernelGGL(k_edt_envelope, dim3(grid_for(H, 64)), dim3(64), 0, c->stream,
                       (const float*)g, m->d_edt, H, W, (long long)H, 1LL, v, z, 1, rows);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      rc = fail(RL_ERR_HIP, "distance transform: %s", hipGetErrorString(e));
      break;
    }
    (void)hipEventElapsedTime(&m->edt_ms, e0, e1);
    rc = build_compact(m, flags);
  } while (0);
  if (g) (void)hipFree(g);
  if (v) (void)hipFree(v);
  if (z) (void)hipFree(z);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int check_map(const rl_map* m, int method, float max_range) {
  TRY(check_method(method));
  if (method == RL_METHOD_RM && !m->d_edt) return fail(RL_ERR_STATE, "map has no distance transform");
  if (method == RL_METHOD_CDDT && !m->cddt_vals)
    return fail(RL_ERR_STATE, "no CDDT tables (call rl_map_build_cddt first)");
  if (method == RL_METHOD_GLT) {
    if (!m->d_glt) return fail(RL_ERR_STATE, "no GiantLUTCast table (call rl_map_build_giant_lut first)");
    if (!(max_range == m->glt_R))
      return fail(RL_ERR_ARG, "max_range %g differs from the GiantLUTCast table's %g", (double)max_range,
                  (double)m->glt_R);
  }
  return RL_OK;
}

int use_device(rl_ctx* c) {
  int cur = -1;
  if (hipGetDevice(&cur) == hipSuccess && cur == c->device) return RL_OK;
  HIP_TRY(hipSetDevice(c->device));
  return RL_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

const char* rl_last_error(void) { return g_err.c_str(); }

int rl_abi_version(void) { return RL_ABI_VERSION; }

int rl_png_occupancy(const char* path, float threshold, uint8_t* occ_xmajor, int* width, int* height) {
  CHECK_ARG(path && width && height, "null argument");
  rlpng::Image img;
  std::string err;
  if (!rlpng::decode_file(path, &img, &err)) return fail(RL_ERR_IO, "%s: %s", path, err.c_str());
  if (img.width > 65535 || img.height > 65535)
    return fail(RL_ERR_ARG, "%s: map size %ux%u outside 1..65535", path, img.width, img.height);
  *width = (int)img.width;
  *height = (int)img.height;
  if (!occ_xmajor) return RL_OK;
  const size_t W = img.width, H = img.height;
  for (size_t y = 0; y < H; ++y)
    for (size_t x = 0; x < W; ++x) {
      const uint8_t* px = &img.rgba[4 * (y * W + x)];
      // channels taken as (r, g, b) = (third, second, first) byte, as
      // OMap(filename, threshold) reads them (RangeLib.h:190-193); gray =
      // 0.229 r + 0.587 g + 0.114 b in double, returned as float and
      // truncated (RangeUtils.h:30-32, RangeLib.h:194)
      double v = 0.229 * (double)px[2];
      v = v + 0.587 * (double)px[1];
      v = v + 0.114 * (double)px[0];
      const int gray = (int)(float)v;
      occ_xmajor[x * H + y] = ((float)gray < threshold) ? 1 : 0;
    }
  return RL_OK;
}

int rl_png_save_occupancy(const char* path, const uint8_t* occ_xmajor, int width, int height) {
  CHECK_ARG(path && occ_xmajor, "null argument");
  CHECK_ARG(width >= 1 && height >= 1, "bad map size");
  const size_t W = (size_t)width, H = (size_t)height;
  std::vector<uint8_t> rgba(4 * W * H, 255);  // opaque white (RangeLib.h:274-277)
  for (size_t y = 0; y < H; ++y)
    for (size_t x = 0; x < W; ++x)
      if (occ_xmajor[x * H + y]) {  // black where occupied (RangeLib.h:279-283)
        uint8_t* px = &rgba[4 * (y * W + x)];
        px[0] = px[1] = px[2] = 0;
      }
  std::string err;
  if (!rlpng::encode_file(path, rgba.data(), (unsigned)W, (unsigned)H, &err)) return fail(RL_ERR_IO, "%s", err.c_str());
  return RL_OK;
}

int rl_trig_table_info(char* table_libc, char* running_libc, int len) {
  CHECK_ARG(len >= 0 && (len == 0 || (table_libc && running_libc)), "bad argument");
  std::string err;
  const TrigFile* tf = trig_table(&err);
  if (!tf) return fail(RL_ERR_IO, "%s", err.c_str());
  if (len > 0) {
    snprintf(table_libc, (size_t)len, "%s", tf->libc_version.c_str());
    snprintf(running_libc, (size_t)len, "%s", gnu_get_libc_version());
  }
  return check_trig_libc(tf);
}

int rl_create(int device, rl_ctx** out) {
  CHECK_ARG(out, "out is NULL");
  *out = nullptr;
  int ndev = 0;
  HIP_TRY(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(RL_ERR_ARG, "device %d not present (%d devices)", device, ndev);
  std::string err;
  const TrigFile* tf = trig_table(&err);
  if (!tf) return fail(RL_ERR_IO, "%s", err.c_str());
  TRY(check_trig_libc(tf));
  HIP_TRY(hipSetDevice(device));
  rl_ctx* c = new rl_ctx;
  c->device = device;
  if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return fail(RL_ERR_HIP, "hipStreamCreate failed");
  }
  c->stream = c->own_stream;
  // this device's address of the no-allocation staging array
  if (hipGetSymbolAddress(&c->d_scratch, HIP_SYMBOL(scratch)) != hipSuccess || !c->d_scratch) {
    (void)hipGetLastError();
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return fail(RL_ERR_HIP, "hipGetSymbolAddress of the staging array failed");
  }
  // One device buffer: sin bitmap | cos bitmap
  size_t b_s = tf->bits_s.size(), b_c = tf->bits_c.size();
  size_t total = b_s + b_c + 256;
  if (hipMalloc(&c->trig_buf, total) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return fail(RL_ERR_NOMEM, "hipMalloc of the trig exception bitmaps failed");
  }
  char* base = (char*)c->trig_buf;
  hipError_t e = hipSuccess;
  if (b_s) e = hipMemcpy(base, tf->bits_s.data(), b_s, hipMemcpyHostToDevice);
  if (e == hipSuccess && b_c) e = hipMemcpy(base + b_s, tf->bits_c.data(), b_c, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(c->trig_buf);
    (void)hipStreamDestroy(c->own_stream);
    delete c;
    return fail(RL_ERR_HIP, "trig table upload: %s", hipGetErrorString(e));
  }
  c->trig.sin_exc = rltrig::Bitmap{(const uint8_t*)base, tf->e_min_s, tf->delta_s};
  c->trig.cos_exc = rltrig::Bitmap{(const uint8_t*)(base + b_s), tf->e_min_c, tf->delta_c};
  // distance table: sqrtf of the 255 smallest sums of two squares
  {
    std::vector<char> is_sum(1024, 0);
    for (int a = 0; a * a < 1024; ++a)
      for (int b = a; a * a + b * b < 1024; ++b) is_sum[a * a + b * b] = 1;
    // lut[0, RL_LUT_SIZE): the distances; lut[RL_LUT_SIZE + i]: the march step
    // from distance lut[i], as the march computes it
    float lut[2 * RL_LUT_SIZE];
    int k = 0;
    for (int n = 0; n < 1024 && k < (int)RL_IDX8_ESCAPE; ++n)
      if (is_sum[n]) lut[k++] = sqrtf((float)n);
    for (; k < RL_LUT_SIZE; ++k) lut[k] = -1.0f;  // never matches a distance
    for (int i = 0; i < RL_LUT_SIZE; ++i) {
      const float st = lut[i] * 0.999f;
      lut[RL_LUT_SIZE + i] = (st < 1.0f) ? 1.0f : st;
    }
    if (hipMalloc(&c->d_lut, sizeof(lut)) != hipSuccess ||
        hipMemcpy(c->d_lut, lut, sizeof(lut), hipMemcpyHostToDevice) != hipSuccess) {
      (void)hipGetLastError();
      rl_destroy(c);
      return fail(RL_ERR_HIP, "distance table upload failed");
    }
  }
  *out = c;
  return RL_OK;
}

int rl_destroy(rl_ctx* c) {
  if (!c) return RL_OK;
  if (c->refs.fetch_sub(1) != 1) return RL_OK;  // maps or batches still hold it
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  for (int i = 0; i < 6; ++i)
    if (c->stage[i]) (void)hipFree(c->stage[i]);
  if (c->pin) (void)hipHostFree(c->pin);
  if (c->trig_buf) (void)hipFree(c->trig_buf);
  if (c->d_lut) (void)hipFree(c->d_lut);
  if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
  delete c;
  return RL_OK;
}

int rl_set_stream(rl_ctx* c, void* stream) {
  CHECK_ARG(c, "ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  c->stream = stream ? (hipStream_t)stream : c->own_stream;
  return RL_OK;
}

int rl_sync(rl_ctx* c) {
  CHECK_ARG(c, "ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(use_device(c));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RL_OK;
}

int rl_map_create_ex(rl_ctx* c, const uint8_t* occ, int width, int height, unsigned flags,
                     rl_map** out) {
  CHECK_ARG(c && occ && out, "null argument");
  *out = nullptr;
  if (width < 1 || height < 1 || width > 65535 || height > 65535)
    return fail(RL_ERR_ARG, "map size %dx%d outside 1..65535", width, height);
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(use_device(c));
  rl_map* m = new rl_map;
  m->ctx = c;
  c->refs.fetch_add(1);
  m->width = width;
  m->height = height;
  set_world(m, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f);  // RangeLibc.pyx:174-180
  const size_t N = (size_t)width * height;
  if (hipMalloc(&m->d_occ, N) != hipSuccess || hipMalloc(&m->d_edt, N * sizeof(float)) != hipSuccess) {
    (void)hipGetLastError();
    rl_map_destroy(m);
    return fail(RL_ERR_NOMEM, "device allocation for a %dx%d map failed", width, height);
  }
  if (hipMemcpyAsync(m->d_occ, occ, N, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
    rl_map_destroy(m);
    return fail(RL_ERR_HIP, "occupancy upload failed");
  }
  int rc = build_edt(m, flags);
  if (rc == RL_OK) rc = build_bits(m);
  if (rc != RL_OK) {
    std::string msg = g_err;
    rl_map_destroy(m);
    return fail(rc, "%s", msg.c_str());
  }
  *out = m;
  return RL_OK;
}

int rl_map_create(rl_ctx* c, const uint8_t* occ, int width, int height, rl_map** out) {
  return rl_map_create_ex(c, occ, width, height, 0u, out);
}

int rl_map_build_giant_lut(rl_map* m, float max_range, int theta_disc) {
  CHECK_ARG(m, "map is NULL");
  CHECK_ARG(theta_disc >= 1 && theta_disc <= 65536, "theta_discretization outside 1..65536");
  CHECK_ARG(max_range > 0.0f && max_range < 1e9f, "max_range outside (0, 1e9)");
  rl_ctx* c = m->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(use_device(c));
  if (!m->d_edt) return fail(RL_ERR_STATE, "map has no distance transform");
  const long long total = (long long)m->width * m->height * theta_disc;
  m->gen.fetch_add(1);
  if (m->d_glt) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipFree(m->d_glt));
    m->d_glt = nullptr;
  }
  if (hipMalloc(&m->d_glt, (size_t)total * sizeof(uint16_t)) != hipSuccess) {
    (void)hipGetLastError();
    m->d_glt = nullptr;
    return fail(RL_ERR_NOMEM, "device allocation of the %lld-entry GiantLUTCast table failed", total);
  }
  // RangeLib.h:1783-1786 (int / double, double / float, float / int, int / float)
  const double two_pi = 6.28318530718;
  m->glt_td = theta_disc;
  m->glt_R = max_range;
  m->glt_td_div_2pi = (float)((double)theta_disc / two_pi);
  m->glt_scale = max_range / (float)65535;
  const float step = (float)(two_pi / (double)(float)theta_disc);
  const float limits_div_max = (float)65535 / max_range;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  long long blocks = (total + 255) / 256;
  if (blocks > (1LL << 20)) blocks = 1LL << 20;
  float2* dirs = nullptr;
  hipError_t e = hipMalloc(&dirs, (size_t)theta_disc * sizeof(float2));
  if (e == hipSuccess) e = hipEventRecord(e0, c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_glt_dirs, dim3((unsigned)((theta_disc + 255) / 256)), dim3(256), 0, c->stream, c->trig,
                       theta_disc, step, dirs);
    const long long ncells = (long long)m->width * m->height;
    if (theta_disc <= RL_GLT_TILE_TD) {
      long long tiles = (ncells + RL_GLT_TILE - 1) / RL_GLT_TILE;
      if (tiles > (1LL << 20)) tiles = 1LL << 20;
      hipLaunchKernelGGL(k_glt_build_tiles, dim3((unsigned)tiles), dim3(256),
                         (unsigned)(RL_GLT_TILE * theta_disc * sizeof(uint16_t)), c->stream, m->dev(), max_range,
                         theta_disc, (const float2*)dirs, limits_div_max, m->d_glt, ncells);
    } else if (total <= 0xFFFFFFFFLL)
      hipLaunchKernelGGL(k_glt_build<false>, dim3((unsigned)blocks), dim3(256), 0, c->stream, m->dev(), max_range,
                         theta_disc, (const float2*)dirs, limits_div_max, m->d_glt, total);
    else
      hipLaunchKernelGGL(k_glt_build<true>, dim3((unsigned)blocks), dim3(256), 0, c->stream, m->dev(), max_range,
                         theta_disc, (const float2*)dirs, limits_div_max, m->d_glt, total);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) (void)hipEventElapsedTime(&m->glt_ms, e0, e1);
  if (dirs) (void)hipFree(dirs);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) {
    (void)hipFree(m->d_glt);
    m->d_glt = nullptr;
    return fail(RL_ERR_HIP, "GiantLUTCast build: %s", hipGetErrorString(e));
  }
  return RL_OK;
}

int rl_map_get_giant_lut(rl_map* m, uint16_t* out, int* theta_disc, float* max_range, float* build_ms) {
  CHECK_ARG(m, "map is NULL");
  rl_ctx* c = m->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(use_device(c));
  if (!m->d_glt) return fail(RL_ERR_STATE, "no GiantLUTCast table (call rl_map_build_giant_lut first)");
  if (theta_disc) *theta_disc = m->glt_td;
  if (max_range) *max_range = m->glt_R;
  if (build_ms) *build_ms = m->glt_ms;
  if (out) {
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipMemcpy(out, m->d_glt, (size_t)m->width * m->height * m->glt_td * sizeof(uint16_t),
                      hipMemcpyDeviceToHost));
  }
  return RL_OK;
}

int rl_map_create_from_edt(rl_ctx* c, const uint8_t* occ, const float* edt, int width, int height,
                           unsigned flags, rl_map** out) {
  CHECK_ARG(c && occ && edt && out, "null argument");
  *out = nullptr;
  if (width < 1 || height < 1 || width > 65535 || height > 65535)
    return fail(RL_ERR_ARG, "map size %dx%d outside 1..65535", width, height);
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(use_device(c));
  rl_map* m = new rl_map;
  m->ctx = c;
  c->refs.fetch_add(1);
  m->width = width;
  m->height = height;
  set_world(m, 1.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f);
  const size_t N = (size_t)width * height;
  if (hipMalloc(&m->d_occ, N) != hipSuccess || hipMalloc(&m->d_edt, N * sizeof(float)) != hipSuccess) {
    (void)hipGetLastError();
    rl_map_destroy(m);
    return fail(RL_ERR_NOMEM, "device allocation for a %dx%d map failed", width, height);
  }
  if (hipMemcpyAsync(m->d_occ, occ, N, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(m->d_edt, edt, N * sizeof(float), hipMemcpyHostToDevice, c->stream) != hipSuccess) {
    (void)hipGetLastError();
    rl_map_destroy(m);
    return fail(RL_ERR_HIP, "map upload failed");
  }
  int rc = build_compact(m, flags);
  if (rc == RL_OK) rc = build_bits(m);
  if (rc == RL_OK && hipStreamSynchronize(c->stream) != hipSuccess) rc = fail(RL_ERR_HIP, "map upload failed");
  if (rc != RL_OK) {
    std::string msg = g_err;
    rl_map_destroy(m);
    return fail(rc, "%s", msg.c_str());
  }
  *out = m;
  return RL_OK;
}

int rl_map_destroy(rl_map* m) {
  if (!m) return RL_OK;
  if (m->refs.fetch_sub(1) != 1) return RL_OK;  // a surrogate still holds it
  rl_ctx* c = m->ctx;
  (void)hipSetDevice(c->device);
  (void)hipStreamSynchronize(c->stream);
  if (m->d_occ) (void)hipFree(m->d_occ);
  if (m->d_edt) (void)hipFree(m->d_edt);
  if (m->d_idx8) (void)hipFree(m->d_idx8);
  if (m->d_bits) (void)hipFree(m->d_bits);
  if (m->d_sched) (void)hipFree(m->d_sched);
  if (m->d_table) (void)hipFree(m->d_table);
  if (m->d_glt) (void)hipFree(m->d_glt);
  {
    void* ps[] = {m->cddt_width, m->cddt_trans, m->cddt_cos, m->cddt_sin, m->cddt_vals, m->cddt_bin_off,
                  m->cddt_start};
    for (void* p : ps)
      if (p) (void)hipFree(p);
  }
  delete m;
  return rl_destroy(c);  // drops the map's reference
}

int rl_map_set_world(rl_map* m, float scale, float angle, float ox, float oy, float sin_a,
                     float cos_a) {
  CHECK_ARG(m, "map is NULL");
  std::lock_guard<std::mutex> lk(m->ctx->mu);
  set_world(m, scale, angle, ox, oy, sin_a, cos_a);
  m->gen.fetch_add(1);
  return RL_OK;
}

int rl_map_set_sensor_table(rl_map* m, const double* table, int k) {
  CHECK_ARG(m && table, "null argument");
  CHECK_ARG(k >= 1, "table width must be >= 1");
  std::lock_guard<std::mutex> lk(m->ctx->mu);
  TRY(use_device(m->ctx));
  m->gen.fetch_add(1);
  if (m->d_table) {
    HIP_TRY(hipStreamSynchronize(m->ctx->stream));
    HIP_TRY(hipFree(m->d_table));
    m->d_table = nullptr;
  }
  size_t bytes = (size_t)k * k * sizeof(double);
  if (hipMalloc(&m->d_table, bytes) != hipSuccess) {
    (void)hipGetLastError();
    m->d_table = nullptr;
    return fail(RL_ERR_NOMEM, "hipMalloc of a %dx%d sensor table failed", k, k);
  }
  HIP_TRY(hipMemcpyAsync(m->d_table, table, bytes, hipMemcpyHostToDevice, m->ctx->stream));
  HIP_TRY(hipStreamSynchronize(m->ctx->stream));
  m->table_width = k;
  return RL_OK;
}

int rl_map_get_edt(rl_map* m, float* out) {
  CHECK_ARG(m && out, "null argument");
  std::lock_guard<std::mutex> lk(m->ctx->mu);
  TRY(use_device(m->ctx));
  HIP_TRY(hipMemcpyAsync(out, m->d_edt, (size_t)m->width * m->height * sizeof(float),
                         hipMemcpyDeviceToHost, m->ctx->stream));
  HIP_TRY(hipStreamSynchronize(m->ctx->stream));
  return RL_OK;
}

int rl_map_device_buffers(rl_map* m, const float** d_edt, const uint8_t** d_occ) {
  CHECK_ARG(m, "map is NULL");
  if (d_edt) *d_edt = m->d_edt;
  if (d_occ) *d_occ = m->d_occ;
  return RL_OK;
}

int rl_map_edt_build_ms(rl_map* m, float* ms) {
  CHECK_ARG(m && ms, "null argument");
  *ms = m->edt_ms;
  return RL_OK;
}

// --- host-pointer batches ----------------------------------------------------

static size_t pin_part(size_t bytes) { return (bytes + 255) & ~(size_t)255; }

static int host_cast(rl_map* m, int method, int frame, float max_range, const float* ins,
                     const float* angles, float* outs, int num_poses, int num_angles) {
  CHECK_ARG(m && ins && outs, "null argument");
  CHECK_ARG(num_poses >= 0, "negative batch size");
  if (frame == RL_FRAME_ANGLES) {
    CHECK_ARG(angles, "angles is NULL");
    CHECK_ARG(num_angles >= 1, "num_angles must be >= 1");
  }
  TRY(check_method(method));
  rl_ctx* c = m->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(use_device(c));
  long long n_rays = frame == RL_FRAME_ANGLES ? (long long)num_poses * num_angles : num_poses;
  if (n_rays > 0x7fffffffLL) return fail(RL_ERR_ARG, "batch of %lld rays exceeds 2^31-1", n_rays);
  if (n_rays == 0) return RL_OK;
  const size_t b_in = pin_part((size_t)num_poses * 3 * sizeof(float));
  const size_t b_ang = frame == RL_FRAME_ANGLES ? pin_part((size_t)num_angles * sizeof(float)) : 0;
  const size_t b_out = pin_part((size_t)n_rays * sizeof(float));
  if (b_in + b_ang + b_out <= RL_PINNED_BYTES) {
    char* p;
    TRY(c->pinned((void**)&p));
    memcpy(p, ins, (size_t)num_poses * 3 * sizeof(float));
    if (b_ang) memcpy(p + b_in, angles, (size_t)num_angles * sizeof(float));
    TRY(launch_cast(m, method, frame, false, max_range, (const float*)p, b_ang ? (const float*)(p + b_in) : nullptr,
                    (float*)(p + b_in + b_ang), nullptr, (int)n_rays, num_angles));
    HIP_TRY(hipStreamSynchronize(c->stream));
    memcpy(outs, p + b_in + b_ang, (size_t)n_rays * sizeof(float));
    return RL_OK;
  }
  void *d_ins, *d_ang = nullptr, *d_out;
  TRY(c->staging(0, (size_t)num_poses * 3 * sizeof(float), &d_ins));
  TRY(c->staging(1, (size_t)n_rays * sizeof(float), &d_out));
  HIP_TRY(hipMemcpyAsync(d_ins, ins, (size_t)num_poses * 3 * sizeof(float), hipMemcpyHostToDevice, c->stream));
  if (frame == RL_FRAME_ANGLES) {
    TRY(c->staging(2, (size_t)num_angles * sizeof(float), &d_ang));
    HIP_TRY(hipMemcpyAsync(d_ang, angles, (size_t)num_angles * sizeof(float), hipMemcpyHostToDevice, c->stream));
  }
  TRY(launch_cast(m, method, frame, false, max_range, (const float*)d_ins, (const float*)d_ang,
                  (float*)d_out, nullptr, (int)n_rays, num_angles));
  HIP_TRY(hipMemcpyAsync(outs, d_out, (size_t)n_rays * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RL_OK;
}

int rl_calc_range(rl_map* m, int method, float max_range, float x, float y, float heading,
                  float* out) {
  CHECK_ARG(out, "out is NULL");
  float in[3] = {x, y, heading};
  return host_cast(m, method, RL_FRAME_GRID, max_range, in, nullptr, out, 1, 0);
}

int rl_calc_range_many(rl_map* m, int method, float max_range, const float* ins, float* outs, int n) {
  return host_cast(m, method, RL_FRAME_GRID, max_range, ins, nullptr, outs, n, 0);
}

int rl_numpy_calc_range(rl_map* m, int method, float max_range, const float* ins, float* outs, int n) {
  return host_cast(m, method, RL_FRAME_WORLD, max_range, ins, nullptr, outs, n, 0);
}

int rl_numpy_calc_range_angles(rl_map* m, int method, float max_range, const float* ins,
                               const float* angles, float* outs, int np_, int na) {
  return host_cast(m, method, RL_FRAME_ANGLES, max_range, ins, angles, outs, np_, na);
}

int rl_eval_sensor_model(rl_map* m, const float* obs, const float* ranges, double* outs, int M,
                         int N) {
  CHECK_ARG(m && obs && ranges && outs, "null argument");
  CHECK_ARG(M >= 1 && N >= 0, "bad rays_per_particle / particles");
  CHECK_ARG((long long)M * N <= 0x7fffffffLL, "batch too large");
  rl_ctx* c = m->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(use_device(c));
  if (N == 0) return RL_OK;
  const size_t b_obs = pin_part((size_t)M * sizeof(float)), b_rng = pin_part((size_t)M * N * sizeof(float));
  if (b_obs + b_rng + pin_part((size_t)N * sizeof(double)) <= RL_PINNED_BYTES) {
    char* p;
    TRY(c->pinned((void**)&p));
    memcpy(p, obs, (size_t)M * sizeof(float));
    memcpy(p + b_obs, ranges, (size_t)M * N * sizeof(float));
    TRY(launch_sensor(m, (const float*)p, (const float*)(p + b_obs), (double*)(p + b_obs + b_rng), M, N));
    HIP_TRY(hipStreamSynchronize(c->stream));
    memcpy(outs, p + b_obs + b_rng, (size_t)N * sizeof(double));
    return RL_OK;
  }
  void *d_obs, *d_rng, *d_w;
  TRY(c->staging(2, (size_t)M * sizeof(float), &d_obs));
  TRY(c->staging(3, (size_t)M * N * sizeof(float), &d_rng));
  TRY(c->staging(4, (size_t)N * sizeof(double), &d_w));
  HIP_TRY(hipMemcpyAsync(d_obs, obs, (size_t)M * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(d_rng, ranges, (size_t)M * N * sizeof(float), hipMemcpyHostToDevice, c->stream));
  TRY(launch_sensor(m, (const float*)d_obs, (const float*)d_rng, (double*)d_w, M, N));
  HIP_TRY(hipMemcpyAsync(outs, d_w, (size_t)N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RL_OK;
}

int rl_calc_range_repeat_angles_eval_sensor_model(rl_map* m, int method, float max_range,
                                                  const float* ins, const float* angles,
                                                  const float* obs, double* weights, int N, int M) {
  CHECK_ARG(m && ins && angles && obs && weights, "null argument");
  CHECK_ARG(M >= 1 && N >= 0, "bad num_angles / num_particles");
  rl_ctx* c = m->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(use_device(c));
  if (N == 0) return RL_OK;
  const size_t b_in = pin_part((size_t)N * 3 * sizeof(float)), b_ang = pin_part((size_t)M * sizeof(float));
  if (b_in + 2 * b_ang + pin_part((size_t)N * sizeof(double)) <= RL_PINNED_BYTES) {
    char* p;
    TRY(c->pinned((void**)&p));
    memcpy(p, ins, (size_t)N * 3 * sizeof(float));
    memcpy(p + b_in, angles, (size_t)M * sizeof(float));
    memcpy(p + b_in + b_ang, obs, (size_t)M * sizeof(float));
    TRY(launch_fused(m, method, max_range, (const float*)p, (const float*)(p + b_in),
                     (const float*)(p + b_in + b_ang), (double*)(p + b_in + 2 * b_ang), N, M));
    HIP_TRY(hipStreamSynchronize(c->stream));
    memcpy(weights, p + b_in + 2 * b_ang, (size_t)N * sizeof(double));
    return RL_OK;
  }
  void *d_ins, *d_ang, *d_obs, *d_w;
  TRY(c->staging(0, (size_t)N * 3 * sizeof(float), &d_ins));
  TRY(c->staging(2, (size_t)M * sizeof(float), &d_ang));
  TRY(c->staging(3, (size_t)M * sizeof(float), &d_obs));
  TRY(c->staging(4, (size_t)N * sizeof(double), &d_w));
  HIP_TRY(hipMemcpyAsync(d_ins, ins, (size_t)N * 3 * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(d_ang, angles, (size_t)M * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(d_obs, obs, (size_t)M * sizeof(float), hipMemcpyHostToDevice, c->stream));
  TRY(launch_fused(m, method, max_range, (const float*)d_ins, (const float*)d_ang,
                   (const float*)d_obs, (double*)d_w, N, M));
  HIP_TRY(hipMemcpyAsync(weights, d_w, (size_t)N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RL_OK;
}

// The fused call above with its staging in the static `scratch` array instead
// of the context's grown buffers: no device allocation.  Layout in the array,
// each part 256-byte aligned: poses N x 3, angles M, observations M, weights N.
static size_t noalloc_part(size_t bytes) { return (bytes + 255) & ~(size_t)255; }

int rl_calc_range_repeat_angles_eval_sensor_model_noalloc(rl_map* m, int method, float max_range,
                                                          const float* ins, const float* angles,
                                                          const float* obs, double* weights, int N, int M) {
  CHECK_ARG(m && ins && angles && obs && weights, "null argument");
  CHECK_ARG(M >= 1 && N >= 0, "bad num_angles / num_particles");
  const size_t b_ins = noalloc_part((size_t)N * 3 * sizeof(float));
  const size_t b_ang = noalloc_part((size_t)M * sizeof(float));
  const size_t b_w = noalloc_part((size_t)N * sizeof(double));
  const size_t need = b_ins + 2 * b_ang + b_w;
  if (need > (size_t)RL_NOALLOC_SCRATCH_BYTES)
    return fail(RL_ERR_ARG, "%d particles x %d angles need %zu staging bytes, above the %u of the "
                            "no-allocation entry point", N, M, need, (unsigned)RL_NOALLOC_SCRATCH_BYTES);
  rl_ctx* c = m->ctx;
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(use_device(c));
  if (N == 0) return RL_OK;
  char* base = (char*)c->d_scratch;
  float* d_ins = (float*)base;
  float* d_ang = (float*)(base + b_ins);
  float* d_obs = (float*)(base + b_ins + b_ang);
  double* d_w = (double*)(base + b_ins + 2 * b_ang);
  HIP_TRY(hipMemcpyAsync(d_ins, ins, (size_t)N * 3 * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(d_ang, angles, (size_t)M * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIP_TRY(hipMemcpyAsync(d_obs, obs, (size_t)M * sizeof(float), hipMemcpyHostToDevice, c->stream));
  TRY(launch_fused(m, method, max_range, d_ins, d_ang, d_obs, d_w, N, M));
  HIP_TRY(hipMemcpyAsync(weights, d_w, (size_t)N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return RL_OK;
}

// --- device-pointer batches ---------------------------------------------------

static int dev_cast(rl_map* m, int method, int frame, float max_range, const float* ins,
                    const float* angles, float* outs, int* counts, int num_poses, int num_angles) {
  CHECK_ARG(m && ins && (outs || counts), "null argument");
  CHECK_ARG(num_poses >= 0, "negative batch size");
  if (frame == RL_FRAME_ANGLES) {
    CHECK_ARG(angles, "angles is NULL");
    CHECK_ARG(num_angles >= 1, "num_angles must be >= 1");
  }
  long long n_rays = frame == RL_FRAME_ANGLES ? (long long)num_poses * num_angles : num_poses;
  if (n_rays > 0x7fffffffLL) return fail(RL_ERR_ARG, "batch of %lld rays exceeds 2^31-1", n_rays);
  std::lock_guard<std::mutex> lk(m->ctx->mu);
  TRY(use_device(m->ctx));
  return launch_cast(m, method, frame, counts != nullptr, max_range, ins, angles, outs, counts,
                     (int)n_rays, num_angles);
}

int rl_calc_range_many_dev(rl_map* m, int method, float max_range, const float* ins, float* outs, int n) {
  return dev_cast(m, method, RL_FRAME_GRID, max_range, ins, nullptr, outs, nullptr, n, 0);
}

int rl_numpy_calc_range_dev(rl_map* m, int method, float max_range, const float* ins, float* outs, int n) {
  return dev_cast(m, method, RL_FRAME_WORLD, max_range, ins, nullptr, outs, nullptr, n, 0);
}

int rl_numpy_calc_range_angles_dev(rl_map* m, int method, float max_range, const float* ins,
                                   const float* angles, float* outs, int np_, int na) {
  return dev_cast(m, method, RL_FRAME_ANGLES, max_range, ins, angles, outs, nullptr, np_, na);
}

int rl_eval_sensor_model_dev(rl_map* m, const float* obs, const float* ranges, double* outs, int M,
                             int N) {
  CHECK_ARG(m && obs && ranges && outs, "null argument");
  CHECK_ARG(M >= 1 && N >= 0, "bad rays_per_particle / particles");
  CHECK_ARG((long long)M * N <= 0x7fffffffLL, "batch too large");
  std::lock_guard<std::mutex> lk(m->ctx->mu);
  TRY(use_device(m->ctx));
  return launch_sensor(m, obs, ranges, outs, M, N);
}

int rl_calc_range_repeat_angles_eval_sensor_model_dev(rl_map* m, int method, float max_range,
                                                      const float* ins, const float* angles,
                                                      const float* obs, double* weights, int N,
                                                      int M) {
  CHECK_ARG(m && ins && angles && obs && weights, "null argument");
  CHECK_ARG(M >= 1 && N >= 0, "bad num_angles / num_particles");
  std::lock_guard<std::mutex> lk(m->ctx->mu);
  TRY(use_device(m->ctx));
  return launch_fused(m, method, max_range, ins, angles, obs, weights, N, M);
}

int rl_pf_step_dev(rl_map* m, int method, float max_range, const float* ins, const float* angles,
                   const float* obs, float* ranges, double* weights, int N, int M) {
  CHECK_ARG(m && ins && angles && obs && ranges && weights, "null argument");
  CHECK_ARG(M >= 1 && N >= 0, "bad num_angles / num_particles");
  CHECK_ARG((long long)M * N <= 0x7fffffffLL, "batch too large");
  std::lock_guard<std::mutex> lk(m->ctx->mu);
  TRY(use_device(m->ctx));
  return launch_particles(m, method, true, max_range, ins, angles, obs, ranges, weights, N, M);
}

int rl_set_option(rl_ctx* c, int option, long long value) {
  CHECK_ARG(c, "ctx is NULL");
  std::lock_guard<std::mutex> lk(c->mu);
  if (option == RL_OPT_XCD_SCHED) {
    c->xcd_sched = value != 0;
  } else if (option == RL_OPT_LDS_BOX) {
    c->lds_box = value != 0;
  } else {
    return fail(RL_ERR_ARG, "unknown option %d", option);
  }
  return RL_OK;
}

int rl_count_steps_dev(rl_map* m, int method, int frame, float max_range, const float* ins,
                       const float* angles, int* counts, int np_, int na) {
  CHECK_ARG(counts, "counts is NULL");
  return dev_cast(m, method, frame, max_range, ins, angles, nullptr, counts, np_, na);
}

int rl_debug_sincosf_dev(rl_ctx* c, const float* x, float* s, float* co, int n) {
  CHECK_ARG(c && x && s && co && n >= 0, "bad argument");
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(use_device(c));
  if (n == 0) return RL_OK;
  hipLaunchKernelGGL(k_debug_sincos, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, c->trig, x, s, co, n);
  HIP_TRY(hipGetLastError());
  return RL_OK;
}

int rl_debug_sqrtf_dev(rl_ctx* c, const float* x, float* out, int n) {
  CHECK_ARG(c && x && out && n >= 0, "bad argument");
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(use_device(c));
  if (n == 0) return RL_OK;
  hipLaunchKernelGGL(k_debug_sqrt, dim3(grid_for(n, 256)), dim3(256), 0, c->stream, x, out, n);
  HIP_TRY(hipGetLastError());
  return RL_OK;
}

int rl_debug_chase_dev(rl_ctx* c, const unsigned* next, int steps, long long* out, double* tick_ns) {
  CHECK_ARG(c && next && out && steps >= 0, "bad argument");
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(use_device(c));
  if (tick_ns) {
    int khz = 0;
    HIP_TRY(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
    *tick_ns = khz > 0 ? 1e6 / (double)khz : 0.0;
  }
  hipLaunchKernelGGL(k_debug_chase, dim3(1), dim3(64), 0, c->stream, next, steps, out);
  HIP_TRY(hipGetLastError());
  return RL_OK;
}

}  // extern "C"

// Multi-map batches (uses the types above).
#include <algorithm>
#include "rl_batch_impl.h"

// Learned range surrogate (uses the types above).
#include <cmath>
#include <random>
#include "rl_surrogate_impl.h"

// CDDTCast / PCDDT tables built on the GPU (uses the types above).
#include "rl_cddt_impl.h"
