// Device-wide exclusive sum of 64-bit counts and per-segment ascending sort of
// float keys, used by the CDDT table build (bin offsets from bin counts, then
// each bin's entries in order).
//
// Included by rl_cddt_impl.h (uses rl_lib.hip's fail/TRY/HIP_TRY and DevBuf).
//
// Scan: tiles of 2048 counts (256 threads x 8 consecutive items); pass 1 sums
// each tile, pass 2 (one workgroup) turns the tile sums into tile offsets,
// pass 3 rescans each tile from its offset.  Wave scans use 64-lane shuffles.
//
// Sort: floats are sorted by their order-preserving unsigned keys (sign bit
// set: all bits inverted; clear: sign bit set), the order of a radix sort of
// float bit patterns (-0 before +0).  Segments of up to 256 keys are sorted by
// one wave each in LDS, up to 2048 by a 256-thread workgroup, up to 16384 by a
// 1024-thread workgroup (64 KiB of LDS), all with bitonic networks over the
// segment padded to a power of two with the largest key; longer segments
// (maps far above 4096 cells a side) are sorted in global memory, one
// bitonic stage per launch.

#define RL_SCAN_ITEMS 8
#define RL_SCAN_BLOCK 256
#define RL_SCAN_TILE (RL_SCAN_ITEMS * RL_SCAN_BLOCK)

__device__ __forceinline__ unsigned long long wave_incl_scan_u64(unsigned long long v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned long long o = __shfl_up(v, (unsigned)d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

// Exclusive scan of one value per thread over a RL_SCAN_BLOCK-thread block;
// *total receives the block's sum.  Every thread of the block must call it.
__device__ __forceinline__ unsigned long long block_excl_scan_u64(unsigned long long v, unsigned long long* wsum,
                                                                  unsigned long long* total) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long inc = wave_incl_scan_u64(v);
  if (lane == 63) wsum[w] = inc;
  __syncthreads();
  unsigned long long off = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < RL_SCAN_BLOCK / 64; ++i) {
    const unsigned long long s = wsum[i];
    off += (i < w) ? s : 0ull;
    tot += s;
  }
  __syncthreads();  // wsum is rewritten by the next call
  *total = tot;
  return off + inc - v;
}

__global__ void __launch_bounds__(RL_SCAN_BLOCK) k_scan_tiles(const unsigned long long* __restrict__ in, long long n,
                                                              unsigned long long* __restrict__ tile_sum) {
  __shared__ unsigned long long wsum[RL_SCAN_BLOCK / 64];
  const long long base = (long long)blockIdx.x * RL_SCAN_TILE + (long long)threadIdx.x * RL_SCAN_ITEMS;
  unsigned long long s = 0;
#pragma unroll
  for (int k = 0; k < RL_SCAN_ITEMS; ++k) s += (base + k < n) ? in[base + k] : 0ull;
  unsigned long long tot;
  (void)block_excl_scan_u64(s, wsum, &tot);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}

// One workgroup: tile_sum[0..nt) becomes its own exclusive prefix sum.
__global__ void __launch_bounds__(RL_SCAN_BLOCK) k_scan_spine(unsigned long long* __restrict__ tile_sum, long long nt) {
  __shared__ unsigned long long wsum[RL_SCAN_BLOCK / 64];
  unsigned long long carry = 0;
  for (long long b0 = 0; b0 < nt; b0 += RL_SCAN_TILE) {
    const long long base = b0 + (long long)threadIdx.x * RL_SCAN_ITEMS;
    unsigned long long v[RL_SCAN_ITEMS], s = 0;
#pragma unroll
    for (int k = 0; k < RL_SCAN_ITEMS; ++k) {
      v[k] = (base + k < nt) ? tile_sum[base + k] : 0ull;
      s += v[k];
    }
    unsigned long long tot;
    unsigned long long ex = block_excl_scan_u64(s, wsum, &tot) + carry;
#pragma unroll
    for (int k = 0; k < RL_SCAN_ITEMS; ++k) {
      if (base + k < nt) tile_sum[base + k] = ex;
      ex += v[k];
    }
    carry += tot;
  }
}

__global__ void __launch_bounds__(RL_SCAN_BLOCK) k_scan_out(const unsigned long long* __restrict__ in, long long n,
                                                            const unsigned long long* __restrict__ tile_off,
                                                            unsigned long long* __restrict__ out) {
  __shared__ unsigned long long wsum[RL_SCAN_BLOCK / 64];
  const long long base = (long long)blockIdx.x * RL_SCAN_TILE + (long long)threadIdx.x * RL_SCAN_ITEMS;
  unsigned long long v[RL_SCAN_ITEMS], s = 0;
#pragma unroll
  for (int k = 0; k < RL_SCAN_ITEMS; ++k) {
    v[k] = (base + k < n) ? in[base + k] : 0ull;
    s += v[k];
  }
  unsigned long long tot;
  unsigned long long ex = block_excl_scan_u64(s, wsum, &tot) + tile_off[blockIdx.x];
#pragma unroll
  for (int k = 0; k < RL_SCAN_ITEMS; ++k) {
    if (base + k < n) out[base + k] = ex;
    ex += v[k];
  }
}

// out[i] = in[0] + ... + in[i - 1] for i < n, on the stream (returns after it ran).
int exclusive_sum_u64(hipStream_t st, const unsigned long long* in, unsigned long long* out, long long n) {
  if (n <= 0) return RL_OK;
  const long long nt = (n + RL_SCAN_TILE - 1) / RL_SCAN_TILE;
  if (nt > 0x7fffffffLL) return fail(RL_ERR_ARG, "scan: %lld items is too many", n);
  DevBuf tiles;
  TRY(tiles.alloc((size_t)nt * sizeof(unsigned long long)));
  hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)nt), dim3(RL_SCAN_BLOCK), 0, st, in, n,
                     tiles.as<unsigned long long>());
  hipLaunchKernelGGL(k_scan_spine, dim3(1), dim3(RL_SCAN_BLOCK), 0, st, tiles.as<unsigned long long>(), nt);
  hipLaunchKernelGGL(k_scan_out, dim3((unsigned)nt), dim3(RL_SCAN_BLOCK), 0, st, in, n,
                     (const unsigned long long*)tiles.p, out);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  return RL_OK;
}

__device__ __forceinline__ unsigned sort_key(float f) {
  const unsigned u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float sort_key_value(unsigned k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

struct WaveSync {
  __device__ void operator()() const {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
};
struct BlockSync {
  __device__ void operator()() const { __syncthreads(); }
};

// Ascending bitonic sort of s[0..P), P a power of two >= 2, by threads
// tid = 0..nt-1; sync() orders consecutive stages.
template <class Sync>
__device__ __forceinline__ void bitonic_lds(unsigned* s, int P, int tid, int nt, Sync sync) {
  for (int k = 2; k <= P; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < (P >> 1); i += nt) {
        const int a = ((i & ~(j - 1)) << 1) | (i & (j - 1)), b = a + j;
        const unsigned x = s[a], y = s[b];
        if ((x > y) == ((a & k) == 0)) {
          s[a] = y;
          s[b] = x;
        }
      }
      sync();
    }
}

__device__ __forceinline__ int pow2_at_least(long long n) {
  int p = 2;
  while (p < n) p <<= 1;
  return p;
}

#define RL_SORT_WAVE_MAX 256
#define RL_SORT_MID_MAX 2048
#define RL_SORT_BIG_MAX 16384

// Segments of at most RL_SORT_WAVE_MAX keys, one wave each (4 per workgroup);
// longer segments are left to the other kernels.
__global__ void __launch_bounds__(256) k_seg_sort_wave(const float* __restrict__ in, float* __restrict__ out,
                                                       const long long* __restrict__ start, long long nseg) {
  __shared__ unsigned sh[4][RL_SORT_WAVE_MAX];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long seg = (long long)blockIdx.x * 4 + w;
  if (seg >= nseg) return;  // no workgroup barrier below: a finished wave may leave
  const long long b = start[seg], n = start[seg + 1] - b;
  if (n > RL_SORT_WAVE_MAX || n <= 0) return;
  if (n == 1) {
    if (lane == 0) out[b] = in[b];
    return;
  }
  const int P = pow2_at_least(n);
  unsigned* s = sh[w];
  for (int i = lane; i < P; i += 64) s[i] = (i < n) ? sort_key(in[b + i]) : 0xffffffffu;
  WaveSync()();
  bitonic_lds(s, P, lane, 64, WaveSync());
  for (int i = lane; i < n; i += 64) out[b + i] = sort_key_value(s[i]);
}

// One workgroup of NT threads per listed segment (at most MAXN keys).
template <int NT, int MAXN>
__global__ void __launch_bounds__(NT) k_seg_sort_block(const float* __restrict__ in, float* __restrict__ out,
                                                       const long long* __restrict__ start,
                                                       const unsigned* __restrict__ list) {
  __shared__ unsigned s[MAXN];
  const long long seg = list[blockIdx.x];
  const long long b = start[seg], n = start[seg + 1] - b;
  const int P = pow2_at_least(n);
  for (int i = threadIdx.x; i < P; i += NT) s[i] = (i < n) ? sort_key(in[b + i]) : 0xffffffffu;
  __syncthreads();
  bitonic_lds(s, P, threadIdx.x, NT, BlockSync());
  for (int i = threadIdx.x; i < n; i += NT) out[b + i] = sort_key_value(s[i]);
}

// Lists the segments longer than RL_SORT_WAVE_MAX by size class: class 0
// (<= RL_SORT_MID_MAX), 1 (<= RL_SORT_BIG_MAX), 2 (longer).  One atomic per
// wave and class.
__global__ void __launch_bounds__(256) k_seg_classify(const long long* __restrict__ start, long long nseg,
                                                      unsigned* __restrict__ lists, unsigned* __restrict__ counts) {
  const long long seg = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long n = 0;
  if (seg < nseg) n = start[seg + 1] - start[seg];
  const int cls = n <= RL_SORT_WAVE_MAX ? -1 : (n <= RL_SORT_MID_MAX ? 0 : (n <= RL_SORT_BIG_MAX ? 1 : 2));
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const unsigned long long m = __ballot(cls == c);
    if (!m) continue;
    const int leader = __builtin_ctzll(m);
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(&counts[c], (unsigned)__popcll(m));
    base = __shfl(base, leader, 64);
    if (cls == c) lists[(size_t)c * (size_t)nseg + base + (unsigned)__popcll(m & ((1ull << lane) - 1ull))] = (unsigned)seg;
  }
}

__global__ void k_sort_global_load(const float* __restrict__ in, unsigned* __restrict__ keys, long long n, long long P) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < P) keys[i] = (i < n) ? sort_key(in[i]) : 0xffffffffu;
}
__global__ void k_sort_global_step(unsigned* __restrict__ keys, long long P, long long j, long long k) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (P >> 1)) return;
  const long long a = ((i & ~(j - 1)) << 1) | (i & (j - 1)), b = a + j;
  const unsigned x = keys[a], y = keys[b];
  if ((x > y) == ((a & k) == 0)) {
    keys[a] = y;
    keys[b] = x;
  }
}
__global__ void k_sort_global_store(const unsigned* __restrict__ keys, float* __restrict__ out, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = sort_key_value(keys[i]);
}

// out[start[s] .. start[s + 1]) = in[start[s] .. start[s + 1]) sorted ascending,
// for every segment s < nseg (start on the device, nseg + 1 entries, nseg < 2^32).
int segmented_sort_f32(hipStream_t st, const float* in, float* out, const long long* start, long long nseg) {
  if (nseg <= 0) return RL_OK;
  if (nseg >= 0xffffffffLL) return fail(RL_ERR_ARG, "sort: %lld segments is too many", nseg);
  hipLaunchKernelGGL(k_seg_sort_wave, dim3((unsigned)((nseg + 3) / 4)), dim3(256), 0, st, in, out, start, nseg);
  DevBuf lists, counts;
  TRY(lists.alloc((size_t)nseg * 3 * sizeof(unsigned)));
  TRY(counts.alloc(4 * sizeof(unsigned)));
  HIP_TRY(hipMemsetAsync(counts.p, 0, 4 * sizeof(unsigned), st));
  hipLaunchKernelGGL(k_seg_classify, dim3((unsigned)((nseg + 255) / 256)), dim3(256), 0, st, start, nseg,
                     lists.as<unsigned>(), counts.as<unsigned>());
  HIP_TRY(hipGetLastError());
  unsigned cnt[4] = {0, 0, 0, 0};
  HIP_TRY(hipMemcpyAsync(cnt, counts.p, sizeof(cnt), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  const unsigned* L = lists.as<unsigned>();
  if (cnt[0])
    hipLaunchKernelGGL((k_seg_sort_block<256, RL_SORT_MID_MAX>), dim3(cnt[0]), dim3(256), 0, st, in, out, start, L);
  if (cnt[1])
    hipLaunchKernelGGL((k_seg_sort_block<1024, RL_SORT_BIG_MAX>), dim3(cnt[1]), dim3(1024), 0, st, in, out, start,
                       L + nseg);
  HIP_TRY(hipGetLastError());
  if (cnt[2]) {
    std::vector<unsigned> big(cnt[2]);
    std::vector<long long> se(2);
    HIP_TRY(hipMemcpyAsync(big.data(), L + 2 * nseg, cnt[2] * sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    for (unsigned s : big) {
      HIP_TRY(hipMemcpy(se.data(), start + s, 2 * sizeof(long long), hipMemcpyDeviceToHost));
      const long long b = se[0], n = se[1] - se[0];
      long long P = 2;
      while (P < n) P <<= 1;
      DevBuf keys;
      TRY(keys.alloc((size_t)P * sizeof(unsigned)));
      const unsigned gp = (unsigned)((P + 255) / 256), gh = (unsigned)((P / 2 + 255) / 256);
      hipLaunchKernelGGL(k_sort_global_load, dim3(gp), dim3(256), 0, st, in + b, keys.as<unsigned>(), n, P);
      for (long long k = 2; k <= P; k <<= 1)
        for (long long j = k >> 1; j > 0; j >>= 1)
          hipLaunchKernelGGL(k_sort_global_step, dim3(gh), dim3(256), 0, st, keys.as<unsigned>(), P, j, k);
      hipLaunchKernelGGL(k_sort_global_store, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                         (const unsigned*)keys.p, out + b, n);
      HIP_TRY(hipGetLastError());
      HIP_TRY(hipStreamSynchronize(st));
    }
  }
  HIP_TRY(hipStreamSynchronize(st));
  return RL_OK;
}
