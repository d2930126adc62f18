// Learned range surrogate: a per-map Fourier-feature MLP that predicts a ray's
// range (and a bound on its own error), trained on the GPU on rays cast by the
// exact marcher, and used as a checked start point for exact marching.
//
// Included at the end of rl_lib.hip.  No counterpart in the reference.
//
// Network (per map):
//   v   = (x0 / W, y0 / H, sin h, cos h)       ray origin (grid frame) and heading
//   phi = [sin(2 pi B v), cos(2 pi B v)]       B: 32 x 4 fixed Gaussian frequencies -> 64 features
//   a1 = relu(W1 phi + b1), a2 = relu(W2 a1 + b2), a3 = relu(W3 a2 + b3)   (256 wide)
//   o  = W4 a3 + b4,  range = R * o[0],  error bound = exp(o[1]) cells
// Weights are stored in fp32 (master copy, Adam) and bf16 (MFMA operands);
// products accumulate in fp32 (v_mfma_f32_16x16x32_bf16 for the 256-wide
// layers, v_mfma_f32_16x16x16_bf16 elsewhere).
//
// Loss per ray: Huber (delta 1 cell) on range - marched range, plus the
// pinball loss at quantile 0.95 between the bound and the observed absolute
// error (the error is treated as a constant for this term).
//
// Checked inference: a ray whose bound b is <= bound_max cells and whose
// range r lies in [0, max_range] is a candidate for finishing from the start
// point t0 = max(0, r - b - 1):
//  1. certificate that the march from the origin cannot hit before t0: the
//     distance map is read at points of [0, t0] spaced at most
//     RL_SG_CERT_STEP apart (independent reads); a point p of the ray lies in
//     a cell less than |p - a| + sqrt(2) from the cell of sample a, so every
//     cell the segment crosses is free when each pair of neighbouring samples
//     a, b has (d_a - sqrt2)+ + (d_b - sqrt2)+ >= |b - a| (+ a rounding
//     margin) and d > 0 at each sample;
//  2. RL_SG_PROBES marches from t0 + k / RL_SG_PROBES (k = 0, 1, ...), with
//     the reference's step rule, each until its first hit, stopped (and the
//     ray rejected) past r + b + 2 or max_range.  The exact march reaches the
//     wall with its own sample phase, which decides between neighbouring hit
//     cells on oblique walls; the ray is accepted only when every probe hits
//     the same cell, and its range is that cell's distance.
// Rejected candidates and the other rays are marched from the origin (exact).
// Accepted ranges can still differ from the exact march where its phase picks
// a cell that no probe picked; the error is measured against the marcher.
// Reads: every read of both steps and of the fallback march is counted, and
// the dependent rounds of the path (one for the certificate, the longest
// probe, the fallback march) separately.

#define RL_SG_F 64    // Fourier features
#define RL_SG_NF 32   // frequencies
#define RL_SG_H 256   // hidden width
#define RL_SG_OUT 2
#define RL_SG_MAX_SPLIT 64  // split-K slices of the weight-gradient GEMMs (>= column-sum slices / 256 * 256)
#define RL_SG_NPARAM (RL_SG_H * RL_SG_F + RL_SG_H + 2 * (RL_SG_H * RL_SG_H + RL_SG_H) + RL_SG_OUT * RL_SG_H + RL_SG_OUT)

typedef uint16_t bf16_t;

__host__ __device__ __forceinline__ bf16_t f2bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)0x7fc0;  // NaN
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (bf16_t)(u >> 16);
}
__host__ __device__ __forceinline__ float bf2f(bf16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// Parameter offsets in the flat fp32 / bf16 vectors.
struct SgLayout {
  static constexpr int W1 = 0;
  static constexpr int B1 = W1 + RL_SG_H * RL_SG_F;
  static constexpr int W2 = B1 + RL_SG_H;
  static constexpr int B2 = W2 + RL_SG_H * RL_SG_H;
  static constexpr int W3 = B2 + RL_SG_H;
  static constexpr int B3 = W3 + RL_SG_H * RL_SG_H;
  static constexpr int W4 = B3 + RL_SG_H;
  static constexpr int B4 = W4 + RL_SG_OUT * RL_SG_H;
  static constexpr int N = B4 + RL_SG_OUT;
};
static_assert(SgLayout::N == RL_SG_NPARAM, "surrogate parameter layout");

// ---------------------------------------------------------------------------
// GEMM: C(m, n) = sum_k A(m, k) * B(k, n), bf16 operands, fp32 accumulation.
// A(m, k) = A[m * sam + k * sak], B(k, n) = B[k * sbk + n * sbn] (any
// transposition).  Epilogue: + bias[n], relu, zero where mask(m, n) <= 0 (a
// bf16 matrix with the output's strides), then store fp32 and/or bf16.
// Tile 64 x 64 per 256-thread workgroup, each wave 32 x 32 (2 x 2 MFMA tiles
// of 16 x 16), k in steps of 32 through LDS.
struct GemmArgs {
  const bf16_t* A;
  long long sam, sak;
  const bf16_t* B;
  long long sbk, sbn;
  int M, N, K;
  const float* bias;
  int relu;
  const bf16_t* mask;
  float* C;       // optional fp32 output, C[m * scm + n * scn]
  bf16_t* Cb;     // optional bf16 output, same strides
  long long scm, scn;
  // split-K (gridDim.z > 1): slice z sums k in [z * kchunk, (z + 1) * kchunk)
  // and stores its fp32 partial at C + z * scz (no epilogue; see k_sg_reduce)
  int kchunk;
  long long scz;
};

typedef short sg_short4 __attribute__((ext_vector_type(4)));
typedef float sg_float4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_gemm_bf16(GemmArgs g) {
  __shared__ bf16_t As[64][32 + 4];
  __shared__ bf16_t Bs[32][64 + 4];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  const bool split = gridDim.z > 1;
  const int kbeg = split ? (int)blockIdx.z * g.kchunk : 0;
  const int kend = split ? min(g.K, kbeg + g.kchunk) : g.K;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  sg_float4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = sg_float4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kbeg; k0 < kend; k0 += 32) {
    // stage A (64 x 32) and B (32 x 64): 2048 elements each, 8 per thread.
    // (Taking neighbouring m per thread when A is m-contiguous, as in the
    // weight-gradient GEMMs, made training 6% slower: 0.967 vs 0.910 s; 4%
    // slower when limited to the 256-row ones.)
    for (int e = tid; e < 64 * 32; e += 256) {
      int r = e >> 5, c = e & 31;
      int m = m0 + r, k = k0 + c;
      As[r][c] = (m < g.M && k < kend) ? g.A[(long long)m * g.sam + (long long)k * g.sak] : (bf16_t)0;
      int kr = e >> 6, nc = e & 63;
      int kk = k0 + kr, n = n0 + nc;
      Bs[kr][nc] = (kk < kend && n < g.N) ? g.B[(long long)kk * g.sbk + (long long)n * g.sbn] : (bf16_t)0;
    }
    __syncthreads();
    for (int ks = 0; ks < 32; ks += 16) {
      sg_short4 a[2], b[2];
      for (int i = 0; i < 2; ++i) {
        int row = wm + i * 16 + (lane & 15);
        int kb = ks + 4 * (lane >> 4);
        for (int q = 0; q < 4; ++q) a[i][q] = (short)As[row][kb + q];
      }
      for (int j = 0; j < 2; ++j) {
        int col = wn + j * 16 + (lane & 15);
        int kb = ks + 4 * (lane >> 4);
        for (int q = 0; q < 4; ++q) b[j][q] = (short)Bs[kb + q][col];
      }
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  // D layout of a 16x16 tile: lane holds column (lane & 15), rows 4*(lane >> 4) + q
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j)
      for (int q = 0; q < 4; ++q) {
        int m = m0 + wm + i * 16 + 4 * (lane >> 4) + q;
        int n = n0 + wn + j * 16 + (lane & 15);
        if (m >= g.M || n >= g.N) continue;
        float v = acc[i][j][q];
        if (split) {
          g.C[(long long)blockIdx.z * g.scz + (long long)m * g.scm + (long long)n * g.scn] = v;
          continue;
        }
        if (g.bias) v += g.bias[n];
        if (g.relu && v < 0.0f) v = 0.0f;
        long long o = (long long)m * g.scm + (long long)n * g.scn;
        if (g.mask && !(bf2f(g.mask[o]) > 0.0f)) v = 0.0f;
        if (g.C) g.C[o] = v;
        if (g.Cb) g.Cb[o] = f2bf(v);
      }
}

// Forward-layer GEMM, both operands K-contiguous ("NT"): C(m, n) = sum_k
// A[m * K + k] * Bt[n * K + k], + bias[n], optional relu, stored fp32 and/or
// bf16 at C[m * N + n].  Requires K % 32 == 0, N % 128 == 0 and 16-byte
// aligned rows.  Tile 128 x (32 * WN) per 256-thread workgroup, 64 x (16 * WN)
// per wave (tiles of v_mfma_f32_16x16x32_bf16), K in steps of 32 through LDS
// rows padded to 40 elements (16-byte reads land on distinct banks).
// Operand layout of the 16x16x32 instruction: lane l supplies row (l & 15)
// of A and column (l & 15) of B, k = 8 * (l >> 4) .. + 7; it receives column
// (l & 15), rows 4 * (l >> 4) .. + 3 of D.
typedef __bf16 sg_bf16x8 __attribute__((ext_vector_type(8)));
#define RL_SG_NT_PAD 40

template <int WN>  // 16-wide MFMA tiles per wave along N: the workgroup tile is 128 x (32 * WN)
__global__ __launch_bounds__(256) void k_gemm_bf16_nt(const bf16_t* __restrict__ A, const bf16_t* __restrict__ Bt,
                                                      int M, int N, int K, const float* __restrict__ bias,
                                                      int relu, float* __restrict__ C, bf16_t* __restrict__ Cb) {
  constexpr int BN = 32 * WN;
  __shared__ __align__(16) bf16_t As[128 * RL_SG_NT_PAD];
  __shared__ __align__(16) bf16_t Bs[BN * RL_SG_NT_PAD];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int m0 = blockIdx.y * 128, n0 = blockIdx.x * BN;
  const int wm = (wave >> 1) * 64, wn = (wave & 1) * (16 * WN);
  sg_float4 acc[4][WN];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < WN; ++j) acc[i][j] = sg_float4{0.f, 0.f, 0.f, 0.f};
  const int kq = (lane >> 4) * 8, r16 = lane & 15;
  for (int k0 = 0; k0 < K; k0 += 32) {
    // 128 rows of A and BN rows of Bt, 4 vectors of 8 elements each
#pragma unroll
    for (int v = tid; v < 512; v += 256) {
      const int r = v >> 2, c = (v & 3) * 8;
      const int m = m0 + r;
      uint4 a = make_uint4(0u, 0u, 0u, 0u);
      if (m < M) a = *(const uint4*)(A + (size_t)m * K + k0 + c);
      *(uint4*)(As + r * RL_SG_NT_PAD + c) = a;
    }
#pragma unroll
    for (int v = tid; v < BN * 4; v += 256) {
      const int r = v >> 2, c = (v & 3) * 8;
      *(uint4*)(Bs + r * RL_SG_NT_PAD + c) = *(const uint4*)(Bt + (size_t)(n0 + r) * K + k0 + c);
    }
    __syncthreads();
    sg_bf16x8 a[4], b[WN];
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = *(const sg_bf16x8*)(As + (wm + i * 16 + r16) * RL_SG_NT_PAD + kq);
#pragma unroll
    for (int j = 0; j < WN; ++j) b[j] = *(const sg_bf16x8*)(Bs + (wn + j * 16 + r16) * RL_SG_NT_PAD + kq);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < WN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    __syncthreads();
  }
  // Fully unrolled so that every accumulator index is a constant (a loop the
  // compiler keeps would index the accumulators through scratch memory).
#pragma unroll
  for (int j = 0; j < WN; ++j) {
    const int n = n0 + wn + j * 16 + r16;
    const float bn = bias ? bias[n] : 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int m = m0 + wm + i * 16 + 4 * (lane >> 4) + q;
        float v = acc[i][j][q] + bn;
        v = (relu && v < 0.0f) ? 0.0f : v;
        const size_t o = (size_t)m * N + n;
        if (m < M) {
          if (C) C[o] = v;
          if (Cb) Cb[o] = f2bf(v);
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Elementwise kernels

// Fourier features of n rays: rays[i] = (x0, y0, sin h, cos h) in grid units.
__global__ void k_sg_features(const float4* __restrict__ rays, const float* __restrict__ freq, float inv_w,
                              float inv_h, bf16_t* __restrict__ out, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 r = rays[i];
  float v0 = r.x * inv_w, v1 = r.y * inv_h, v2 = r.z, v3 = r.w;
  for (int j = 0; j < RL_SG_NF; ++j) {
    float p = 6.283185307179586f * (freq[4 * j] * v0 + freq[4 * j + 1] * v1 + freq[4 * j + 2] * v2 +
                                    freq[4 * j + 3] * v3);
    out[(size_t)i * RL_SG_F + j] = f2bf(sinf(p));
    out[(size_t)i * RL_SG_F + RL_SG_NF + j] = f2bf(cosf(p));
  }
}

// ---------------------------------------------------------------------------
// Fused inference: features, the three hidden layers and the heads of 64
// rays per 256-thread workgroup, the activations kept in LDS (one 64 x 256
// bf16 tile, rewritten in place by each layer).  The outputs are bitwise those
// of k_sg_features followed by the per-layer GEMMs: the same feature
// expression, the same matrix instructions over k in the same ascending order
// from zero accumulators (16x16x32 for the hidden layers as in
// k_gemm_bf16_nt, 16x16x16 for the heads as in k_gemm_bf16), the same bias /
// relu / bf16 epilogue.  Each row's outputs depend on that row alone, so the
// rows past n (zero features) are computed and never stored.
#define RL_SG_TM 64
#define RL_SG_ACT_PAD (RL_SG_H + 8)  // row stride in elements: 16-byte reads of 16 rows spread over the banks

// f2bf through the hardware's round-to-nearest-even conversion, NaN mapped to
// 0x7fc0 as f2bf does: equal to f2bf for every float (all 2^32 bit patterns
// compared by rl_debug_bf16_check).
__device__ __forceinline__ bf16_t f2bf_hw(float v) {
  const bf16_t r = __builtin_bit_cast(bf16_t, (__bf16)v);
  return v != v ? (bf16_t)0x7fc0 : r;
}

__global__ void k_sg_bf16_check(unsigned long long* __restrict__ bad) {
  unsigned long long local = 0;
  const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
  for (unsigned long long u = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; u < (1ull << 32);
       u += stride) {
    const float f = __uint_as_float((unsigned)u);
    local += f2bf_hw(f) != f2bf(f) ? 1u : 0u;
  }
  if (local) atomicAdd(bad, local);
}

// One hidden layer: act (64 x K) -> act (64 x 256), wave w computing columns
// 64w .. 64w + 63.  W is the bf16 layer matrix [256][K] (K-contiguous), sb the
// layer's biases in LDS.  b0 holds the wave's fragments of the first k step,
// loaded before the layer starts; the layer loads those of the next layer (Wn,
// KN columns; KN = 0: none) before its own epilogue, so that no layer starts
// with an L2 round trip.
template <int K, int KN>
__device__ __forceinline__ void sg_hidden_layer(bf16_t* act, const bf16_t* __restrict__ W, const float* sb,
                                                const bf16_t* __restrict__ Wn, sg_bf16x8 (&b0)[4], int wave,
                                                int lane) {
  constexpr int RT = RL_SG_TM / 16, KS = K / 32;
  const int kq = (lane >> 4) * 8, r16 = lane & 15;
  const int nw = wave * 64;
  sg_float4 acc[RT][4];
#pragma unroll
  for (int i = 0; i < RT; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = sg_float4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int s = 0; s < KS; ++s) {
    sg_bf16x8 b[4];
#pragma unroll
    for (int j = 0; j < 4; ++j)
      b[j] = s == 0 ? b0[j] : *(const sg_bf16x8*)(W + (size_t)(nw + j * 16 + r16) * K + s * 32 + kq);
#pragma unroll
    for (int i = 0; i < RT; ++i) {
      const sg_bf16x8 a = *(const sg_bf16x8*)(act + (i * 16 + r16) * RL_SG_ACT_PAD + s * 32 + kq);
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b[j], acc[i][j], 0, 0, 0);
    }
  }
  if (KN) {
#pragma unroll
    for (int j = 0; j < 4; ++j) b0[j] = *(const sg_bf16x8*)(Wn + (size_t)(nw + j * 16 + r16) * KN + kq);
  }
  __syncthreads();  // every wave has read the layer's input before it is overwritten
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int n = nw + j * 16 + r16;
    const float bb = sb[n];
#pragma unroll
    for (int i = 0; i < RT; ++i)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        float v = acc[i][j][q] + bb;
        v = v < 0.0f ? 0.0f : v;
        act[(i * 16 + 4 * (lane >> 4) + q) * RL_SG_ACT_PAD + n] = f2bf_hw(v);
      }
  }
  __syncthreads();
}

// rays: n x (x0, y0, sin h, cos h); P / Pf: the bf16 / fp32 parameter vectors;
// out: n x 2 fp32.  Launched with ceil(n / 64) workgroups of 256 threads.
// Register use is held to 128 so that four waves share each SIMD (measured on
// 240,000 rays: 3 waves, 144 registers, 5% slower).
__global__ __launch_bounds__(256, 4) void k_sg_infer(const float4* __restrict__ rays, const float* __restrict__ freq,
                                                  float inv_w, float inv_h, const bf16_t* __restrict__ P,
                                                  const float* __restrict__ Pf, float* __restrict__ out, int n) {
  __shared__ __align__(16) bf16_t act[RL_SG_TM * RL_SG_ACT_PAD];
  __shared__ float sbias[3 * RL_SG_H + RL_SG_OUT];         // b1, b2, b3, b4
  __shared__ __align__(16) bf16_t sw4[RL_SG_OUT * RL_SG_H];  // W4
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int r0 = blockIdx.x * RL_SG_TM;
  sg_bf16x8 b0[4];
#pragma unroll
  for (int j = 0; j < 4; ++j)
    b0[j] = *(const sg_bf16x8*)(P + SgLayout::W1 + (size_t)(wave * 64 + j * 16 + (lane & 15)) * RL_SG_F +
                                (lane >> 4) * 8);
  sbias[tid] = Pf[SgLayout::B1 + tid];
  sbias[RL_SG_H + tid] = Pf[SgLayout::B2 + tid];
  sbias[2 * RL_SG_H + tid] = Pf[SgLayout::B3 + tid];
  if (tid < RL_SG_OUT) sbias[3 * RL_SG_H + tid] = Pf[SgLayout::B4 + tid];
  sw4[tid] = P[SgLayout::W4 + tid];
  sw4[RL_SG_H + tid] = P[SgLayout::W4 + RL_SG_H + tid];
  for (int e = tid; e < RL_SG_TM * RL_SG_NF; e += 256) {
    const int r = e / RL_SG_NF, j = e - r * RL_SG_NF;
    bf16_t s = 0, c = 0;
    if (r0 + r < n) {
      const float4 rr = rays[r0 + r];
      const float v0 = rr.x * inv_w, v1 = rr.y * inv_h, v2 = rr.z, v3 = rr.w;
      const float p = 6.283185307179586f * (freq[4 * j] * v0 + freq[4 * j + 1] * v1 + freq[4 * j + 2] * v2 +
                                            freq[4 * j + 3] * v3);
      s = f2bf(sinf(p));
      c = f2bf(cosf(p));
    }
    act[r * RL_SG_ACT_PAD + j] = s;
    act[r * RL_SG_ACT_PAD + RL_SG_NF + j] = c;
  }
  __syncthreads();
  sg_hidden_layer<RL_SG_F, RL_SG_H>(act, P + SgLayout::W1, sbias, P + SgLayout::W2, b0, wave, lane);
  sg_hidden_layer<RL_SG_H, RL_SG_H>(act, P + SgLayout::W2, sbias + RL_SG_H, P + SgLayout::W3, b0, wave, lane);
  sg_hidden_layer<RL_SG_H, 0>(act, P + SgLayout::W3, sbias + 2 * RL_SG_H, nullptr, b0, wave, lane);
  // heads: wave w takes rows 16w .. 16w + 15 (RL_SG_TM / 64 tiles of 16 rows
  // per wave); columns 0 and 1 of one 16-wide tile
  const int col = lane & 15, kb = 4 * (lane >> 4);
#pragma unroll
  for (int t = 0; t < RL_SG_TM / 64; ++t) {
    const int rt = (wave * (RL_SG_TM / 64) + t) * 16;
    sg_float4 acc = sg_float4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < RL_SG_H; k += 16) {
      const sg_short4 a = *(const sg_short4*)(act + (rt + col) * RL_SG_ACT_PAD + k + kb);
      const sg_short4 b = col < RL_SG_OUT ? *(const sg_short4*)(sw4 + col * RL_SG_H + k + kb) : sg_short4{0, 0, 0, 0};
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    if (col < RL_SG_OUT) {
      const float bb = sbias[3 * RL_SG_H + col];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int m = r0 + rt + kb + q;
        if (m < n) out[(size_t)m * RL_SG_OUT + col] = acc[q] + bb;
      }
    }
  }
}

// Output gradient and per-block loss partial sums.  out: n x 2 fp32.
__global__ void k_sg_loss(const float* __restrict__ out, const float* __restrict__ target, float max_range,
                          float inv_n, bf16_t* __restrict__ dout, float* __restrict__ grad_f32,
                          float* __restrict__ partial, int n) {
  __shared__ float red[256];
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.0f;
  if (i < n) {
    const float tau = 0.95f;
    float pred = max_range * out[2 * i];
    float e = pred - target[i];
    float ae = fabsf(e);
    float lr = ae <= 1.0f ? 0.5f * e * e : ae - 0.5f;
    float ge = ae <= 1.0f ? e : (e > 0.0f ? 1.0f : -1.0f);
    float lb0 = out[2 * i + 1];
    float lb = lb0 > 20.0f ? 20.0f : (lb0 < -20.0f ? -20.0f : lb0);
    float b = expf(lb);
    float lq = ae > b ? tau * (ae - b) : (1.0f - tau) * (b - ae);
    float gb = lb == lb0 ? (ae > b ? -tau : (1.0f - tau)) * b : 0.0f;
    l = lr + 0.1f * lq;
    float g0 = ge * max_range * inv_n, g1 = 0.1f * gb * inv_n;
    dout[2 * i] = f2bf(g0);
    dout[2 * i + 1] = f2bf(g1);
    if (grad_f32) {
      grad_f32[2 * i] = g0;
      grad_f32[2 * i + 1] = g1;
    }
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0] * inv_n;
}

// Column sums of D (rows x cols, bf16 row-major) over row chunk blockIdx.y of
// `chunk` rows: part[blockIdx.y * cols + c].  One thread per column.
__global__ void k_sg_colsum(const bf16_t* __restrict__ D, int rows, int cols, int chunk,
                            float* __restrict__ part) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  int r0 = blockIdx.y * chunk, r1 = min(rows, r0 + chunk);
  float s = 0.0f;
#pragma unroll 8
  for (int r = r0; r < r1; ++r) s += bf2f(D[(size_t)r * cols + c]);  // loads issued ahead, sums in order
  part[(size_t)blockIdx.y * cols + c] = s;
}

// dst[i] = sum over z of part[z * count + i], z in [0, slices), in order.
__global__ void k_sg_reduce(const float* __restrict__ part, int slices, int count, float* __restrict__ dst) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float s = 0.0f;
#pragma unroll 8
  for (int z = 0; z < slices; ++z) s += part[(size_t)z * count + i];  // loads issued ahead, sums in order
  dst[i] = s;
}

// Adam on the flat parameter vector; writes the bf16 copy used by the GEMMs.
__global__ void k_sg_adam(float* __restrict__ p, float* __restrict__ m1, float* __restrict__ m2,
                          const float* __restrict__ g, bf16_t* __restrict__ pb, int n, float lr, float b1,
                          float b2, float c1, float c2, float eps) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float a = b1 * m1[i] + (1.0f - b1) * gi;
  float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  float v = p[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
  p[i] = v;
  pb[i] = f2bf(v);
}

__global__ void k_sg_to_bf16(const float* __restrict__ p, bf16_t* __restrict__ pb, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) pb[i] = f2bf(p[i]);
}

// Uniform free-space rays for training: cell drawn by a counter-based hash,
// redrawn while occupied (at most 64 draws; the map must have free cells).
__device__ __forceinline__ uint32_t sg_hash(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352du;
  x ^= x >> 15;
  x *= 0x846ca68bu;
  x ^= x >> 16;
  return x;
}

__global__ void k_sg_sample(MapDev m, TrigDev trig, uint32_t seed, long long first, int n,
                            float* __restrict__ poses, float4* __restrict__ rays) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t base = (uint32_t)(first + i) * 4u + seed * 0x9e3779b9u;
  uint32_t ix = 0, iy = 0;
  float fx = 0.5f, fy = 0.5f;
  for (int a = 0; a < 64; ++a) {
    uint32_t h1 = sg_hash(base + 0x10001u * (uint32_t)a), h2 = sg_hash(h1 ^ 0x5bd1e995u);
    uint32_t h3 = sg_hash(h2 + 0x27d4eb2du);
    ix = h1 % (uint32_t)m.width;
    iy = h3 % (uint32_t)m.height;
    fx = (float)(h2 >> 8) * (1.0f / 16777216.0f);
    fy = (float)(sg_hash(h3) >> 8) * (1.0f / 16777216.0f);
    if (!m.occ[(size_t)ix * m.height + iy]) break;
  }
  // cell + offset may round up to the next cell edge in fp32: keep it inside
  float x = (float)ix + fx, y = (float)iy + fy;
  if (x >= (float)(ix + 1u)) x = __int_as_float(__float_as_int((float)(ix + 1u)) - 1);
  if (y >= (float)(iy + 1u)) y = __int_as_float(__float_as_int((float)(iy + 1u)) - 1);
  float h = (float)(sg_hash(base + 3u) >> 8) * (6.283185307179586f / 16777216.0f);
  float s, c;
  rl::sincos_libc(h, trig, &s, &c);
  poses[3 * (size_t)i] = x;
  poses[3 * (size_t)i + 1] = y;
  poses[3 * (size_t)i + 2] = h;
  rays[i] = make_float4(x, y, s, c);
}

// Rays of a particle batch (world frame poses + beam angles) as surrogate inputs.
__global__ void k_sg_rays_from_particles(MapDev m, TrigDev trig, const float* __restrict__ ins,
                                         const float* __restrict__ angles, int N, int M,
                                         float4* __restrict__ rays) {
  int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= N * M) return;
  int p = r / M, a = r - p * M;
  float gx, gy;
  rl::to_grid(m.world, ins[3 * (size_t)p], ins[3 * (size_t)p + 1], &gx, &gy);
  float theta = -ins[3 * (size_t)p + 2] + m.world.rotation_const - angles[a];
  float s, c;
  rl::sincos_libc(theta, trig, &s, &c);
  rays[r] = make_float4(gy, gx, s, c);  // march origin (gy, gx), RangeLib.h:475
}

// Checked march from the surrogate's start point (see file comment).
// out: n x 2 network outputs; stats[0..4] += fallbacks (bound too large),
// fallbacks (check failed), rays finished from the start point, map reads,
// dependent read rounds.
#define RL_SG_CERT_STEP 8.0f  // largest gap between certificate samples (cells)
#define RL_SG_PROBES 4
#define RL_SG_PROBE_READS 64  // reads per probe before the ray is rejected
__global__ __launch_bounds__(256) void k_sg_checked_march(MapDev m, float max_range,
                                                          const float4* __restrict__ rays,
                                                          const float* __restrict__ out, float bound_max,
                                                          float* __restrict__ ranges, float scale,
                                                          unsigned* __restrict__ stats, int n) {
  __shared__ float lut[RL_LUT_SIZE];
  __shared__ unsigned cnt[5];
  for (int i = threadIdx.x; i < RL_LUT_SIZE; i += blockDim.x) lut[i] = m.steps[i];
  if (threadIdx.x < 5) cnt[threadIdx.x] = 0u;
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const float4 r = rays[i];
    const float x0 = r.x, y0 = r.y, dx = r.w, dy = r.z;
    const float pred = max_range * out[2 * i];
    float lb = out[2 * i + 1];
    lb = lb > 20.0f ? 20.0f : (lb < -20.0f ? -20.0f : lb);
    const float bound = expf(lb);
    const unsigned W = (unsigned)m.width, H = (unsigned)m.height;
    float result = -1.0f;
    int reads = 0, rounds = 0;
    bool tried = false;
    if (bound <= bound_max && pred >= 0.0f && pred <= max_range && rl::fast_march_ok(x0, y0, max_range)) {
      tried = true;
      float t0 = pred - bound - 1.0f;
      if (t0 < 0.0f) t0 = 0.0f;
      const float t_end = pred + bound + 2.0f;
      // 1. certificate over [0, t0]: nc + 1 evenly spaced samples
      bool ok = true;
      const int nc = t0 > 0.0f ? (int)ceilf(t0 / RL_SG_CERT_STEP) : 0;
      const float gap = nc > 0 ? t0 / (float)nc : 0.0f;
      float dprev = 0.0f;
      for (int k0 = 0; k0 <= nc; k0 += 16) {  // 16 independent reads in flight per round
        float d[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const int k = k0 + j;
          const float tk = gap * (float)k;
          const float fx = x0 + dx * tk, fy = y0 + dy * tk;
          const int px = (int)fx, py = (int)fy;
          d[j] = 0.0f;  // off the map or in (-1, 0), where truncation is not floor: no certificate
          if (k <= nc && fx >= 0.0f && fy >= 0.0f && (unsigned)px < W && (unsigned)py < H)
            d[j] = rl::glb(m.edt)[(size_t)px * H + py];
        }
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const int k = k0 + j;
          const float cover = fmaxf(d[j] - 1.41422f, 0.0f) + fmaxf(dprev - 1.41422f, 0.0f);
          if (k <= nc) {
            if (!(d[j] > 0.0f) || (k > 0 && !(cover >= gap + 0.01f))) ok = false;
            dprev = d[j];
            ++reads;
          }
        }
        ++rounds;
      }
      // 2. probes from t0 at RL_SG_PROBES phases, marched side by side
      if (ok) {
        float tp[RL_SG_PROBES], rp[RL_SG_PROBES];
        int ep[RL_SG_PROBES];
#pragma unroll
        for (int k = 0; k < RL_SG_PROBES; ++k) {
          tp[k] = t0 + (float)k * (1.0f / (float)RL_SG_PROBES);
          ep[k] = 0;
          rp[k] = 0.0f;
        }
        for (int step = 0; step < RL_SG_PROBE_READS; ++step) {
          bool any = false;
#pragma unroll
          for (int k = 0; k < RL_SG_PROBES; ++k) {
            // a probe past its window (r + b + 2), max_range or the map ends
            // without a hit (1)
            const int px = (int)(x0 + dx * tp[k]), py = (int)(y0 + dy * tp[k]);
            const bool inb = (unsigned)px < W && (unsigned)py < H;
            if (ep[k] == 0 && (!(tp[k] < max_range) || tp[k] > t_end || !inb)) ep[k] = 1;
            if (ep[k] == 0) {
              const float d = rl::glb(m.edt)[(size_t)px * H + py];
              ++reads;
              any = true;
              if (d <= 0.0f) {
                const float xd = (float)px - x0, yd = (float)py - y0;
                rp[k] = __builtin_sqrtf(xd * xd + yd * yd);
                ep[k] = 2;
              } else {
                const float st = d * 0.999f;
                tp[k] += (st < 1.0f) ? 1.0f : st;
              }
            }
          }
          if (!any) break;
          ++rounds;
        }
        bool agree = true;
#pragma unroll
        for (int k = 0; k < RL_SG_PROBES; ++k)
          agree = agree && ep[k] == 2 && __float_as_uint(rp[k]) == __float_as_uint(rp[0]);
        if (agree) result = rp[0];
      }
    }
    if (result < 0.0f) {
      int c = 0;
      result = (m.idx8 && rl::fast_march_ok(x0, y0, max_range))
                   ? rl::rm_march_idx8<true>(m, lut, rl::Box{nullptr, 0, 0}, max_range, x0, y0, dx, dy, &c)
                   : rl::rm_march<true>(m.edt, m.width, m.height, max_range, x0, y0, dx, dy, &c);
      reads += c;
      rounds += c;
      atomicAdd(&cnt[tried ? 1 : 0], 1u);
    } else {
      atomicAdd(&cnt[2], 1u);
    }
    atomicAdd(&cnt[3], (unsigned)reads);
    atomicAdd(&cnt[4], (unsigned)rounds);
    ranges[i] = result * scale;
  }
  __syncthreads();
  if (threadIdx.x < 5 && cnt[threadIdx.x]) atomicAdd(&stats[threadIdx.x], cnt[threadIdx.x]);
}

// ---------------------------------------------------------------------------
// Host side

struct rl_surrogate {
  rl_map* map = nullptr;
  float* d_p = nullptr;    // fp32 parameters
  float* d_m1 = nullptr;   // Adam moments
  float* d_m2 = nullptr;
  float* d_g = nullptr;    // gradient
  bf16_t* d_pb = nullptr;  // bf16 parameters
  float* d_freq = nullptr; // RL_SG_NF x 4
  long long adam_t = 0;
  // activations for `cap` rays
  int cap = 0;
  bf16_t *d_x = nullptr, *d_a1 = nullptr, *d_a2 = nullptr, *d_a3 = nullptr, *d_do = nullptr;
  bf16_t *d_dz_a = nullptr, *d_dz_b = nullptr;
  float* d_out = nullptr;
  float* d_partial = nullptr;
  float4* d_rays = nullptr;
  float* d_poses = nullptr;
  float* d_target = nullptr;
  float* d_split = nullptr;  // split-K / column-sum partials
  unsigned* d_stats = nullptr;
};

namespace {

void sg_free(rl_surrogate* s) {
  void* ps[] = {s->d_p, s->d_m1, s->d_m2, s->d_g, s->d_pb, s->d_freq, s->d_x, s->d_a1, s->d_a2, s->d_a3,
                s->d_do, s->d_dz_a, s->d_dz_b, s->d_out, s->d_partial, s->d_rays, s->d_poses, s->d_target,
                s->d_split, s->d_stats};
  for (void* p : ps)
    if (p) (void)hipFree(p);
}

template <typename T>
int sg_alloc(T** p, size_t count) {
  if (hipMalloc((void**)p, count * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return fail(RL_ERR_NOMEM, "surrogate: device allocation of %zu bytes failed", count * sizeof(T));
  }
  return RL_OK;
}

// Activation buffers for n rays.
int sg_reserve(rl_surrogate* s, int n) {
  if (s->cap >= n) return RL_OK;
  void* old[] = {s->d_x, s->d_a1, s->d_a2, s->d_a3, s->d_do, s->d_dz_a, s->d_dz_b, s->d_out, s->d_partial,
                 s->d_rays, s->d_poses, s->d_target, s->d_split};
  HIP_TRY(hipStreamSynchronize(s->map->ctx->stream));
  for (void* p : old)
    if (p) (void)hipFree(p);
  s->d_x = s->d_a1 = s->d_a2 = s->d_a3 = s->d_do = s->d_dz_a = s->d_dz_b = nullptr;
  s->d_out = s->d_partial = s->d_poses = s->d_target = nullptr;
  s->d_rays = nullptr;
  s->d_split = nullptr;
  s->cap = 0;
  size_t c = (size_t)n;
  TRY(sg_alloc(&s->d_x, c * RL_SG_F));
  TRY(sg_alloc(&s->d_a1, c * RL_SG_H));
  TRY(sg_alloc(&s->d_a2, c * RL_SG_H));
  TRY(sg_alloc(&s->d_a3, c * RL_SG_H));
  TRY(sg_alloc(&s->d_dz_a, c * RL_SG_H));
  TRY(sg_alloc(&s->d_dz_b, c * RL_SG_H));
  TRY(sg_alloc(&s->d_do, c * RL_SG_OUT));
  TRY(sg_alloc(&s->d_out, c * RL_SG_OUT));
  TRY(sg_alloc(&s->d_partial, c / 256 + 2));
  TRY(sg_alloc(&s->d_rays, c));
  TRY(sg_alloc(&s->d_poses, c * 3));
  TRY(sg_alloc(&s->d_target, c));
  TRY(sg_alloc(&s->d_split, (size_t)RL_SG_MAX_SPLIT * RL_SG_H * RL_SG_H));
  s->cap = n;
  return RL_OK;
}

// The NT kernel applies when both operands are K-contiguous with 16-byte
// aligned rows, the output is dense row-major and nothing is masked.
bool nt_ok(const GemmArgs& g) {
  return g.sak == 1 && g.sbk == 1 && g.sam == g.K && g.sbn == g.K && g.K % 32 == 0 && g.N % 128 == 0 &&
         g.scn == 1 && g.scm == g.N && !g.mask && ((uintptr_t)g.A & 15) == 0 && ((uintptr_t)g.B & 15) == 0;
}

int sg_gemm(hipStream_t st, const GemmArgs& g) {
  if (nt_ok(g)) {  // 128 x 128 tiles: 128 x 256 (WN = 8) needs 280 registers, one wave per SIMD, 2x slower
    dim3 grid(g.N / 128, (g.M + 127) / 128);
    hipLaunchKernelGGL(k_gemm_bf16_nt<4>, grid, dim3(256), 0, st, g.A, g.B, g.M, g.N, g.K, g.bias, g.relu, g.C, g.Cb);
  } else {
    dim3 grid((g.N + 63) / 64, (g.M + 63) / 64);
    hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, st, g);
  }
  HIP_TRY(hipGetLastError());
  return RL_OK;
}

GemmArgs gemm_args() {
  GemmArgs g;
  memset(&g, 0, sizeof(g));
  return g;
}

// Forward pass of n rays already in d_rays: d_x -> d_a1..d_a3 -> d_out.
int sg_forward(rl_surrogate* s, int n) {
  rl_map* m = s->map;
  hipStream_t st = m->ctx->stream;
  hipLaunchKernelGGL(k_sg_features, dim3(grid_for(n, 256)), dim3(256), 0, st, (const float4*)s->d_rays,
                     (const float*)s->d_freq, 1.0f / (float)m->width, 1.0f / (float)m->height, s->d_x, n);
  HIP_TRY(hipGetLastError());
  const bf16_t* P = s->d_pb;
  const bf16_t* in[4] = {s->d_x, s->d_a1, s->d_a2, s->d_a3};
  bf16_t* outb[3] = {s->d_a1, s->d_a2, s->d_a3};
  const int w_off[4] = {SgLayout::W1, SgLayout::W2, SgLayout::W3, SgLayout::W4};
  const int b_off[4] = {SgLayout::B1, SgLayout::B2, SgLayout::B3, SgLayout::B4};
  const int fin[4] = {RL_SG_F, RL_SG_H, RL_SG_H, RL_SG_H};
  const int fout[4] = {RL_SG_H, RL_SG_H, RL_SG_H, RL_SG_OUT};
  for (int l = 0; l < 4; ++l) {
    GemmArgs g = gemm_args();
    g.A = in[l];
    g.sam = fin[l];
    g.sak = 1;
    g.B = P + w_off[l];  // B(k = i, n = o) = W[o][i]
    g.sbk = 1;
    g.sbn = fin[l];
    g.M = n;
    g.N = fout[l];
    g.K = fin[l];
    g.bias = s->d_p + b_off[l];
    g.relu = l < 3;
    g.scm = fout[l];
    g.scn = 1;
    if (l < 3)
      g.Cb = outb[l];
    else
      g.C = s->d_out;
    TRY(sg_gemm(st, g));
  }
  return RL_OK;
}

// Network outputs of n rays already in d_rays into d_out, one fused launch
// (no intermediate activations in HBM); bitwise equal to sg_forward's d_out.
int sg_infer(rl_surrogate* s, int n) {
  rl_map* m = s->map;
  hipLaunchKernelGGL(k_sg_infer, dim3((unsigned)((n + RL_SG_TM - 1) / RL_SG_TM)), dim3(256), 0, m->ctx->stream,
                     (const float4*)s->d_rays, (const float*)s->d_freq, 1.0f / (float)m->width,
                     1.0f / (float)m->height, (const bf16_t*)s->d_pb, (const float*)s->d_p, s->d_out, n);
  HIP_TRY(hipGetLastError());
  return RL_OK;
}

// Loss, gradient into d_g (fp32, layout of the parameters) for n rays with
// targets in d_target.  grad_out (optional) receives d loss / d out (n x 2).
int sg_backward(rl_surrogate* s, int n, float max_range, float* grad_out) {
  rl_map* m = s->map;
  hipStream_t st = m->ctx->stream;
  hipLaunchKernelGGL(k_sg_loss, dim3(grid_for(n, 256)), dim3(256), 0, st, (const float*)s->d_out,
                     (const float*)s->d_target, max_range, 1.0f / (float)n, s->d_do, grad_out, s->d_partial, n);
  HIP_TRY(hipGetLastError());
  const bf16_t* P = s->d_pb;
  // layer l: input act_in (n x fin), output grad dz (n x fout); W at w_off (fout x fin)
  const bf16_t* acts[4] = {s->d_x, s->d_a1, s->d_a2, s->d_a3};
  const int w_off[4] = {SgLayout::W1, SgLayout::W2, SgLayout::W3, SgLayout::W4};
  const int b_off[4] = {SgLayout::B1, SgLayout::B2, SgLayout::B3, SgLayout::B4};
  const int fin[4] = {RL_SG_F, RL_SG_H, RL_SG_H, RL_SG_H};
  const int fout[4] = {RL_SG_H, RL_SG_H, RL_SG_H, RL_SG_OUT};
  const bf16_t* dz = s->d_do;
  bf16_t* bufs[2] = {s->d_dz_a, s->d_dz_b};
  // split-K over the batch for the weight gradients: ~1024 rows per slice
  const int splits = std::max(1, std::min(RL_SG_MAX_SPLIT, (n + 1023) / 1024));
  const int kchunk = ((n + splits - 1) / splits + 31) / 32 * 32;
  const int nz = (n + kchunk - 1) / kchunk;
  // column sums over row chunks of >= 256 rows, at most 256 chunks
  const int cchunk = std::max(256, (n + 255) / 256);
  const int ncz = (n + cchunk - 1) / cchunk;
  for (int l = 3; l >= 0; --l) {
    // dW[o][i] = sum_b dz[b][o] act[b][i]
    GemmArgs g = gemm_args();
    g.A = dz;
    g.sam = 1;
    g.sak = fout[l];
    g.B = acts[l];
    g.sbk = fin[l];
    g.sbn = 1;
    g.M = fout[l];
    g.N = fin[l];
    g.K = n;
    g.scm = fin[l];
    g.scn = 1;
    if (nz > 1) {
      g.C = s->d_split;
      g.kchunk = kchunk;
      g.scz = (long long)fout[l] * fin[l];
      dim3 grid((g.N + 63) / 64, (g.M + 63) / 64, nz);
      hipLaunchKernelGGL(k_gemm_bf16, grid, dim3(256), 0, st, g);
      HIP_TRY(hipGetLastError());
      hipLaunchKernelGGL(k_sg_reduce, dim3(grid_for(fout[l] * fin[l], 256)), dim3(256), 0, st,
                         (const float*)s->d_split, nz, fout[l] * fin[l], s->d_g + w_off[l]);
    } else {
      g.C = s->d_g + w_off[l];
      TRY(sg_gemm(st, g));
    }
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_sg_colsum, dim3(grid_for(fout[l], 64), ncz), dim3(64), 0, st, dz, n, fout[l], cchunk,
                       s->d_split);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_sg_reduce, dim3(grid_for(fout[l], 256)), dim3(256), 0, st, (const float*)s->d_split,
                       ncz, fout[l], s->d_g + b_off[l]);
    HIP_TRY(hipGetLastError());
    if (l == 0) break;
    // dz_prev[b][i] = (act[b][i] > 0) * sum_o dz[b][o] W[o][i]
    bf16_t* nxt = bufs[l & 1];
    GemmArgs h = gemm_args();
    h.A = dz;
    h.sam = fout[l];
    h.sak = 1;
    h.B = P + w_off[l];
    h.sbk = fin[l];
    h.sbn = 1;
    h.M = n;
    h.N = fin[l];
    h.K = fout[l];
    h.mask = acts[l];
    h.Cb = nxt;
    h.scm = fin[l];
    h.scn = 1;
    TRY(sg_gemm(st, h));
    dz = nxt;
  }
  return RL_OK;
}

int sg_adam(rl_surrogate* s, float lr) {
  s->adam_t += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  float c1 = (float)(1.0 - pow(0.9, (double)s->adam_t));
  float c2 = (float)(1.0 - pow(0.999, (double)s->adam_t));
  hipLaunchKernelGGL(k_sg_adam, dim3(grid_for(RL_SG_NPARAM, 256)), dim3(256), 0, s->map->ctx->stream, s->d_p,
                     s->d_m1, s->d_m2, (const float*)s->d_g, s->d_pb, RL_SG_NPARAM, lr, b1, b2, c1, c2, 1e-8f);
  HIP_TRY(hipGetLastError());
  return RL_OK;
}

static int debug_gemm(rl_ctx* c, const uint16_t* A, const uint16_t* B, float* C, int M, int N, int K, bool b_nk);

}  // namespace

__global__ void k_sg_zero_u32(unsigned* __restrict__ p, int n) {
  const int i = threadIdx.x;
  if (i < n) p[i] = 0u;
}

extern "C" {

int rl_surrogate_create(rl_map* map, unsigned seed, rl_surrogate** out) {
  CHECK_ARG(map && out, "null argument");
  *out = nullptr;
  std::lock_guard<std::mutex> lk(map->ctx->mu);
  TRY(use_device(map->ctx));
  rl_surrogate* s = new rl_surrogate;
  s->map = map;
  map->refs.fetch_add(1);
  int rc = RL_OK;
  if ((rc = sg_alloc(&s->d_p, RL_SG_NPARAM)) || (rc = sg_alloc(&s->d_m1, RL_SG_NPARAM)) ||
      (rc = sg_alloc(&s->d_m2, RL_SG_NPARAM)) || (rc = sg_alloc(&s->d_g, RL_SG_NPARAM)) ||
      (rc = sg_alloc(&s->d_pb, RL_SG_NPARAM)) || (rc = sg_alloc(&s->d_freq, RL_SG_NF * 4)) ||
      (rc = sg_alloc(&s->d_stats, 5))) {
    sg_free(s);
    delete s;
    map->refs.fetch_sub(1);  // cannot reach zero: the caller holds the map
    return rc;
  }
  // He-normal weights, zero biases; the range head starts at a tenth of
  // max_range, the bound head at e^2 cells.  Frequencies: N(0, 2^2) for the
  // normalised position, N(0, 1) for the direction components (low
  // frequencies train fastest on both small rooms and large corridor maps).
  std::mt19937_64 rng(seed);
  std::vector<float> p(RL_SG_NPARAM, 0.0f), fr(RL_SG_NF * 4);
  auto fill = [&](int off, int fout, int fin) {
    std::normal_distribution<double> nd(0.0, sqrt(2.0 / fin));
    for (int i = 0; i < fout * fin; ++i) p[off + i] = (float)nd(rng);
  };
  fill(SgLayout::W1, RL_SG_H, RL_SG_F);
  fill(SgLayout::W2, RL_SG_H, RL_SG_H);
  fill(SgLayout::W3, RL_SG_H, RL_SG_H);
  {
    std::normal_distribution<double> nd(0.0, 0.01);
    for (int i = 0; i < RL_SG_OUT * RL_SG_H; ++i) p[SgLayout::W4 + i] = (float)nd(rng);
  }
  p[SgLayout::B4] = 0.1f;
  p[SgLayout::B4 + 1] = 2.0f;
  {
    std::normal_distribution<double> pos(0.0, 2.0), dir(0.0, 1.0);
    for (int j = 0; j < RL_SG_NF; ++j) {
      fr[4 * j] = (float)pos(rng);
      fr[4 * j + 1] = (float)pos(rng);
      fr[4 * j + 2] = (float)dir(rng);
      fr[4 * j + 3] = (float)dir(rng);
    }
  }
  hipStream_t st = map->ctx->stream;
  hipError_t e = hipMemcpyAsync(s->d_p, p.data(), p.size() * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemcpyAsync(s->d_freq, fr.data(), fr.size() * 4, hipMemcpyHostToDevice, st);
  if (e == hipSuccess) e = hipMemsetAsync(s->d_m1, 0, RL_SG_NPARAM * 4, st);
  if (e == hipSuccess) e = hipMemsetAsync(s->d_m2, 0, RL_SG_NPARAM * 4, st);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_sg_to_bf16, dim3(grid_for(RL_SG_NPARAM, 256)), dim3(256), 0, st, (const float*)s->d_p,
                       s->d_pb, RL_SG_NPARAM);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    sg_free(s);
    delete s;
    map->refs.fetch_sub(1);
    return fail(RL_ERR_HIP, "surrogate init: %s", hipGetErrorString(e));
  }
  *out = s;
  return RL_OK;
}

int rl_surrogate_destroy(rl_surrogate* s) {
  if (!s) return RL_OK;
  rl_map* m = s->map;
  (void)hipSetDevice(m->ctx->device);
  (void)hipStreamSynchronize(m->ctx->stream);
  sg_free(s);
  delete s;
  return rl_map_destroy(m);  // drops the surrogate's reference
}

int rl_surrogate_num_params(void) { return RL_SG_NPARAM; }

int rl_surrogate_get_params(rl_surrogate* s, float* params, float* freq) {
  CHECK_ARG(s, "null argument");
  std::lock_guard<std::mutex> lk(s->map->ctx->mu);
  TRY(use_device(s->map->ctx));
  hipStream_t st = s->map->ctx->stream;
  if (params) HIP_TRY(hipMemcpyAsync(params, s->d_p, RL_SG_NPARAM * 4, hipMemcpyDeviceToHost, st));
  if (freq) HIP_TRY(hipMemcpyAsync(freq, s->d_freq, RL_SG_NF * 16, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return RL_OK;
}

int rl_surrogate_set_params(rl_surrogate* s, const float* params, const float* freq) {
  CHECK_ARG(s, "null argument");
  std::lock_guard<std::mutex> lk(s->map->ctx->mu);
  TRY(use_device(s->map->ctx));
  hipStream_t st = s->map->ctx->stream;
  if (params) HIP_TRY(hipMemcpyAsync(s->d_p, params, RL_SG_NPARAM * 4, hipMemcpyHostToDevice, st));
  if (freq) HIP_TRY(hipMemcpyAsync(s->d_freq, freq, RL_SG_NF * 16, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_sg_to_bf16, dim3(grid_for(RL_SG_NPARAM, 256)), dim3(256), 0, st, (const float*)s->d_p,
                     s->d_pb, RL_SG_NPARAM);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(st));
  return RL_OK;
}

// Network outputs (n x 2) for n rays (x0, y0, sin h, cos h); host pointers.
static int surrogate_forward(rl_surrogate* s, const float* rays4, int n, float* out, bool layers) {
  CHECK_ARG(s && rays4 && out && n >= 0, "bad argument");
  std::lock_guard<std::mutex> lk(s->map->ctx->mu);
  TRY(use_device(s->map->ctx));
  if (n == 0) return RL_OK;
  TRY(sg_reserve(s, n));
  hipStream_t st = s->map->ctx->stream;
  HIP_TRY(hipMemcpyAsync(s->d_rays, rays4, (size_t)n * 16, hipMemcpyHostToDevice, st));
  TRY(layers ? sg_forward(s, n) : sg_infer(s, n));
  HIP_TRY(hipMemcpyAsync(out, s->d_out, (size_t)n * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  return RL_OK;
}

int rl_surrogate_forward(rl_surrogate* s, const float* rays4, int n, float* out) {
  return surrogate_forward(s, rays4, n, out, false);
}

// Loss and gradient for n rays with target ranges (host pointers); the bf16
// features the network saw are returned in `features` (n x 64, optional).
int rl_surrogate_gradient(rl_surrogate* s, float max_range, const float* rays4, const float* targets, int n,
                          float* loss, float* grad, uint16_t* features) {
  CHECK_ARG(s && rays4 && targets && n > 0, "bad argument");
  std::lock_guard<std::mutex> lk(s->map->ctx->mu);
  TRY(use_device(s->map->ctx));
  TRY(sg_reserve(s, n));
  hipStream_t st = s->map->ctx->stream;
  HIP_TRY(hipMemcpyAsync(s->d_rays, rays4, (size_t)n * 16, hipMemcpyHostToDevice, st));
  HIP_TRY(hipMemcpyAsync(s->d_target, targets, (size_t)n * 4, hipMemcpyHostToDevice, st));
  TRY(sg_forward(s, n));
  TRY(sg_backward(s, n, max_range, nullptr));
  if (grad) HIP_TRY(hipMemcpyAsync(grad, s->d_g, RL_SG_NPARAM * 4, hipMemcpyDeviceToHost, st));
  if (features) HIP_TRY(hipMemcpyAsync(features, s->d_x, (size_t)n * RL_SG_F * 2, hipMemcpyDeviceToHost, st));
  std::vector<float> part(grid_for(n, 256));
  HIP_TRY(hipMemcpyAsync(part.data(), s->d_partial, part.size() * 4, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (loss) {
    double t = 0;
    for (float v : part) t += v;
    *loss = (float)t;
  }
  return RL_OK;
}

// One Adam step (lr; beta 0.9 / 0.999, eps 1e-8) with the gradient of the last
// rl_surrogate_gradient call.
int rl_surrogate_step(rl_surrogate* s, float lr) {
  CHECK_ARG(s && lr > 0.0f, "bad argument");
  std::lock_guard<std::mutex> lk(s->map->ctx->mu);
  TRY(use_device(s->map->ctx));
  TRY(sg_adam(s, lr));
  HIP_TRY(hipStreamSynchronize(s->map->ctx->stream));
  return RL_OK;
}

// Trains on n_pairs uniform free-space rays cast by the exact marcher, in
// minibatches of `batch`; loss_curve (optional, n_pairs / batch entries)
// receives each step's loss.
int rl_surrogate_train(rl_surrogate* s, float max_range, long long n_pairs, int batch, float lr,
                       unsigned seed, float* loss_curve) {
  CHECK_ARG(s && n_pairs > 0 && batch >= 64 && batch <= (1 << 20) && lr > 0.0f, "bad argument");
  std::lock_guard<std::mutex> lk(s->map->ctx->mu);
  rl_ctx* c = s->map->ctx;
  TRY(use_device(c));
  if (!s->map->d_edt) return fail(RL_ERR_STATE, "map has no distance transform");
  TRY(sg_reserve(s, batch));
  hipStream_t st = c->stream;
  const long long steps = n_pairs / batch;
  std::vector<float> part(grid_for(batch, 256));
  for (long long k = 0; k < steps; ++k) {
    hipLaunchKernelGGL(k_sg_sample, dim3(grid_for(batch, 256)), dim3(256), 0, st, s->map->dev(), c->trig, seed,
                       k * (long long)batch, batch, s->d_poses, s->d_rays);
    HIP_TRY(hipGetLastError());
    TRY(launch_cast(s->map, RL_METHOD_RM, RL_FRAME_GRID, false, max_range, s->d_poses, nullptr, s->d_target,
                    nullptr, batch, 0));
    TRY(sg_forward(s, batch));
    TRY(sg_backward(s, batch, max_range, nullptr));
    TRY(sg_adam(s, lr));
    if (loss_curve) {
      HIP_TRY(hipMemcpyAsync(part.data(), s->d_partial, part.size() * 4, hipMemcpyDeviceToHost, st));
      HIP_TRY(hipStreamSynchronize(st));
      double t = 0;
      for (float v : part) t += v;
      loss_curve[k] = (float)t;
    }
  }
  HIP_TRY(hipStreamSynchronize(st));
  return RL_OK;
}

// Checked surrogate cast of a particle batch (numpy_calc_range_angles
// semantics for inputs and outputs; ranges approximate, see file comment).
// stats (host, 5 entries): fallbacks (bound), fallbacks (check), rays finished
// from the surrogate's start point, distance-map reads, dependent read rounds.
int rl_surrogate_calc_range_angles_dev(rl_surrogate* s, float max_range, const float* d_ins,
                                       const float* d_angles, float* d_ranges, int N, int M, float bound_max,
                                       unsigned* stats) {
  CHECK_ARG(s && d_ins && d_angles && d_ranges && N >= 0 && M >= 1, "bad argument");
  CHECK_ARG((long long)N * M <= 0x7fffffffLL, "batch too large");
  std::lock_guard<std::mutex> lk(s->map->ctx->mu);
  rl_ctx* c = s->map->ctx;
  TRY(use_device(c));
  if (!s->map->d_edt) return fail(RL_ERR_STATE, "map has no distance transform");
  const int n = N * M;
  if (n == 0) return RL_OK;
  TRY(sg_reserve(s, n));
  hipStream_t st = c->stream;
  MapDev md = s->map->dev();
  if (stats) {  // a fill kernel, not a memset node, so that the launch sequence stays graph-safe
    hipLaunchKernelGGL(k_sg_zero_u32, dim3(1), dim3(64), 0, st, s->d_stats, 5);
    HIP_TRY(hipGetLastError());
  }
  hipLaunchKernelGGL(k_sg_rays_from_particles, dim3(grid_for(n, 256)), dim3(256), 0, st, md, c->trig, d_ins,
                     d_angles, N, M, s->d_rays);
  HIP_TRY(hipGetLastError());
  TRY(sg_infer(s, n));
  hipLaunchKernelGGL(k_sg_checked_march, dim3(grid_for(n, 256)), dim3(256), 0, st, md, max_range,
                     (const float4*)s->d_rays, (const float*)s->d_out, bound_max, d_ranges, md.world.scale,
                     s->d_stats, n);
  HIP_TRY(hipGetLastError());
  if (stats) {
    HIP_TRY(hipMemcpyAsync(stats, s->d_stats, 5 * sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
  }
  return RL_OK;
}

// Plain bf16 GEMM through the same kernels (test hook): C = A(M x K) * B(K x N),
// row-major host matrices, fp32 result.
int rl_debug_gemm_bf16(rl_ctx* c, const uint16_t* A, const uint16_t* B, float* C, int M, int N, int K) {
  return debug_gemm(c, A, B, C, M, N, K, false);
}

// rl_surrogate_forward through the per-layer kernels the training uses (test
// hook for the fused inference kernel).
int rl_debug_surrogate_forward_layers(rl_surrogate* s, const float* rays4, int n, float* out) {
  return surrogate_forward(s, rays4, n, out, true);
}

// Mismatches between the fused kernel's bf16 conversion and f2bf over all
// 2^32 float bit patterns (test hook).
int rl_debug_bf16_check(rl_ctx* c, unsigned long long* mismatches) {
  CHECK_ARG(c && mismatches, "null argument");
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(use_device(c));
  unsigned long long* d = nullptr;
  HIP_TRY(hipMalloc(&d, sizeof(*d)));
  hipError_t e = hipMemsetAsync(d, 0, sizeof(*d), c->stream);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_sg_bf16_check, dim3(4096), dim3(256), 0, c->stream, d);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(mismatches, d, sizeof(*d), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  HIP_TRY(e);
  return RL_OK;
}

// C = A(M x K) * Bt(N x K)^T: the forward-layer layout (K-contiguous operands).
int rl_debug_gemm_bf16_nt(rl_ctx* c, const uint16_t* A, const uint16_t* Bt, float* C, int M, int N, int K) {
  return debug_gemm(c, A, Bt, C, M, N, K, true);
}

}  // extern "C"

namespace {

static int debug_gemm(rl_ctx* c, const uint16_t* A, const uint16_t* B, float* C, int M, int N, int K, bool b_nk) {
  CHECK_ARG(c && A && B && C && M > 0 && N > 0 && K > 0, "bad argument");
  std::lock_guard<std::mutex> lk(c->mu);
  TRY(use_device(c));
  bf16_t *dA = nullptr, *dB = nullptr;
  float* dC = nullptr;
  int rc = RL_OK;
  if ((rc = sg_alloc(&dA, (size_t)M * K)) || (rc = sg_alloc(&dB, (size_t)K * N)) ||
      (rc = sg_alloc(&dC, (size_t)M * N))) {
    if (dA) (void)hipFree(dA);
    if (dB) (void)hipFree(dB);
    return rc;
  }
  GemmArgs g = gemm_args();
  g.A = dA;
  g.sam = K;
  g.sak = 1;
  g.B = dB;
  g.sbk = b_nk ? 1 : N;
  g.sbn = b_nk ? K : 1;
  g.M = M;
  g.N = N;
  g.K = K;
  g.C = dC;
  g.scm = N;
  g.scn = 1;
  hipError_t e = hipMemcpy(dA, A, (size_t)M * K * 2, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dB, B, (size_t)K * N * 2, hipMemcpyHostToDevice);
  if (e == hipSuccess) rc = sg_gemm(c->stream, g);
  if (e == hipSuccess && rc == RL_OK) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess && rc == RL_OK) e = hipMemcpy(C, dC, (size_t)M * N * 4, hipMemcpyDeviceToHost);
  (void)hipFree(dA);
  (void)hipFree(dB);
  (void)hipFree(dC);
  if (rc != RL_OK) return rc;
  if (e != hipSuccess) return fail(RL_ERR_HIP, "debug gemm: %s", hipGetErrorString(e));
  return RL_OK;
}

}  // namespace
