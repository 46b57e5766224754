// Reduced-precision forms of the two large sketch contractions of a dense
// window (A_w is w x n, FP64, column-major in HBM):
//   left : x        = Upsilon[:, idx] * A_w          (k x n, reduce over w)
//   right: [y | G]  = A_w * [Omega; Psi]^T           (w x (k+s), reduce over n)
//
// Each window is read once in FP64 by a quantize pass that writes two compact
// copies with the reduction index contiguous: L (column-major, for the left
// product, K = window rows) and R (row-major, for the right product, K =
// columns).  Both products are then the same "NT" GEMM, C = A * B^T with A
// (M x K) and B (N x K) both K-contiguous, fed by 16-byte loads and K-major
// LDS tiles.  The maps are converted once at sketch creation.  Products
// accumulate in FP32 on the matrix cores; the epilogue widens to FP64 and
// applies the nu/eta update into the FP64 state.  The small core product
// Z += Phi[:, idx] * G and the whole finalize stay FP64.
//
// Forms:
//   fp32   : v_mfma_f32_16x16x4_f32 on FP32-rounded operands
//   bf16   : v_mfma_f32_16x16x32_bf16, one product
//   bf16x3 : same instruction, x = hi + lo (both bf16), a*b ~ hi*hi + hi*lo + lo*hi
//   fp8    : v_mfma_scale_f32_16x16x128_f8f6f4 with unit block scales on e4m3
//            operands.  Every row of each operand carries one power-of-two
//            scale over its whole reduction range (per row of a map and of
//            the window, per column of the window), chosen from that row's
//            max |x| so it lands in [128, 256) before rounding; the FP32
//            accumulator is multiplied by the two scales in the FP64 epilogue.
#include <cmath>

#include "lowp.hpp"
#include "util.hpp"

namespace skema {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef int i8v __attribute__((ext_vector_type(8)));
typedef int i4v __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
constexpr int kQ = 64;  // quantize tile

__device__ __forceinline__ int xcd_remap(int bid, int nb) {
  const int q = nb >> 3, r = nb & 7;
  const int xcd = bid & 7, local = bid >> 3;
  return xcd * q + (xcd < r ? xcd : r) + local;
}

__device__ __forceinline__ void store_out(const GemmOut& o, int64_t i, int64_t j, double v) {
  if (j < o.nsplit) {
    double* c = o.C1 + i + j * o.ldc1;
    *c = (o.beta1 != 0.0) ? o.alpha1 * v + o.beta1 * *c : o.alpha1 * v;
  } else {
    double* c = o.C2 + i + (j - o.nsplit) * o.ldc2;
    *c = (o.beta2 != 0.0) ? o.alpha2 * v + o.beta2 * *c : o.alpha2 * v;
  }
}

// ---- e4m3 with a per-row power-of-two scale ----------------------------------
// Scale exponent e with max|x| * 2^-e in [128, 256) (<= 448, the largest e4m3
// normal).  A zero row gets e = 0.
__device__ __forceinline__ int block_exponent(float amax) {
  if (!(amax > 0.f)) return 0;
  int ex;
  (void)frexpf(amax, &ex);  // amax = f * 2^ex, f in [0.5, 1)
  int e = ex - 8;
  e = e < -127 ? -127 : (e > 127 ? 127 : e);
  return e;
}
__device__ __forceinline__ uint32_t pack4_e4m3(float a, float b, float c, float d) {
  int lo = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
  return (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(c, d, lo, true);
}

// ---- window quantization ------------------------------------------------------
// 64 x 64 tile of A_w -> L (column-major, ld ldl) and R (row-major, ld ldr),
// elements of type T (bf16 or fp32); SPLIT adds the bf16 residual planes.
template <class T, bool SPLIT>
__global__ __launch_bounds__(kThreads) void k_quant(const double* A, int64_t lda, int64_t rows, int64_t n, T* Lh,
                                                    T* Ll, int64_t ldl, T* Rh, T* Rl, int64_t ldr) {
  __shared__ float t[kQ][kQ + 1];  // t[j][i]
  const int tid = threadIdx.x;
  const int64_t i0 = (int64_t)blockIdx.x * kQ, j0 = (int64_t)blockIdx.y * kQ;
  for (int e = 0; e < kQ * kQ / kThreads; ++e) {
    const int i = tid % kQ, j = tid / kQ + e * (kThreads / kQ);
    const int64_t gi = i0 + i, gj = j0 + j;
    const float x = (gi < rows && gj < n) ? (float)A[gi + gj * lda] : 0.f;
    t[j][i] = x;
    if (gi < rows && gj < n) {
      const T h = (T)x;
      Lh[gj * ldl + gi] = h;
      if (SPLIT) Ll[gj * ldl + gi] = (T)(x - (float)h);
    }
  }
  __syncthreads();
  for (int e = 0; e < kQ * kQ / kThreads; ++e) {
    const int j = tid % kQ, i = tid / kQ + e * (kThreads / kQ);
    const int64_t gi = i0 + i, gj = j0 + j;
    if (gi < rows && gj < n) {
      const float x = t[j][i];
      const T h = (T)x;
      Rh[gi * ldr + gj] = h;
      if (SPLIT) Rl[gi * ldr + gj] = (T)(x - (float)h);
    }
  }
}

// FP8 scales of a window: per-row and per-column max |x| (as float bits,
// which order like unsigned integers for x >= 0), folded in with atomics
// from 64 x 64 tiles.  rmax and cmax start at zero.
__global__ __launch_bounds__(kThreads) void k_absmax_fp8(const double* A, int64_t lda, int64_t rows, int64_t n,
                                                         unsigned* rmax, unsigned* cmax) {
  __shared__ float t[kQ][kQ + 1];  // t[j][i]
  const int tid = threadIdx.x;
  const int64_t i0 = (int64_t)blockIdx.x * kQ, j0 = (int64_t)blockIdx.y * kQ;
  for (int e = 0; e < kQ * kQ / kThreads; ++e) {
    const int i = tid % kQ, j = tid / kQ + e * (kThreads / kQ);
    const int64_t gi = i0 + i, gj = j0 + j;
    t[j][i] = (gi < rows && gj < n) ? fabsf((float)A[gi + gj * lda]) : 0.f;
  }
  __syncthreads();
  if (tid < kQ) {  // row i0 + tid over the tile's columns
    float m = 0.f;
    for (int j = 0; j < kQ; ++j) m = fmaxf(m, t[j][tid]);
    if (i0 + tid < rows && m > 0.f) atomicMax(rmax + i0 + tid, __float_as_uint(m));
  } else if (tid < 2 * kQ) {  // column j0 + tid - 64 over the tile's rows
    const int j = tid - kQ;
    float m = 0.f;
    for (int i = 0; i < kQ; ++i) m = fmaxf(m, t[j][i]);
    if (j0 + j < n && m > 0.f) atomicMax(cmax + j0 + j, __float_as_uint(m));
  }
}

// Window -> e4m3 copies: L (column-major; column j scaled by 2^-e(cmax_j)) and
// R (row-major; row i scaled by 2^-e(rmax_i)).  Four elements per thread.
__global__ __launch_bounds__(kThreads) void k_quant_fp8(const double* A, int64_t lda, int64_t rows, int64_t n,
                                                        const float* rmax, const float* cmax, uint8_t* Lq,
                                                        int64_t ldl, uint8_t* Rq, int64_t ldr) {
  __shared__ float t[kQ][kQ + 1];  // t[j][i]
  const int tid = threadIdx.x;
  const int64_t i0 = (int64_t)blockIdx.x * kQ, j0 = (int64_t)blockIdx.y * kQ;
  for (int e = 0; e < kQ * kQ / kThreads; ++e) {
    const int i = tid % kQ, j = tid / kQ + e * (kThreads / kQ);
    const int64_t gi = i0 + i, gj = j0 + j;
    t[j][i] = (gi < rows && gj < n) ? (float)A[gi + gj * lda] : 0.f;
  }
  __syncthreads();
  // L: column j, rows 4c..4c+3 of the tile
  for (int e = tid; e < kQ * kQ / 4; e += kThreads) {
    const int j = e / (kQ / 4), c = e % (kQ / 4);
    const int64_t gj = j0 + j, gi = i0 + 4 * c;
    if (gj < n && gi < rows) {
      const float inv = ldexpf(1.f, -block_exponent(cmax[gj]));
      const uint32_t w = pack4_e4m3(t[j][4 * c] * inv, t[j][4 * c + 1] * inv, t[j][4 * c + 2] * inv,
                                    t[j][4 * c + 3] * inv);
      uint8_t* dst = Lq + gj * ldl + gi;
      const int lim = (int)((rows - gi) < 4 ? (rows - gi) : 4);
      if (lim == 4) *reinterpret_cast<uint32_t*>(dst) = w;
      else for (int q = 0; q < lim; ++q) dst[q] = (uint8_t)(w >> (8 * q));
    }
  }
  // R: row i, columns 4c..4c+3 of the tile
  for (int e = tid; e < kQ * kQ / 4; e += kThreads) {
    const int i = e / (kQ / 4), c = e % (kQ / 4);
    const int64_t gi = i0 + i, gj = j0 + 4 * c;
    if (gi < rows && gj < n) {
      const float inv = ldexpf(1.f, -block_exponent(rmax[gi]));
      const uint32_t w = pack4_e4m3(t[4 * c][i] * inv, t[4 * c + 1][i] * inv, t[4 * c + 2][i] * inv,
                                    t[4 * c + 3][i] * inv);
      uint8_t* dst = Rq + gi * ldr + gj;
      const int lim = (int)((n - gj) < 4 ? (n - gj) : 4);
      if (lim == 4) *reinterpret_cast<uint32_t*>(dst) = w;
      else for (int q = 0; q < lim; ++q) dst[q] = (uint8_t)(w >> (8 * q));
    }
  }
}

// Map -> row-major reduced copy (rows x cols logical, column-major source).
template <class T>
__global__ void k_convert_map(const double* src, int64_t ld, int64_t rows, int64_t cols, T* hi, T* lo, int64_t ldo) {
  const int64_t total = rows * cols;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e % rows, c = e / rows;
    const float x = (float)src[r + c * ld];
    const T h = (T)x;
    hi[r * ldo + c] = h;
    if (lo) lo[r * ldo + c] = (T)(x - (float)h);
  }
}
// Map row r -> e4m3 row (row-major, ld ldo) scaled by 2^-e(max_r), with
// max_r written to amax[r].  One block per row.
__global__ __launch_bounds__(kThreads) void k_convert_map_fp8(const double* src, int64_t ld, int64_t rows,
                                                              int64_t cols, uint8_t* q, int64_t ldo, float* amax) {
  __shared__ float red[kThreads];
  const int64_t r = blockIdx.x;
  const int tid = threadIdx.x;
  float m = 0.f;
  for (int64_t c = tid; c < cols; c += kThreads) m = fmaxf(m, fabsf((float)src[r + c * ld]));
  red[tid] = m;
  __syncthreads();
  for (int h = kThreads / 2; h > 0; h >>= 1) {
    if (tid < h) red[tid] = fmaxf(red[tid], red[tid + h]);
    __syncthreads();
  }
  const float mx = red[0];
  const float inv = ldexpf(1.f, -block_exponent(mx));
  for (int64_t c0 = 4 * (int64_t)tid; c0 < cols; c0 += 4 * kThreads) {
    float v[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) v[t] = (c0 + t < cols) ? (float)src[r + (c0 + t) * ld] * inv : 0.f;
    const uint32_t w = pack4_e4m3(v[0], v[1], v[2], v[3]);
    const int64_t lim = cols - c0 < 4 ? cols - c0 : 4;
    if (lim == 4) *reinterpret_cast<uint32_t*>(q + r * ldo + c0) = w;
    else for (int t = 0; t < lim; ++t) q[r * ldo + c0 + t] = (uint8_t)(w >> (8 * t));
  }
  if (tid == 0) amax[r] = mx;
}

// ---- NT GEMM: C(i, j) = sum_k A(i, k) B(j, k), A and B K-contiguous ----------
struct NtArgs {
  int64_t M, N, K;
  const void* ah;  // A planes (hi, lo) or fp8 bytes
  const void* al;
  const float* amax_a;  // fp8: max |x| of each row of A (its scale is 2^e(max))
  int64_t lda, koff_a;
  const void* bh;
  const void* bl;
  const float* amax_b;  // fp8: same for the rows of B
  int64_t ldb;
  GemmOut out;
  int tiles_m;
};

template <int BM, int BN, bool SPLIT>
__global__ __launch_bounds__(kThreads) void k_gemm_nt_bf16(NtArgs g) {
  constexpr int BK = 64, LDK = BK + 8;
  constexpr int TM = BM / 32, TN = BN / 32;
  constexpr int NP = SPLIT ? 2 : 1;
  constexpr int CA = BM * BK / 8 / kThreads, CB = BN * BK / 8 / kThreads;
  __shared__ __attribute__((aligned(16))) __bf16 sA[NP][BM][LDK];
  __shared__ __attribute__((aligned(16))) __bf16 sB[NP][BN][LDK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t m0 = (int64_t)(t % g.tiles_m) * BM;
  const int64_t n0 = (int64_t)(t / g.tiles_m) * BN;
  const __bf16* A[2] = {static_cast<const __bf16*>(g.ah), static_cast<const __bf16*>(g.al)};
  const __bf16* B[2] = {static_cast<const __bf16*>(g.bh), static_cast<const __bf16*>(g.bl)};

  f4v acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = f4v{0.f, 0.f, 0.f, 0.f};
  bf16x8 ra[NP][CA], rb[NP][CB];

  auto load_chunk = [&](const __bf16* base, int64_t row, int64_t nrows, int64_t ld, int64_t koff, int64_t k) {
    const __bf16* src = base + row * ld + koff + k;
    bf16x8 v;
    if (row < nrows && k + 8 <= g.K && (((uintptr_t)src) & 15) == 0) {
      v = *reinterpret_cast<const bf16x8*>(src);
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = (row < nrows && k + q < g.K) ? src[q] : (__bf16)0.0f;
    }
    return v;
  };
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int e = 0; e < CA; ++e) {
      const int c = tid + e * kThreads, row = c / (BK / 8), kc = (c % (BK / 8)) * 8;
#pragma unroll
      for (int h = 0; h < NP; ++h) ra[h][e] = load_chunk(A[h], m0 + row, g.M, g.lda, g.koff_a, k0 + kc);
    }
#pragma unroll
    for (int e = 0; e < CB; ++e) {
      const int c = tid + e * kThreads, row = c / (BK / 8), kc = (c % (BK / 8)) * 8;
#pragma unroll
      for (int h = 0; h < NP; ++h) rb[h][e] = load_chunk(B[h], n0 + row, g.N, g.ldb, 0, k0 + kc);
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int e = 0; e < CA; ++e) {
      const int c = tid + e * kThreads, row = c / (BK / 8), kc = (c % (BK / 8)) * 8;
#pragma unroll
      for (int h = 0; h < NP; ++h) *reinterpret_cast<bf16x8*>(&sA[h][row][kc]) = ra[h][e];
    }
#pragma unroll
    for (int e = 0; e < CB; ++e) {
      const int c = tid + e * kThreads, row = c / (BK / 8), kc = (c % (BK / 8)) * 8;
#pragma unroll
      for (int h = 0; h < NP; ++h) *reinterpret_cast<bf16x8*>(&sB[h][row][kc]) = rb[h][e];
    }
  };

  gload(0);
  for (int64_t k0 = 0; k0 < g.K; k0 += BK) {
    __syncthreads();
    sstore();
    __syncthreads();
    if (k0 + BK < g.K) gload(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 32) {
      const int kk = ks + 8 * (lane >> 4);
      bf16x8 a[NP][TM], b[NP][TN];
#pragma unroll
      for (int h = 0; h < NP; ++h) {
#pragma unroll
        for (int x = 0; x < TM; ++x)
          a[h][x] = *reinterpret_cast<const bf16x8*>(&sA[h][wm * (BM / 2) + x * 16 + (lane & 15)][kk]);
#pragma unroll
        for (int y = 0; y < TN; ++y)
          b[h][y] = *reinterpret_cast<const bf16x8*>(&sB[h][wn * (BN / 2) + y * 16 + (lane & 15)][kk]);
      }
#pragma unroll
      for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TN; ++y) {
          acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][x], b[0][y], acc[x][y], 0, 0, 0);
          if (SPLIT) {
            acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][x], b[NP - 1][y], acc[x][y], 0, 0, 0);
            acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[NP - 1][x], b[0][y], acc[x][y], 0, 0, 0);
          }
        }
    }
  }
  // 16x16 f32 MFMA result: lane l holds rows 4*(l/16)+v of column l%16.
#pragma unroll
  for (int x = 0; x < TM; ++x)
#pragma unroll
    for (int y = 0; y < TN; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t i = m0 + wm * (BM / 2) + x * 16 + 4 * (lane >> 4) + v;
        const int64_t j = n0 + wn * (BN / 2) + y * 16 + (lane & 15);
        if (i < g.M && j < g.N) store_out(g.out, i, j, (double)acc[x][y][v]);
      }
}

// FP32 NT GEMM on v_mfma_f32_16x16x4_f32.  Each lane reads 4 consecutive K
// elements of its row (K quarter l/16 of a 16-wide chunk) with one 16-byte LDS
// read and issues 4 MFMAs, element q of A with element q of B: every K index
// of the chunk is met exactly once, so the sum is the same contraction.
template <int BM, int BN>
__global__ __launch_bounds__(kThreads) void k_gemm_nt_f32(NtArgs g) {
  constexpr int BK = 32, LDK = BK + 4;
  constexpr int TM = BM / 32, TN = BN / 32;
  constexpr int CA = BM * BK / 4 / kThreads, CB = BN * BK / 4 / kThreads;
  __shared__ __attribute__((aligned(16))) float sA[BM][LDK];
  __shared__ __attribute__((aligned(16))) float sB[BN][LDK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t m0 = (int64_t)(t % g.tiles_m) * BM;
  const int64_t n0 = (int64_t)(t / g.tiles_m) * BN;
  const float* A = static_cast<const float*>(g.ah);
  const float* B = static_cast<const float*>(g.bh);

  f4v acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = f4v{0.f, 0.f, 0.f, 0.f};
  f4v ra[CA], rb[CB];

  auto load4 = [&](const float* base, int64_t row, int64_t nrows, int64_t ld, int64_t koff, int64_t k) {
    const float* src = base + row * ld + koff + k;
    f4v v;
    if (row < nrows && k + 4 <= g.K && (((uintptr_t)src) & 15) == 0) {
      v = *reinterpret_cast<const f4v*>(src);
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = (row < nrows && k + q < g.K) ? src[q] : 0.f;
    }
    return v;
  };
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int e = 0; e < CA; ++e) {
      const int c = tid + e * kThreads, row = c / (BK / 4), kc = (c % (BK / 4)) * 4;
      ra[e] = load4(A, m0 + row, g.M, g.lda, g.koff_a, k0 + kc);
    }
#pragma unroll
    for (int e = 0; e < CB; ++e) {
      const int c = tid + e * kThreads, row = c / (BK / 4), kc = (c % (BK / 4)) * 4;
      rb[e] = load4(B, n0 + row, g.N, g.ldb, 0, k0 + kc);
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int e = 0; e < CA; ++e) {
      const int c = tid + e * kThreads, row = c / (BK / 4), kc = (c % (BK / 4)) * 4;
      *reinterpret_cast<f4v*>(&sA[row][kc]) = ra[e];
    }
#pragma unroll
    for (int e = 0; e < CB; ++e) {
      const int c = tid + e * kThreads, row = c / (BK / 4), kc = (c % (BK / 4)) * 4;
      *reinterpret_cast<f4v*>(&sB[row][kc]) = rb[e];
    }
  };

  gload(0);
  for (int64_t k0 = 0; k0 < g.K; k0 += BK) {
    __syncthreads();
    sstore();
    __syncthreads();
    if (k0 + BK < g.K) gload(k0 + BK);
#pragma unroll
    for (int ks = 0; ks < BK; ks += 16) {
      const int kk = ks + 4 * (lane >> 4);
      f4v a[TM], b[TN];
#pragma unroll
      for (int x = 0; x < TM; ++x) a[x] = *reinterpret_cast<const f4v*>(&sA[wm * (BM / 2) + x * 16 + (lane & 15)][kk]);
#pragma unroll
      for (int y = 0; y < TN; ++y) b[y] = *reinterpret_cast<const f4v*>(&sB[wn * (BN / 2) + y * 16 + (lane & 15)][kk]);
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int x = 0; x < TM; ++x)
#pragma unroll
          for (int y = 0; y < TN; ++y)
            acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[x][q], b[y][q], acc[x][y], 0, 0, 0);
    }
  }
#pragma unroll
  for (int x = 0; x < TM; ++x)
#pragma unroll
    for (int y = 0; y < TN; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t i = m0 + wm * (BM / 2) + x * 16 + 4 * (lane >> 4) + v;
        const int64_t j = n0 + wn * (BN / 2) + y * 16 + (lane & 15);
        if (i < g.M && j < g.N) store_out(g.out, i, j, (double)acc[x][y][v]);
      }
}

// FP8 NT GEMM: BK = 128 = one MFMA; lane l feeds 32 bytes (row l%16, K
// indices 32*(l/16) .. +31).  Block scales are all 2^0 (E8M0 127); the row
// scales of A and B multiply the FP32 sum in the epilogue.
template <int BM, int BN>
__global__ __launch_bounds__(kThreads) void k_gemm_nt_fp8(NtArgs g) {
  constexpr int BK = 128, LDK = BK + 16;
  constexpr int TM = BM / 32, TN = BN / 32;
  constexpr int CA = BM * BK / 16 / kThreads, CB = BN * BK / 16 / kThreads;
  __shared__ __attribute__((aligned(16))) uint8_t sA[BM][LDK];
  __shared__ __attribute__((aligned(16))) uint8_t sB[BN][LDK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t m0 = (int64_t)(t % g.tiles_m) * BM;
  const int64_t n0 = (int64_t)(t / g.tiles_m) * BN;
  const uint8_t* A = static_cast<const uint8_t*>(g.ah);
  const uint8_t* B = static_cast<const uint8_t*>(g.bh);

  f4v acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = f4v{0.f, 0.f, 0.f, 0.f};
  i4v ra[CA], rb[CB];

  auto load16 = [&](const uint8_t* base, int64_t row, int64_t nrows, int64_t ld, int64_t koff, int64_t k) {
    const uint8_t* src = base + row * ld + koff + k;
    i4v v;
    if (row < nrows && k + 16 <= g.K && (((uintptr_t)src) & 15) == 0) {
      v = *reinterpret_cast<const i4v*>(src);
    } else {
      uint8_t b[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) b[q] = (row < nrows && k + q < g.K) ? src[q] : (uint8_t)0;
      v = *reinterpret_cast<const i4v*>(b);
    }
    return v;
  };
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int e = 0; e < CA; ++e) {
      const int c = tid + e * kThreads, row = c / (BK / 16), kc = (c % (BK / 16)) * 16;
      ra[e] = load16(A, m0 + row, g.M, g.lda, g.koff_a, k0 + kc);
    }
#pragma unroll
    for (int e = 0; e < CB; ++e) {
      const int c = tid + e * kThreads, row = c / (BK / 16), kc = (c % (BK / 16)) * 16;
      rb[e] = load16(B, n0 + row, g.N, g.ldb, 0, k0 + kc);
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int e = 0; e < CA; ++e) {
      const int c = tid + e * kThreads, row = c / (BK / 16), kc = (c % (BK / 16)) * 16;
      *reinterpret_cast<i4v*>(&sA[row][kc]) = ra[e];
    }
#pragma unroll
    for (int e = 0; e < CB; ++e) {
      const int c = tid + e * kThreads, row = c / (BK / 16), kc = (c % (BK / 16)) * 16;
      *reinterpret_cast<i4v*>(&sB[row][kc]) = rb[e];
    }
  };

  gload(0);
  for (int64_t k0 = 0; k0 < g.K; k0 += BK) {
    __syncthreads();
    sstore();
    __syncthreads();
    if (k0 + BK < g.K) gload(k0 + BK);
    const int kb = lane >> 4;
    i8v a[TM], b[TN];
#pragma unroll
    for (int x = 0; x < TM; ++x) {
      const i4v* p = reinterpret_cast<const i4v*>(&sA[wm * (BM / 2) + x * 16 + (lane & 15)][32 * kb]);
      const i4v lo = p[0], hi = p[1];
      a[x] = i8v{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    }
#pragma unroll
    for (int y = 0; y < TN; ++y) {
      const i4v* p = reinterpret_cast<const i4v*>(&sB[wn * (BN / 2) + y * 16 + (lane & 15)][32 * kb]);
      const i4v lo = p[0], hi = p[1];
      b[y] = i8v{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    }
#pragma unroll
    for (int x = 0; x < TM; ++x)
#pragma unroll
      for (int y = 0; y < TN; ++y)
        acc[x][y] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a[x], b[y], acc[x][y], 0, 0, 0, 127, 0, 127);
  }
  double sb[TN];
#pragma unroll
  for (int y = 0; y < TN; ++y) {
    const int64_t j = n0 + wn * (BN / 2) + y * 16 + (lane & 15);
    sb[y] = (j < g.N) ? ldexp(1.0, block_exponent(g.amax_b[j])) : 0.0;
  }
#pragma unroll
  for (int x = 0; x < TM; ++x)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t i = m0 + wm * (BM / 2) + x * 16 + 4 * (lane >> 4) + v;
      if (i >= g.M) continue;
      const double sa = ldexp(1.0, block_exponent(g.amax_a[i]));
#pragma unroll
      for (int y = 0; y < TN; ++y) {
        const int64_t j = n0 + wn * (BN / 2) + y * 16 + (lane & 15);
        if (j < g.N) store_out(g.out, i, j, (double)acc[x][y][v] * sa * sb[y]);
      }
    }
}

struct Tile {
  int bm, bn;
};
Tile pick_tile(int64_t M, int64_t N) {
  const Tile cfgs[] = {{128, 128}, {128, 64}, {64, 128}, {64, 64}};
  Tile best = cfgs[3];
  int64_t best_cost = INT64_MAX;
  for (const Tile& c : cfgs) {
    const int64_t tiles = cdiv(M, c.bm) * cdiv(N, c.bn);
    const int64_t cost = cdiv(tiles, 256) * c.bm * c.bn;
    if (cost < best_cost) { best_cost = cost; best = c; }
  }
  return best;
}

void gemm_nt(hipStream_t st, int prec, NtArgs a) {
  const Tile tl = pick_tile(a.M, a.N);
  a.tiles_m = (int)cdiv(a.M, tl.bm);
  const unsigned tiles = (unsigned)(a.tiles_m * cdiv(a.N, tl.bn));
  const dim3 grid(tiles), block(kThreads);
#define SKEMA_NT_LAUNCH(BM_, BN_)                                                                   \
  do {                                                                                              \
    if (prec == SKEMA_PREC_FP8) hipLaunchKernelGGL((k_gemm_nt_fp8<BM_, BN_>), grid, block, 0, st, a); \
    else if (prec == SKEMA_PREC_FP32) hipLaunchKernelGGL((k_gemm_nt_f32<BM_, BN_>), grid, block, 0, st, a); \
    else if (prec == SKEMA_PREC_BF16X3)                                                             \
      hipLaunchKernelGGL((k_gemm_nt_bf16<BM_, BN_, true>), grid, block, 0, st, a);                  \
    else hipLaunchKernelGGL((k_gemm_nt_bf16<BM_, BN_, false>), grid, block, 0, st, a);              \
  } while (0)
  if (tl.bm == 128 && tl.bn == 128) SKEMA_NT_LAUNCH(128, 128);
  else if (tl.bm == 128) SKEMA_NT_LAUNCH(128, 64);
  else if (tl.bn == 128) SKEMA_NT_LAUNCH(64, 128);
  else SKEMA_NT_LAUNCH(64, 64);
#undef SKEMA_NT_LAUNCH
  check_launch("k_gemm_nt");
}

unsigned grid1(int64_t e) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(e, 256), 65536)); }

}  // namespace

void prepare_lowp_maps(Sketch& sk) {
  const int prec = sk.p.precision;
  SKEMA_REQUIRE(prec == SKEMA_PREC_BF16 || prec == SKEMA_PREC_BF16X3 || prec == SKEMA_PREC_FP8 ||
                    prec == SKEMA_PREC_FP32,
                kErrUnsupported, "unknown reduced-precision form");
  hipStream_t st = sk.ctx->stream;
  const int64_t k = sk.k, s = sk.s, m = sk.p.m, n = sk.p.n;
  // Row-major reduced maps, K contiguous, leading dimension padded to 128.
  sk.lp_ldups = cdiv(m, 128) * 128;
  sk.lp_ldomps = cdiv(n, 128) * 128;
  const size_t ue = (size_t)k * sk.lp_ldups, oe = (size_t)(k + s) * sk.lp_ldomps;
  if (prec == SKEMA_PREC_FP8) {
    // e4m3 rows followed by one float (row max |x|) per row
    sk.lp_ups.alloc(align_up(ue, 256) + (size_t)k * sizeof(float));
    sk.lp_omps.alloc(align_up(oe, 256) + (size_t)(k + s) * sizeof(float));
    hipLaunchKernelGGL(k_convert_map_fp8, dim3((unsigned)k), dim3(kThreads), 0, st, sk.upsL, sk.ldups, k, m,
                       sk.lp_ups.as<uint8_t>(), sk.lp_ldups, (float*)(sk.lp_ups.as<uint8_t>() + align_up(ue, 256)));
    hipLaunchKernelGGL(k_convert_map_fp8, dim3((unsigned)(k + s)), dim3(kThreads), 0, st, sk.omps.as<double>(),
                       k + s, k + s, n, sk.lp_omps.as<uint8_t>(), sk.lp_ldomps,
                       (float*)(sk.lp_omps.as<uint8_t>() + align_up(oe, 256)));
    check_launch("k_convert_map_fp8");
  } else if (prec == SKEMA_PREC_FP32) {
    sk.lp_ups.alloc(ue * 4);
    sk.lp_omps.alloc(oe * 4);
    hipLaunchKernelGGL(k_convert_map<float>, dim3(grid1(k * m)), dim3(256), 0, st, sk.upsL, sk.ldups, k, m,
                       sk.lp_ups.as<float>(), (float*)nullptr, sk.lp_ldups);
    hipLaunchKernelGGL(k_convert_map<float>, dim3(grid1((k + s) * n)), dim3(256), 0, st, sk.omps.as<double>(), k + s,
                       k + s, n, sk.lp_omps.as<float>(), (float*)nullptr, sk.lp_ldomps);
    check_launch("k_convert_map<float>");
  } else {
    const bool split = prec == SKEMA_PREC_BF16X3;
    sk.lp_ups.alloc(ue * 2 * (split ? 2 : 1));
    sk.lp_omps.alloc(oe * 2 * (split ? 2 : 1));
    __bf16* uh = sk.lp_ups.as<__bf16>();
    __bf16* oh = sk.lp_omps.as<__bf16>();
    hipLaunchKernelGGL(k_convert_map<__bf16>, dim3(grid1(k * m)), dim3(256), 0, st, sk.upsL, sk.ldups, k, m, uh,
                       split ? uh + ue : nullptr, sk.lp_ldups);
    hipLaunchKernelGGL(k_convert_map<__bf16>, dim3(grid1((k + s) * n)), dim3(256), 0, st, sk.omps.as<double>(),
                       k + s, k + s, n, oh, split ? oh + oe : nullptr, sk.lp_ldomps);
    check_launch("k_convert_map<bf16>");
  }
  SKEMA_HIP(hipStreamSynchronize(st));
}

void update_dense_window_lowp(Sketch& sk, const double* H, int64_t ldh, int64_t row0, int64_t rows) {
  Ctx* c = sk.ctx;
  hipStream_t st = c->stream;
  const int prec = sk.p.precision;
  const int64_t k = sk.k, s = sk.s, m = sk.p.m, n = sk.p.n;
  const double eta = sk.p.eta, nu = sk.p.nu;
  const bool fp8 = prec == SKEMA_PREC_FP8, split = prec == SKEMA_PREC_BF16X3, f32 = prec == SKEMA_PREC_FP32;
  double* G = sk.G.as<double>();
  // window copies: L (column-major, ld ldl = rows padded), R (row-major, ld ldr = n padded);
  // fp8 adds the row and column max |x| of the window (floats)
  const int64_t ldl = cdiv(rows, 128) * 128, ldr = cdiv(n, 128) * 128;
  const size_t esz = fp8 ? 1 : (f32 ? 4 : 2), planes = split ? 2 : 1;
  const size_t le = (size_t)n * ldl, re = (size_t)rows * ldr;
  const size_t o_l = 0, o_r = align_up(le * esz * planes, 256), o_rmax = o_r + align_up(re * esz * planes, 256),
               o_cmax = o_rmax + (fp8 ? align_up((size_t)rows * 4, 256) : 0),
               total = o_cmax + (fp8 ? align_up((size_t)n * 4, 256) : 0);
  sk.win.reserve(total);
  char* base = sk.win.as<char>();
  float* rmax = (float*)(base + o_rmax);
  float* cmax = (float*)(base + o_cmax);
  {
    Timed t(c, "window");
    dim3 grid((unsigned)cdiv(rows, kQ), (unsigned)cdiv(n, kQ));
    if (fp8) {
      fill_i32(st, (int*)rmax, rows, 0);
      fill_i32(st, (int*)cmax, n, 0);
      hipLaunchKernelGGL(k_absmax_fp8, grid, dim3(kThreads), 0, st, H, ldh, rows, n, (unsigned*)rmax,
                         (unsigned*)cmax);
      hipLaunchKernelGGL(k_quant_fp8, grid, dim3(kThreads), 0, st, H, ldh, rows, n, (const float*)rmax,
                         (const float*)cmax, (uint8_t*)(base + o_l), ldl, (uint8_t*)(base + o_r), ldr);
    } else if (f32) {
      hipLaunchKernelGGL((k_quant<float, false>), grid, dim3(kThreads), 0, st, H, ldh, rows, n, (float*)(base + o_l),
                         (float*)nullptr, ldl, (float*)(base + o_r), (float*)nullptr, ldr);
    } else if (split) {
      __bf16* L = (__bf16*)(base + o_l);
      __bf16* R = (__bf16*)(base + o_r);
      hipLaunchKernelGGL((k_quant<__bf16, true>), grid, dim3(kThreads), 0, st, H, ldh, rows, n, L, L + le, ldl, R,
                         R + re, ldr);
    } else {
      hipLaunchKernelGGL((k_quant<__bf16, false>), grid, dim3(kThreads), 0, st, H, ldh, rows, n,
                         (__bf16*)(base + o_l), (__bf16*)nullptr, ldl, (__bf16*)(base + o_r), (__bf16*)nullptr, ldr);
    }
    check_launch("k_quant");
  }
  const size_t ue = (size_t)k * sk.lp_ldups, oe = (size_t)(k + s) * sk.lp_ldomps;
  {
    Timed t(c, "sketch_left");
    NtArgs a{};
    a.M = k; a.N = n; a.K = rows;
    a.ah = sk.lp_ups.p;
    a.al = split ? (const void*)(sk.lp_ups.as<__bf16>() + ue) : nullptr;
    a.amax_a = fp8 ? (const float*)(sk.lp_ups.as<uint8_t>() + align_up(ue, 256)) : nullptr;
    a.lda = sk.lp_ldups; a.koff_a = row0;
    a.bh = base + o_l;
    a.bl = split ? (const void*)((const __bf16*)(base + o_l) + le) : nullptr;
    a.amax_b = fp8 ? cmax : nullptr;
    a.ldb = ldl;
    a.out.C1 = sk.X.as<double>(); a.out.ldc1 = k; a.out.alpha1 = eta; a.out.beta1 = nu; a.out.nsplit = n;
    a.out.C2 = a.out.C1; a.out.ldc2 = k;
    gemm_nt(st, prec, a);
  }
  {
    Timed t(c, "sketch_right");
    NtArgs a{};
    a.M = rows; a.N = k + s; a.K = n;
    a.ah = base + o_r;
    a.al = split ? (const void*)((const __bf16*)(base + o_r) + re) : nullptr;
    a.amax_a = fp8 ? rmax : nullptr;
    a.lda = ldr; a.koff_a = 0;
    a.bh = sk.lp_omps.p;
    a.bl = split ? (const void*)(sk.lp_omps.as<__bf16>() + oe) : nullptr;
    a.amax_b = fp8 ? (const float*)(sk.lp_omps.as<uint8_t>() + align_up(oe, 256)) : nullptr;
    a.ldb = sk.lp_ldomps;
    a.out.C1 = sk.Y.as<double>() + row0; a.out.ldc1 = m; a.out.alpha1 = eta; a.out.beta1 = nu; a.out.nsplit = k;
    a.out.C2 = G; a.out.ldc2 = rows; a.out.alpha2 = 1.0; a.out.beta2 = 0.0;
    gemm_nt(st, prec, a);
  }
  {
    Timed t(c, "sketch_core");
    gemm_f64(st, 'N', 'N', s, s, rows, eta, sk.phiL + row0 * sk.ldphi, sk.ldphi, G, rows, nu, sk.Z.as<double>(), s,
             &c->gws);
  }
}

}  // namespace skema
