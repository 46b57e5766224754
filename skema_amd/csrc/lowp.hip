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
// applies the nu/eta update into the FP64 state.  The right product writes G
// (the window's A_w Psi^T block) straight in the core product's element type,
// column-major, so the core product Z += Phi[:, idx] * G is one more NT GEMM
// (BF16 for the bf16 and fp8 forms, split-BF16 and FP32 for theirs).  The
// whole finalize stays FP64.
//
// Forms:
//   fp32   : v_mfma_f32_16x16x4_f32 on FP32-rounded operands
//   bf16   : v_mfma_f32_16x16x32_bf16, one product
//   bf16x3 : same instruction, x = hi + lo (both bf16), a*b ~ hi*hi + hi*lo + lo*hi
//   fp8    : v_mfma_scale_f32_16x16x128_f8f6f4 on e4m3 operands.  The maps
//            carry one power-of-two scale per row over the whole reduction
//            (applied in the FP64 epilogue).  The window is quantised in one
//            read with an E8M0 scale per 32-element block of each column (L)
//            and each row (R), fed to the instruction's own block-scale
//            operands; windows that start off a 16-row boundary use one
//            scale per row and per column instead (two reads, epilogue).
//            Each scale is chosen from the block's (row's) max |x| so it
//            lands in [128, 256) before rounding.
//
// Products on 16-byte-aligned operands run on the wide NT kernels (k_ntx:
// A streamed to registers, B through LDS); the others on the 2 x 2-wave
// kernels (k_nt_*).
#include <cmath>
#include <type_traits>

#include "dma.hpp"
#include "lowp.hpp"
#include "util.hpp"

namespace skema {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef int i8v __attribute__((ext_vector_type(8)));
typedef int i4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef double d4v __attribute__((ext_vector_type(4)));
typedef double d2v __attribute__((ext_vector_type(2)));

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
    const float x = (gi < rows && gj < n) ? (float)__builtin_nontemporal_load(A + gi + gj * lda) : 0.f;
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

// Window -> e4m3 copies with E8M0 block scales in one read of A_w: L
// (column-major) with column j cut into 32-row blocks and R (row-major) with
// row i cut into 32-column blocks; each block is scaled by 2^-e of its own
// max |x| (max * 2^-e in [128, 256)) and its scale byte 127 + e goes to row
// j of Ls (i of Rs) at position (b % 4) ld/4 + b / 4 for block b (the order
// the NT kernels read: four consecutive K steps of one lane group in one
// dword).  Tile 128 rows x 64 columns, 4 waves.  A wave loads one column's
// 128 rows per instruction (two rows a lane), so 16 lanes hold one L block:
// L is quantised in registers (the block maximum over those 16 lanes by DPP
// moves) and stored two bytes a lane, 128 contiguous bytes per wave.  R
// needs the transpose: the tile goes through LDS once, then each thread owns
// one (row, column block) of R, and a wave's 64 lanes take 64 consecutive
// rows of the same column block, so each of its column reads of the tile
// covers 64 consecutive words.
constexpr int kMxR = 128, kMxC = 64;
__device__ __forceinline__ void quant32_store(const float (&v)[32], int valid, uint8_t* dst, uint8_t* sdst) {
  float m = 0.f;
#pragma unroll
  for (int q = 0; q < 32; ++q) m = fmaxf(m, fabsf(v[q]));
  const int e = block_exponent(m);
  const float inv = ldexpf(1.f, -e);
  uint32_t w[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) w[c] = pack4_e4m3(v[4 * c] * inv, v[4 * c + 1] * inv, v[4 * c + 2] * inv, v[4 * c + 3] * inv);
  if (valid == 32) {
    reinterpret_cast<i4v*>(dst)[0] = i4v{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
    reinterpret_cast<i4v*>(dst)[1] = i4v{(int)w[4], (int)w[5], (int)w[6], (int)w[7]};
  } else {
    for (int q = 0; q < valid; ++q) dst[q] = (uint8_t)(w[q >> 2] >> (8 * (q & 3)));
  }
  *sdst = (uint8_t)(127 + e);
}
// Maximum over the 16 lanes of a DPP row, every lane active: quad_perm
// [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror (lane i of 8 with 7 - i),
// row_mirror (lane i of 16 with 15 - i).
__device__ __forceinline__ float max_over_16_lanes(float m) {
  m = fmaxf(m, __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(m), 0xb1, 0xf, 0xf, false)));
  m = fmaxf(m, __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(m), 0x4e, 0xf, 0xf, false)));
  m = fmaxf(m, __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(m), 0x141, 0xf, 0xf, false)));
  return fmaxf(m, __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(m), 0x140, 0xf, 0xf, false)));
}
__global__ __launch_bounds__(kThreads) void k_quant_fp8_mx(const double* A, int64_t lda, int64_t rows, int64_t n,
                                                           uint8_t* Lq, int64_t ldl, uint8_t* Ls, int64_t ldls,
                                                           uint8_t* Rq, int64_t ldr, uint8_t* Rs, int64_t ldrs) {
  static_assert(kThreads == 2 * kMxR && kMxC == 2 * 32 && kMxR == 2 * 64, "tile / block shape");
  __shared__ __attribute__((aligned(16))) float t[kMxC][kMxR];  // t[j][i]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t i0 = (int64_t)blockIdx.x * kMxR, j0 = (int64_t)blockIdx.y * kMxC;
  // the lane's rows gi, gi + 1 of columns wave + 4 e, e < 16
  const int64_t gi = i0 + 2 * lane;
  const bool vec = ((lda & 1) == 0) && ((reinterpret_cast<uintptr_t>(A) & 15) == 0) && gi + 1 < rows;
  constexpr int NE = kMxC / 4;
  // all 16 loads of a lane are issued before the first use; non-temporal:
  // the window is read once, and streaming it through the caches would
  // evict the maps and the copies the products read next
  double2 v[NE];
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int64_t gj = j0 + wave + 4 * e;
    v[e] = double2{0.0, 0.0};
    if (gj < n) {
      const double* src = A + gi + gj * lda;
      if (vec) {
        const d2v x = __builtin_nontemporal_load(reinterpret_cast<const d2v*>(src));
        v[e] = double2{x[0], x[1]};
      } else {
        if (gi < rows) v[e].x = src[0];
        if (gi + 1 < rows) v[e].y = src[1];
      }
    }
  }
  const int64_t bl = (i0 + 32 * (lane >> 4)) / 32;  // the lane's L block (first row / 32)
#pragma unroll
  for (int e = 0; e < NE; ++e) {
    const int jt = wave + 4 * e;
    const float2 f = float2{(float)v[e].x, (float)v[e].y};
    *reinterpret_cast<float2*>(&t[jt][2 * lane]) = f;
    // rows past the window and columns past n hold zeros: no effect on a maximum
    const float m = max_over_16_lanes(fmaxf(fabsf(f.x), fabsf(f.y)));
    const int64_t gj = j0 + jt;
    if (gj < n && gi < rows) {
      const int ex = block_exponent(m);
      const float inv = ldexpf(1.f, -ex);
      const uint32_t w = (uint32_t)__builtin_amdgcn_cvt_pk_fp8_f32(f.x * inv, f.y * inv, 0, false);
      uint8_t* dst = Lq + gj * ldl + gi;
      if (gi + 1 < rows) *reinterpret_cast<uint16_t*>(dst) = (uint16_t)w;
      else dst[0] = (uint8_t)w;
      // the block's first lane (its first row is inside the window)
      if ((lane & 15) == 0) Ls[gj * ldls + (bl & 3) * (ldls / 4) + (bl >> 2)] = (uint8_t)(127 + ex);
    }
  }
  __syncthreads();
  {  // R: row i, columns 32 cb .. 32 cb + 31 of the tile
    const int i = tid & (kMxR - 1), cb = tid / kMxR;
    const int64_t gr = i0 + i, gc = j0 + 32 * cb;
    if (gr < rows && gc < n) {
      float r[32];
#pragma unroll
      for (int q = 0; q < 32; ++q) r[q] = t[32 * cb + q][i];
      const int valid = (int)(n - gc < 32 ? n - gc : 32);
      const int64_t b = gc / 32;
      quant32_store(r, valid, Rq + gr * ldr + gc, Rs + gr * ldrs + (b & 3) * (ldrs / 4) + (b >> 2));
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
  // the padding columns [cols, ldo) are zero: kernels that copy whole
  // 16-byte chunks past the end of a row multiply them by zero
  const int64_t pad = ldo - cols, ptotal = rows * pad;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < ptotal; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e / pad, c = cols + e % pad;
    hi[r * ldo + c] = (T)0.f;
    if (lo) lo[r * ldo + c] = (T)0.f;
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
  for (int64_t c = cols + tid; c < ldo; c += kThreads) q[r * ldo + c] = 0;  // zero padding
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
  // Reduced output of the columns j >= out.nsplit (G of the right product):
  // gtype 1 = bf16 (gh), 2 = bf16 hi + lo (gh, gl), 3 = fp32 (gh); element
  // (i, j - nsplit) at i + (j - nsplit) * ldg.  gtype 0 writes out.C2.
  int gtype;
  void* gh;
  void* gl;
  int64_t ldg;
  int tiles_m;
  int n_fast;  // wide kernels: tile order (see k_ntx)
  int64_t k_per_split;
  float* part;  // FP32 partials [split][M x N] when K is split
  // fp8 with E8M0 block scales (k_ntx, k_mx): the byte of (row r, 32-element
  // K block b) is at asc[r * ldas + (b % 4) * (ldas / 4) + b / 4] for A
  // (grouped by the lane group b % 4 that supplies it to a 16x16x128 MFMA,
  // so one dword holds four consecutive K steps), bsc[...] likewise for B.
  // An operand with block scales has no row scale in the epilogue.
  const uint8_t* asc;
  int64_t ldas;
  const uint8_t* bsc;
  int64_t ldbs;
  // k_mx (fp8, A with block scales) only: A stored in 64-byte column strips
  // when > 0 -- row r, K byte k at (k / 64) a_strip + 64 r + k % 64; a_strip
  // = bytes per strip
  int64_t a_strip;
  // k_mx<..., SEG> (bf16): K in three segments of seg_steps 128-byte steps
  // each, segment z read at byte offset seg_da[z] (A) / seg_db[z] (B) from
  // ah / bh -- split-BF16 as one bf16 product over [Ah | Ah | Al] and
  // [Bh | Bl | Bh] with nothing copied
  int seg_steps;
  int64_t seg_da0, seg_da1, seg_da2, seg_db0, seg_db1, seg_db2;
};

// Rows i0 .. i0+3 of column j (the four accumulator values of one lane in
// the 16x16 f32 result layout): contiguous in every column-major output, so
// G goes out as one 8-byte (bf16) or 16-byte (fp32) store and the FP64
// outputs as one 32-byte read-modify-write when aligned.
__device__ __forceinline__ void store_nt4(const NtArgs& g, int64_t i0, int64_t j, const double (&v)[4]);
__device__ __forceinline__ void store_nt(const NtArgs& g, int64_t i, int64_t j, double v) {
  if (g.gtype != 0 && j >= g.out.nsplit) {
    const int64_t e = i + (j - g.out.nsplit) * g.ldg;
    if (g.gtype == 3) {
      static_cast<float*>(g.gh)[e] = (float)v;
    } else {
      const float x = (float)v;
      const __bf16 h = (__bf16)x;
      static_cast<__bf16*>(g.gh)[e] = h;
      if (g.gtype == 2) static_cast<__bf16*>(g.gl)[e] = (__bf16)(x - (float)h);
    }
    return;
  }
  store_out(g.out, i, j, v);
}

__device__ __forceinline__ void store_nt4(const NtArgs& g, int64_t i0, int64_t j, const double (&v)[4]) {
  if (i0 + 4 > g.M) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (i0 + q < g.M) store_nt(g, i0 + q, j, v[q]);
    return;
  }
  if (g.gtype != 0 && j >= g.out.nsplit) {
    const int64_t e = i0 + (j - g.out.nsplit) * g.ldg;
    if (g.gtype == 3) {
      float* p = static_cast<float*>(g.gh) + e;
      if ((((uintptr_t)p) & 15) == 0) {
        *reinterpret_cast<f4v*>(p) = f4v{(float)v[0], (float)v[1], (float)v[2], (float)v[3]};
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q) p[q] = (float)v[q];
      }
      return;
    }
    bf16x4 h, l;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float x = (float)v[q];
      h[q] = (__bf16)x;
      l[q] = (__bf16)(x - (float)h[q]);
    }
    __bf16* ph = static_cast<__bf16*>(g.gh) + e;
    const bool al = (((uintptr_t)ph) & 7) == 0;
    if (al) *reinterpret_cast<bf16x4*>(ph) = h;
    else for (int q = 0; q < 4; ++q) ph[q] = h[q];
    if (g.gtype == 2) {
      __bf16* pl = static_cast<__bf16*>(g.gl) + e;
      if ((((uintptr_t)pl) & 7) == 0) *reinterpret_cast<bf16x4*>(pl) = l;
      else for (int q = 0; q < 4; ++q) pl[q] = l[q];
    }
    return;
  }
  const GemmOut& o = g.out;
  double* c;
  double alpha, beta;
  if (j < o.nsplit) { c = o.C1 + i0 + j * o.ldc1; alpha = o.alpha1; beta = o.beta1; }
  else { c = o.C2 + i0 + (j - o.nsplit) * o.ldc2; alpha = o.alpha2; beta = o.beta2; }
  if ((((uintptr_t)c) & 31) == 0) {
    d4v* cv = reinterpret_cast<d4v*>(c);
    if (beta != 0.0) {
      const d4v old = *cv;
      *cv = d4v{alpha * v[0] + beta * old[0], alpha * v[1] + beta * old[1], alpha * v[2] + beta * old[2],
                alpha * v[3] + beta * old[3]};
    } else {
      *cv = d4v{alpha * v[0], alpha * v[1], alpha * v[2], alpha * v[3]};
    }
  } else {
#pragma unroll
    for (int q = 0; q < 4; ++q) c[q] = (beta != 0.0) ? alpha * v[q] + beta * c[q] : alpha * v[q];
  }
}

// Epilogue of the NT kernels: lane l holds rows 4*(l/16)+v of column l%16 of
// each 16x16 tile, i.e. four consecutive rows of a column-major output.  fp8
// (SCALED) multiplies by the two operands' row scales.  With a K split the
// FP32 partial goes to g.part[z] for k_nt_reduce.
template <int TM, int TN, bool SA, bool SB>
__device__ __forceinline__ void nt_epilogue_at(const NtArgs& g, const f4v (&acc)[TM][TN], int64_t ibase,
                                               int64_t jbase, int lane) {
  const int l16 = lane & 15, lq = lane >> 4;
  double sb[TN];
#pragma unroll
  for (int y = 0; y < TN; ++y) {
    const int64_t j = jbase + y * 16 + l16;
    sb[y] = (SB && j < g.N) ? ldexp(1.0, block_exponent(g.amax_b[j])) : 1.0;
  }
#pragma unroll
  for (int x = 0; x < TM; ++x) {
    const int64_t i0 = ibase + x * 16 + 4 * lq;
    if (i0 >= g.M) continue;
    double sa[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) sa[v] = (SA && i0 + v < g.M) ? ldexp(1.0, block_exponent(g.amax_a[i0 + v])) : 1.0;
#pragma unroll
    for (int y = 0; y < TN; ++y) {
      const int64_t j = jbase + y * 16 + l16;
      if (j >= g.N) continue;
      double val[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) val[v] = (double)acc[x][y][v] * sa[v] * sb[y];
      if (g.part) {
        float* p = g.part + (int64_t)blockIdx.z * g.M * g.N + j * g.M;
        if ((g.M & 3) == 0 && i0 + 4 <= g.M) {
          *reinterpret_cast<float4*>(p + i0) = float4{(float)val[0], (float)val[1], (float)val[2], (float)val[3]};
        } else {
#pragma unroll
          for (int v = 0; v < 4; ++v)
            if (i0 + v < g.M) p[i0 + v] = (float)val[v];
        }
      } else {
        store_nt4(g, i0, j, val);
      }
    }
  }
}
template <int TM, int TN, int BM, int BN, bool SCALED>
__device__ __forceinline__ void nt_epilogue(const NtArgs& g, const f4v (&acc)[TM][TN], int64_t m0, int64_t n0, int wm,
                                            int wn, int lane) {
  nt_epilogue_at<TM, TN, SCALED, SCALED>(g, acc, m0 + wm * (BM / 2), n0 + wn * (BN / 2), lane);
}

// ---- NT kernels: C(i, j) = sum over this block's K range (blockIdx.z of the
// split) of A(i, k) B(j, k), A and B K-contiguous, 2 x 2 waves per block ----
template <int BM, int BN, bool SPLIT>
__global__ __launch_bounds__(kThreads) void k_nt_bf16(NtArgs g) {
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
  const int64_t kbeg = (int64_t)blockIdx.z * g.k_per_split;
  const int64_t kend = (g.K < kbeg + g.k_per_split) ? g.K : kbeg + g.k_per_split;
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
    if (row < nrows && k + 8 <= kend && (((uintptr_t)src) & 15) == 0) {
      v = *reinterpret_cast<const bf16x8*>(src);
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = (row < nrows && k + q < kend) ? src[q] : (__bf16)0.0f;
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

  gload(kbeg);
  for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
    __syncthreads();
    sstore();
    __syncthreads();
    if (k0 + BK < kend) gload(k0 + BK);
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
  nt_epilogue<TM, TN, BM, BN, false>(g, acc, m0, n0, wm, wn, lane);
}

// FP32 NT GEMM on v_mfma_f32_16x16x4_f32.  Each lane reads 4 consecutive K
// elements of its row (K quarter l/16 of a 16-wide chunk) with one 16-byte LDS
// read and issues 4 MFMAs, element q of A with element q of B: every K index
// of the chunk is met exactly once, so the sum is the same contraction.
template <int BM, int BN>
__global__ __launch_bounds__(kThreads) void k_nt_f32(NtArgs g) {
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
  const int64_t kbeg = (int64_t)blockIdx.z * g.k_per_split;
  const int64_t kend = (g.K < kbeg + g.k_per_split) ? g.K : kbeg + g.k_per_split;
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
    if (row < nrows && k + 4 <= kend && (((uintptr_t)src) & 15) == 0) {
      v = *reinterpret_cast<const f4v*>(src);
    } else {
#pragma unroll
      for (int q = 0; q < 4; ++q) v[q] = (row < nrows && k + q < kend) ? src[q] : 0.f;
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

  gload(kbeg);
  for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
    __syncthreads();
    sstore();
    __syncthreads();
    if (k0 + BK < kend) gload(k0 + BK);
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
  nt_epilogue<TM, TN, BM, BN, false>(g, acc, m0, n0, wm, wn, lane);
}

// FP8 NT GEMM: BK = 128 = one MFMA; lane l feeds 32 bytes (row l%16, K
// indices 32*(l/16) .. +31).  Block scales are all 2^0 (E8M0 127); the row
// scales of A and B multiply the FP32 sum in the epilogue.
template <int BM, int BN>
__global__ __launch_bounds__(kThreads) void k_nt_fp8(NtArgs g) {
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
  const int64_t kbeg = (int64_t)blockIdx.z * g.k_per_split;
  const int64_t kend = (g.K < kbeg + g.k_per_split) ? g.K : kbeg + g.k_per_split;
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
    if (row < nrows && k + 16 <= kend && (((uintptr_t)src) & 15) == 0) {
      v = *reinterpret_cast<const i4v*>(src);
    } else {
      uint8_t b[16];
#pragma unroll
      for (int q = 0; q < 16; ++q) b[q] = (row < nrows && k + q < kend) ? src[q] : (uint8_t)0;
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

  gload(kbeg);
  for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
    __syncthreads();
    sstore();
    __syncthreads();
    if (k0 + BK < kend) gload(k0 + BK);
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
  nt_epilogue<TM, TN, BM, BN, true>(g, acc, m0, n0, wm, wn, lane);
}

// ---- Wide NT kernels (fp8, bf16): the four waves are stacked along M and
// each streams its own rows of A from global memory straight into MFMA
// registers, one K step ahead; only the B tile goes through LDS (double
// buffered, one barrier per step), so LDS carries B once per wave instead of
// A and B.  A K step is 128 bytes of every row (128 e4m3 or 64 bf16).  Lane
// l (group q = l/16) holds bytes [16q, 16q+16) and [64+16q, 64+16q+16) of
// row l%16 of a tile; B is read from LDS with the same assignment, so every
// MFMA contracts matching K indices (the order of a sum over K is free).
// Requires 16-byte aligned rows (pointers, leading dimensions, K offset).
// This K assignment is the f8f6f4 instruction's own: VGPRs 0-3 of lane l
// carry K [16q, 16q+16) and VGPRs 4-7 K [64+16q, 64+16q+16) of the 128, and
// the E8M0 scale of (row l%16, K block [32b, 32b+32)) is read from lane
// 16b + l%16 (both checked with single-lane probes on gfx950), so fp8
// operands may carry one scale per row and 32-element K block (ASC, BSC).
enum { kNxFp8 = 0, kNxBf16 = 1, kNxBf16x3 = 2 };  // kNxBf16x3: hi + lo planes (k_mx only)

__device__ __forceinline__ i4v ld16_masked(const uint8_t* p, int64_t rem) {
  if (rem >= 16) return *reinterpret_cast<const i4v*>(p);
  uint8_t b[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) b[q] = (q < rem) ? p[q] : (uint8_t)0;
  return *reinterpret_cast<const i4v*>(b);
}

template <int FORM, int TM, int TN, bool ASC, bool BSC>
__global__ __launch_bounds__(kThreads) void k_ntx(NtArgs g) {
  constexpr int BM = 64 * TM, BN = 16 * TN, SB = 128, LDB = SB + 16;
  constexpr int CB = BN * (SB / 16) / kThreads;  // 16-byte chunks of B per thread and step
  static_assert(CB >= 1 && BN * (SB / 16) % kThreads == 0, "B tile must split evenly over the block");
  constexpr int64_t ES = FORM == kNxFp8 ? 1 : 2;
  __shared__ __attribute__((aligned(16))) uint8_t sB[2][BN][LDB];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, l16 = lane & 15, q = lane >> 4;
  // Tiles of one XCD are consecutive in t; n_fast walks all N tiles of an M
  // tile before the next, so an XCD shares its A rows (the larger operand)
  // in its own L2 and only B is read by every XCD.
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int tiles_n = (int)((g.N + BN - 1) / BN);
  const int64_t m0 = (int64_t)(g.n_fast ? t / tiles_n : t % g.tiles_m) * BM;
  const int64_t n0 = (int64_t)(g.n_fast ? t % tiles_n : t / g.tiles_m) * BN;
  // byte range of K for this split
  const int64_t kb0 = (int64_t)blockIdx.z * g.k_per_split * ES;
  const int64_t kb1 = (g.K < ((int64_t)blockIdx.z + 1) * g.k_per_split ? g.K : ((int64_t)blockIdx.z + 1) * g.k_per_split) * ES;
  const int nsteps = (int)((kb1 - kb0 + SB - 1) / SB);

  // A rows of this lane (clamped into the matrix: rows past M are computed
  // and never stored)
  const uint8_t* pa[TM];
#pragma unroll
  for (int x = 0; x < TM; ++x) {
    int64_t row = m0 + wave * (16 * TM) + x * 16 + l16;
    row = row < g.M ? row : g.M - 1;
    pa[x] = static_cast<const uint8_t*>(g.ah) + row * g.lda * ES + g.koff_a * ES + kb0 + 16 * q;
  }
  const uint8_t* pb[CB];
  int brow[CB], bch[CB];
#pragma unroll
  for (int e = 0; e < CB; ++e) {
    const int c = tid + e * kThreads;
    brow[e] = c / (SB / 16);
    bch[e] = (c % (SB / 16)) * 16;
    int64_t row = n0 + brow[e];
    row = row < g.N ? row : g.N - 1;
    pb[e] = static_cast<const uint8_t*>(g.bh) + row * g.ldb * ES + kb0 + bch[e];
  }
  // Block scales (E8M0).  Lane group q needs K block 4 st + q of step st;
  // the scale rows are stored grouped by q (block b at (b % 4) * ldq + b / 4,
  // ldq = ld / 4), so four consecutive steps of a lane are one dword.
  const uint32_t* psa[TM];
  const uint32_t* psb[TN];
  const int64_t st0 = kb0 / SB;  // first step of this split (a multiple of 4)
  if (ASC) {
#pragma unroll
    for (int x = 0; x < TM; ++x) {
      int64_t row = m0 + wave * (16 * TM) + x * 16 + l16;
      row = row < g.M ? row : g.M - 1;
      psa[x] = reinterpret_cast<const uint32_t*>(g.asc + row * g.ldas + q * (g.ldas / 4) + st0);
    }
  }
  if (BSC) {
#pragma unroll
    for (int y = 0; y < TN; ++y) {
      int64_t row = n0 + y * 16 + l16;
      row = row < g.N ? row : g.N - 1;
      psb[y] = reinterpret_cast<const uint32_t*>(g.bsc + row * g.ldbs + q * (g.ldbs / 4) + st0);
    }
  }
  uint32_t swa[TM], swb[TN];

  i4v an[TM][2], rb[CB];
  // the scale dword of step st (re-read on every step of its group of four:
  // a cached load is cheaper than a branch in the pipelined loop)
  auto load_scales = [&](int st) {
#pragma unroll
    for (int x = 0; x < TM; ++x)
      if (ASC) swa[x] = psa[x][st >> 2];
#pragma unroll
    for (int y = 0; y < TN; ++y)
      if (BSC) swb[y] = psb[y][st >> 2];
  };
  auto load_full = [&](int st) {
    load_scales(st);
    const int64_t off = (int64_t)st * SB;
#pragma unroll
    for (int x = 0; x < TM; ++x) {
      an[x][0] = *reinterpret_cast<const i4v*>(pa[x] + off);
      an[x][1] = *reinterpret_cast<const i4v*>(pa[x] + off + 64);
    }
#pragma unroll
    for (int e = 0; e < CB; ++e) rb[e] = *reinterpret_cast<const i4v*>(pb[e] + off);
  };
  auto load_any = [&](int st) {
    const int64_t off = (int64_t)st * SB;
    const int64_t rem = kb1 - kb0 - off;  // valid bytes from the step start
    if (rem >= SB) {
      load_full(st);
      return;
    }
    load_scales(st);
#pragma unroll
    for (int x = 0; x < TM; ++x) {
      an[x][0] = ld16_masked(pa[x] + off, rem - 16 * q);
      an[x][1] = ld16_masked(pa[x] + off + 64, rem - 64 - 16 * q);
    }
#pragma unroll
    for (int e = 0; e < CB; ++e) rb[e] = ld16_masked(pb[e] + off, rem - bch[e]);
  };
  auto store_b = [&](int buf) {
#pragma unroll
    for (int e = 0; e < CB; ++e) *reinterpret_cast<i4v*>(&sB[buf][brow[e]][bch[e]]) = rb[e];
  };

  f4v acc[TM][TN];
#pragma unroll
  for (int x = 0; x < TM; ++x)
#pragma unroll
    for (int y = 0; y < TN; ++y) acc[x][y] = f4v{0.f, 0.f, 0.f, 0.f};

  // this step's operands leave the prefetch registers before the next
  // step's loads are issued into them
  i4v ac[TM][2];
  int sac[TM], sbc[TN];
  auto take = [&](int st) {
    const int sh = 8 * (st & 3);
#pragma unroll
    for (int x = 0; x < TM; ++x) {
      ac[x][0] = an[x][0];
      ac[x][1] = an[x][1];
      sac[x] = ASC ? (int)((swa[x] >> sh) & 0xffu) : 127;
    }
#pragma unroll
    for (int y = 0; y < TN; ++y) sbc[y] = BSC ? (int)((swb[y] >> sh) & 0xffu) : 127;
  };
  auto compute = [&](int buf) {
#pragma unroll
    for (int y = 0; y < TN; ++y) {
      const i4v b0 = *reinterpret_cast<const i4v*>(&sB[buf][y * 16 + l16][16 * q]);
      const i4v b1 = *reinterpret_cast<const i4v*>(&sB[buf][y * 16 + l16][64 + 16 * q]);
#pragma unroll
      for (int x = 0; x < TM; ++x) {
        const i4v a0 = ac[x][0], a1 = ac[x][1];
        if (FORM == kNxFp8) {
          const i8v av = i8v{a0[0], a0[1], a0[2], a0[3], a1[0], a1[1], a1[2], a1[3]};
          const i8v bv = i8v{b0[0], b0[1], b0[2], b0[3], b1[0], b1[1], b1[2], b1[3]};
          acc[x][y] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(av, bv, acc[x][y], 0, 0, 0, sac[x], 0, sbc[y]);
        } else {
          acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a0),
                                                              __builtin_bit_cast(bf16x8, b0), acc[x][y], 0, 0, 0);
          acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a1),
                                                              __builtin_bit_cast(bf16x8, b1), acc[x][y], 0, 0, 0);
        }
      }
    }
  };

  const int nfull = (int)((kb1 - kb0) / SB);
  if (nsteps > 0) {
    load_any(0);
    store_b(0);
  }
  __syncthreads();
  int st = 0;
  // steady state: the next step is a full one, so the body has no branch and
  // the order below is the order issued: next loads, this step's LDS reads
  // and MFMAs, then (sched_barrier: the compiler would otherwise hoist it)
  // the wait for the loads and the LDS store of the next B tile.
  for (; st + 1 < nfull; ++st) {
    take(st);
    load_full(st + 1);
    __builtin_amdgcn_sched_barrier(0);
    compute(st & 1);
    __builtin_amdgcn_sched_barrier(0);
    store_b((st + 1) & 1);
    __syncthreads();
  }
  // last full step and the partial tail
  for (; st < nsteps; ++st) {
    take(st);
    const bool more = st + 1 < nsteps;
    if (more) load_any(st + 1);
    compute(st & 1);
    if (more) store_b((st + 1) & 1);
    __syncthreads();
  }
  nt_epilogue_at<TM, TN, FORM == kNxFp8 && !ASC, FORM == kNxFp8 && !BSC>(g, acc, m0 + wave * (16 * TM), n0, lane);
}

// ---- Direct-to-LDS fp8 NT kernel ---------------------------------------------
// C = A * B^T on e4m3 or bf16 operands with WM x WN waves, wave tile
// (16 TM) x (16 TN), block tile BM = 16 TM WM by BN = 16 TN WN.  Both tiles of a 128-byte K step go
// global -> LDS through global_load_lds_dwordx4 (no staging registers) into
// an NS-slot ring, so NS - 1 steps are in flight while one is consumed.
// LDS rows are 128 bytes with the eight 16-byte chunks XOR-swizzled (chunk c
// of row r sits at position c ^ ((r >> 1) & 7)): a ds_read_b128 of 16
// consecutive rows at one chunk then touches 16 distinct 16-byte bank groups.
// One wave DMA instruction fills 1 KB = 8 rows; lane l lands at byte 16 l,
// i.e. row l / 8, position l % 8, and fetches the chunk that belongs there.
// A slot is ready after an explicit vmcnt wait (loads return in issue order)
// and s_barrier; no workgroup fence is issued in the loop, because a fence
// would drain every load in flight.
// Fragment K assignment and block scales as in k_ntx.  Requires 16-byte
// aligned rows and every split's K range a multiple of 128 bytes.
template <int FORM, int WM, int WN, int TM, int TN, int NS, bool ASC, bool BSC, bool SEG = false>
__global__ __launch_bounds__(64 * WM * WN, 1) void k_mx(NtArgs g) {
  constexpr int NW = WM * WN;  // waves: WM along M by WN along N
  constexpr int BM = 16 * TM * WM, BN = 16 * TN * WN, SB = 128;
  // DMA instructions per wave and step: 8 rows of A or B, or 16 scale rows
  constexpr int IA = BM / 8 / NW, IB = BN / 8 / NW;
  static_assert(IA * 8 * NW == BM && IB * 8 * NW == BN, "tile rows must split evenly over the waves");
  constexpr int64_t ES = FORM == kNxFp8 ? 1 : 2;  // bytes per element (per plane)
  static_assert(FORM == kNxFp8 || (!ASC && !BSC), "block scales are fp8 only");
  // rows with block scales and their DMA instructions per wave and step
  // (16 rows x 4 lane groups = 64 dwords per instruction)
  constexpr int SR = ASC ? BM : (BSC ? BN : 0), SC = SR / 16 / NW;
  static_assert(SC * 16 * NW == SR, "block-scaled tile rows must split evenly over the waves");
  // split-BF16 carries both planes of both operands: A hi | B hi | A lo | B lo
  constexpr int PL = FORM == kNxBf16x3 ? 2 : 1;
  constexpr int D = PL * (IA + IB) + SC;  // DMA instructions per wave and step
  // fp8 with A scales only, B fragments held (the right product): a row's
  // scale bytes of sixteen steps (16 bytes per lane group, four dwords of
  // four steps) are copied once per sixteen steps into the scale area of
  // slot 0, read from there once per four steps (sw) and shifted per step
  constexpr bool kScaleSel = FORM == kNxFp8 && ASC && !BSC && TM > TN && NS == 2;
  constexpr int LOOFF = (BM + BN) * SB, SCOFF = PL * LOOFF, STAGE = SCOFF + SR * (kScaleSel ? 64 : 16);
  static_assert(NS >= 2 && NS <= 4 && (NS - 2) * D < 64, "ring depth");
  __shared__ __attribute__((aligned(1024))) uint8_t smem[NS * STAGE];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, l16 = lane & 15, q = lane >> 4;
  const int wm = wave % WM, wn = wave / WM;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int tiles_n = (int)((g.N + BN - 1) / BN);
  const int64_t m0 = (int64_t)(g.n_fast ? t / tiles_n : t % g.tiles_m) * BM;
  const int64_t n0 = (int64_t)(g.n_fast ? t % tiles_n : t / g.tiles_m) * BN;
  // byte range of K for this split
  const int64_t kb0 = (int64_t)blockIdx.z * g.k_per_split * ES;
  const int64_t kb1 = (g.K < ((int64_t)blockIdx.z + 1) * g.k_per_split ? g.K : ((int64_t)blockIdx.z + 1) * g.k_per_split) * ES;
  const int nsteps = (int)((kb1 - kb0) / SB);

  // DMA sources: wave w fills rows [8 w IA, 8 (w + 1) IA) of the A tile and
  // [8 w IB, 8 (w + 1) IB) of the B tile; rows past M / N repeat the last row
  // (computed, never stored).
  const int lr = lane >> 3, lp = lane & 7;
  const uint8_t* ga[IA];
  const uint8_t* gb[IB];
  int64_t dal = 0, dbl = 0;  // byte distance from the hi to the lo plane (split-BF16)
  if (FORM == kNxBf16x3) {
    dal = static_cast<const uint8_t*>(g.al) - static_cast<const uint8_t*>(g.ah);
    dbl = static_cast<const uint8_t*>(g.bl) - static_cast<const uint8_t*>(g.bh);
  }
  // bytes from one K step of A to the next: 128 in rows, two strips in
  // 64-byte column strips
  const int64_t astep = (FORM == kNxFp8 && ASC && g.a_strip) ? 2 * g.a_strip : SB;
#pragma unroll
  for (int i = 0; i < IA; ++i) {
    const int r = 8 * (wave * IA + i) + lr;
    int64_t row = m0 + r;
    row = row < g.M ? row : g.M - 1;
    const int c = lp ^ ((r >> 1) & 7);  // the 16-byte chunk of the step this lane fetches
    if (FORM == kNxFp8 && ASC && g.a_strip)
      ga[i] = static_cast<const uint8_t*>(g.ah) + (kb0 / SB) * astep + (c >> 2) * g.a_strip + row * 64 + 16 * (c & 3);
    else
      ga[i] = static_cast<const uint8_t*>(g.ah) + (row * g.lda + g.koff_a) * ES + kb0 + 16 * c;
  }
#pragma unroll
  for (int j = 0; j < IB; ++j) {
    const int r = 8 * (wave * IB + j) + lr;
    int64_t row = n0 + r;
    row = row < g.N ? row : g.N - 1;
    gb[j] = static_cast<const uint8_t*>(g.bh) + row * g.ldb * ES + kb0 + 16 * (lp ^ ((r >> 1) & 7));
  }
  // Block scales (E8M0): byte (row, K block b) at row * ld + (b % 4) (ld / 4)
  // + b / 4, so the dword at row * ld + q (ld / 4) + 4 (st / 4) holds lane
  // group q's scales of steps 4 (st / 4) .. +3.  Each step copies that dword
  // of every (row, q) of the tile into the slot (row-major, 16 bytes a row).
  const uint8_t* gs[SC > 0 ? SC : 1];
#pragma unroll
  for (int i = 0; i < SC; ++i) {
    const int r = 16 * (wave * SC + i) + (lane >> 2);
    const int64_t lim = ASC ? g.M : g.N;
    int64_t row = (ASC ? m0 : n0) + r;
    row = row < lim ? row : lim - 1;
    const int64_t ld = ASC ? g.ldas : g.ldbs;
    gs[i] = (ASC ? g.asc : g.bsc) + row * ld + (lane & 3) * (ld / 4) + kb0 / SB;
  }
  auto issue = [&](int st) {
    uint8_t* slot = smem + (st % NS) * STAGE;
    int64_t off = (int64_t)st * SB, offa = (int64_t)st * astep;
    if constexpr (SEG) {
      // K segments: the absolute step's place in its segment, from that
      // segment's plane (ga and gb already hold this split's start kb0)
      const int sa = (int)(kb0 / SB) + st;
      const int z = sa / g.seg_steps;
      const int64_t r = (int64_t)(sa - z * g.seg_steps) * SB - kb0;
      offa = r + (z == 0 ? g.seg_da0 : (z == 1 ? g.seg_da1 : g.seg_da2));
      off = r + (z == 0 ? g.seg_db0 : (z == 1 ? g.seg_db1 : g.seg_db2));
    }
#pragma unroll
    for (int i = 0; i < IA; ++i)
      __builtin_amdgcn_global_load_lds((const void*)(ga[i] + offa),
                                       (SKEMA_LDS void*)(slot + 1024 * (wave * IA + i)), 16, 0, 0);
#pragma unroll
    for (int j = 0; j < IB; ++j)
      __builtin_amdgcn_global_load_lds((const void*)(gb[j] + off),
                                       (SKEMA_LDS void*)(slot + BM * SB + 1024 * (wave * IB + j)), 16, 0, 0);
    if constexpr (FORM == kNxBf16x3) {
#pragma unroll
      for (int i = 0; i < IA; ++i)
        __builtin_amdgcn_global_load_lds((const void*)(ga[i] + dal + off),
                                         (SKEMA_LDS void*)(slot + LOOFF + 1024 * (wave * IA + i)), 16, 0, 0);
#pragma unroll
      for (int j = 0; j < IB; ++j)
        __builtin_amdgcn_global_load_lds((const void*)(gb[j] + dbl + off),
                                         (SKEMA_LDS void*)(slot + LOOFF + BM * SB + 1024 * (wave * IB + j)), 16,
                                         0, 0);
    }
    if constexpr (kScaleSel) {
      // sixteen steps of (row, lane group) scale bytes: 16-byte aligned
      // (splits start on a 16-step boundary, lane-group rows padded to 16)
      if ((st & 15) == 0) {
#pragma unroll
        for (int i = 0; i < SC; ++i)
          __builtin_amdgcn_global_load_lds((const void*)(gs[i] + st), (SKEMA_LDS void*)(slot + SCOFF + 1024 * (wave * SC + i)),
                                           16, 0, 0);
      }
    } else {
#pragma unroll
      for (int i = 0; i < SC; ++i)
        __builtin_amdgcn_global_load_lds((const void*)(gs[i] + 4 * (st >> 2)),
                                         (SKEMA_LDS void*)(slot + SCOFF + 256 * (wave * SC + i)), 4, 0, 0);
    }
  };

  f4v acc[TM][TN];
#pragma unroll
  for (int x = 0; x < TM; ++x)
#pragma unroll
    for (int y = 0; y < TN; ++y) acc[x][y] = f4v{0.f, 0.f, 0.f, 0.f};

  // swizzled byte offsets of this lane's two chunks (q and 4 + q) in a row
  const int so0 = 16 * (q ^ ((l16 >> 1) & 7)), so1 = so0 ^ 64;
  auto frag = [&](const uint8_t* tile, int r) {
    const i4v lo = *reinterpret_cast<const i4v*>(tile + r * SB + so0);
    const i4v hi = *reinterpret_cast<const i4v*>(tile + r * SB + so1);
    return i8v{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
  };

#pragma unroll
  for (int p = 0; p < NS - 1; ++p)
    if (p < nsteps) issue(p);

  int sw[TM];
#pragma unroll
  for (int x = 0; x < TM; ++x) sw[x] = 0x7f7f7f7f;
  for (int st = 0; st < nsteps; ++st) {
    // slot st is complete once at most the DMA groups issued after it are
    // outstanding (fewer of those near the end)
    const int after = nsteps - 1 - st;
    if (NS >= 4 && after >= 2) wait_vm<(NS >= 4 ? (NS - 2) * D : 0)>();
    else if (NS >= 3 && after >= 1) wait_vm<(NS >= 3 ? D : 0)>();
    else wait_vm<0>();
    lds_barrier();
    if (st + NS - 1 < nsteps) issue(st + NS - 1);
    const uint8_t* slot = smem + (st % NS) * STAGE;
    const uint8_t* tA = slot + (wm * 16 * TM) * SB;
    const uint8_t* tB = slot + BM * SB + (wn * 16 * TN) * SB;
    const uint32_t* ts = reinterpret_cast<const uint32_t*>(slot + SCOFF);
    const int sh = 8 * (st & 3);
    int sa[TM], sb[TN];
#pragma unroll
    for (int x = 0; x < TM; ++x)
      sa[x] = ASC ? (int)((ts[(wm * 16 * TM + 16 * x + l16) * 4 + q] >> sh) & 0xffu) : 127;
#pragma unroll
    for (int y = 0; y < TN; ++y)
      sb[y] = BSC ? (int)((ts[(wn * 16 * TN + 16 * y + l16) * 4 + q] >> sh) & 0xffu) : 127;
    // fp8: one 16x16x128 scaled MFMA per 128-byte step; bf16: the step's two
    // 64-byte halves (chunks q and 4 + q of each lane) feed two 16x16x32
    // MFMAs -- A and B use the same K assignment, so the sum is unchanged;
    // split-BF16: hi.hi + hi.lo + lo.hi, each as those two MFMAs.
    auto bf2 = [&](f4v& c, const i8v& a, const i8v& b) {
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, i4v{a[0], a[1], a[2], a[3]}),
                                                  __builtin_bit_cast(bf16x8, i4v{b[0], b[1], b[2], b[3]}), c, 0, 0, 0);
      c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, i4v{a[4], a[5], a[6], a[7]}),
                                                  __builtin_bit_cast(bf16x8, i4v{b[4], b[5], b[6], b[7]}), c, 0, 0, 0);
    };
    auto mma = [&](f4v& c, const i8v& a, const i8v& b, int sx, int sy) {
      if constexpr (FORM == kNxFp8) {
        c = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(a, b, c, 0, 0, 0, sx, 0, sy);
      } else {
        bf2(c, a, b);
      }
    };
    if constexpr (FORM == kNxBf16x3) {
      const uint8_t* tAl = tA + LOOFF;
      const uint8_t* tBl = tB + LOOFF;
      i8v ah[TM], al[TM];
#pragma unroll
      for (int x = 0; x < TM; ++x) {
        ah[x] = frag(tA, 16 * x + l16);
        al[x] = frag(tAl, 16 * x + l16);
      }
#pragma unroll
      for (int y = 0; y < TN; ++y) {
        const i8v bh = frag(tB, 16 * y + l16), bl = frag(tBl, 16 * y + l16);
#pragma unroll
        for (int x = 0; x < TM; ++x) {
          bf2(acc[x][y], ah[x], bh);
          bf2(acc[x][y], ah[x], bl);
          bf2(acc[x][y], al[x], bh);
        }
      }
    } else if constexpr (kScaleSel) {
      // A's scale dword of a row tile holds the bytes of four steps: read
      // from LDS once every four steps, the step's byte shifted out of the
      // register
      if ((st & 3) == 0) {
        // the sixteen-step record (row, q) in slot 0's scale area, this
        // four-step dword of it
        const uint32_t* sr = reinterpret_cast<const uint32_t*>(smem + SCOFF) + ((st & 15) >> 2);
#pragma unroll
        for (int x = 0; x < TM; ++x) sw[x] = (int)sr[((wm * 16 * TM + 16 * x + l16) * 4 + q) * 4];
        // read before the copy of the next sixteen steps may land (three
        // barriers later)
        asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
      }
      i8v bfr[TN];
#pragma unroll
      for (int y = 0; y < TN; ++y) bfr[y] = frag(tB, 16 * y + l16);
#pragma unroll
      for (int x = 0; x < TM; ++x) {
        const i8v af = frag(tA, 16 * x + l16);
        const int sx = (int)(((uint32_t)sw[x] >> sh) & 0xffu);
#pragma unroll
        for (int y = 0; y < TN; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(af, bfr[y], acc[x][y], 0, 0, 0, sx, 0, 127);
      }
    } else if constexpr (TM <= TN) {
      i8v af[TM];
#pragma unroll
      for (int x = 0; x < TM; ++x) af[x] = frag(tA, 16 * x + l16);
#pragma unroll
      for (int y = 0; y < TN; ++y) {
        const i8v bf = frag(tB, 16 * y + l16);
#pragma unroll
        for (int x = 0; x < TM; ++x) mma(acc[x][y], af[x], bf, sa[x], sb[y]);
      }
    } else {
      i8v bfr[TN];
#pragma unroll
      for (int y = 0; y < TN; ++y) bfr[y] = frag(tB, 16 * y + l16);
#pragma unroll
      for (int x = 0; x < TM; ++x) {
        const i8v af = frag(tA, 16 * x + l16);
#pragma unroll
        for (int y = 0; y < TN; ++y) mma(acc[x][y], af, bfr[y], sa[x], sb[y]);
      }
    }
  }
  nt_epilogue_at<TM, TN, FORM == kNxFp8 && !ASC, FORM == kNxFp8 && !BSC>(g, acc, m0 + wm * (16 * TM),
                                                                          n0 + wn * (16 * TN), lane);
}

// ---- fp8 window: quantise and left product in one pass ---------------------
// In the block-scaled fp8 form the left product x = nu x + eta Upsilon_w
// Q(A_w) uses L only as the B operand of its MFMAs.  This kernel quantises
// each 128-row x 64-column tile of the FP64 window into LDS and consumes it
// there: L never goes to HBM, and since a block owns 64 columns of x over all
// rows of the window the product needs no K split.  It also writes R and its
// scales for the right product, as k_quant_fp8_mx does, and its L blocks and
// scales are k_quant_fp8_mx's: each column of a step cut into four 32-row
// blocks, each scaled by 2^-e of its own max |x|.
// Block: 8 waves (two per SIMD) in two groups with separate registers and
// separate load counts.  Per 128-row step:
//  - waves 0-3 (tile waves) hold the tile in registers: wave w loads
//    columns w + 4 e (e < 16), two rows a lane (one contiguous 1 KB column
//    segment per instruction), one step ahead, so the next tile is in flight
//    for a whole step.  Every load is issued (addresses clamped into the
//    window, out-of-window values zeroed afterwards) and every thread stores
//    the same way (out-of-window R blocks go to a sink row), so the
//    compiler's counts agree on every path and a wait for the step's tile
//    never waits for the next one.  They write the tile to LDS as floats (t,
//    rows padded to 132 words) and, after a barrier, quantise it: each thread
//    two halves of L blocks (16 rows of a column: e4m3 bytes into the L tile,
//    128-byte rows with 16-byte chunks XOR-swizzled as in k_mx, and the scale
//    byte) and two halves of R blocks (16 columns of a row, straight to HBM),
//    the block maximum shared with the lane holding the other half.
//  - waves 4-7 (MFMA waves) hold the accumulators: wave 4 + m owns Upsilon
//    rows [16 TM m, 16 TM (m + 1)) against the tile's four 16-column groups
//    (tiles past the map's rows are computed on zero rows, never stored),
//    its Upsilon fragments copied to LDS by DMA one step ahead, between the
//    barriers (the window's slice of Upsilon, at most 448 x w bytes, stays
//    in L2), and the L tile read after the second barrier.  The L tile is double
//    buffered, so the tile waves quantise the next step while the MFMAs run.
// Blocks start at different steps (block b at step b mod the steps of a
// window) and wrap around, so the blocks in flight spread over the rows.
// max |x| over 16 values as float bits (non-negative floats order as
// unsigned integers)
__device__ __forceinline__ uint32_t abs_max_bits(const float (&v)[16]) {
  uint32_t m = 0;
#pragma unroll
  for (int c = 0; c < 16; ++c) {
    const uint32_t a = __float_as_uint(v[c]) & 0x7fffffffu;
    m = m > a ? m : a;
  }
  return m;
}
// max with lane l ^ 1 (DPP quad_perm [1, 0, 3, 2], every lane active)
__device__ __forceinline__ uint32_t pair_max_dpp(uint32_t m) {
  const uint32_t o = (uint32_t)__builtin_amdgcn_mov_dpp((int)m, 0xb1, 0xf, 0xf, false);
  return m > o ? m : o;
}
// max with lane l ^ 32
__device__ __forceinline__ uint32_t half_max_swap(uint32_t m) {
  const uint32_t o = (uint32_t)__shfl_xor((int)m, 32);
  return m > o ? m : o;
}
struct QlArgs {
  const double* A;
  int64_t lda, rows, n;
  uint8_t* R;
  int64_t ldr;
  uint8_t* Rs;
  int64_t ldrs;
  int64_t ldu;  // row length of the fp8 Upsilon (bytes, >= 256 past the last column)
  int64_t rstrip;  // > 0: R in 64-column strips of rstrip bytes ((wrows + 1) x 64)
  // Windows: the pass covers rows in windows of wrows rows (a multiple of
  // 128; the last may be shorter), window w's R and scales at R + w rwin and
  // Rs + w swin (rows of the window, the sink row wrows); x sums the
  // windows' products.  One window: rwin = 0 and wrows = rows.
  int64_t wrows, rwin, swin;
};
constexpr int kQlCols = 64, kQlRows = 128, kQlThreads = 512, kQlPad = 132;
// Rows of the fp8 Upsilon the fused kernel reads for range k: 4 MFMA waves
// of 16 TM rows (TM = 1, 2, 4, 7); 0 when k is beyond it.
__host__ __device__ constexpr int ql_tm(int64_t k) { return k <= 64 ? 1 : (k <= 128 ? 2 : (k <= 256 ? 4 : (k <= 448 ? 7 : 0))); }
template <int TM>
__global__ __launch_bounds__(kQlThreads, 1) void k_quant_left_fp8(NtArgs g, QlArgs q) {
  static_assert(kQlCols == 64 && kQlRows == 128 && kQlThreads == 512, "tile / block shape");
  constexpr int NE = kQlCols / 4;  // tile columns per tile wave
  // LDS: t[j][i] (the tile as floats), the e4m3 L tiles Lt[buf][j] (row j =
  // column j) and their scales sc[buf][j][b] (block b of column j).  After
  // the last step, su holds half of the block's x tile at a time (FP32,
  // column-major, ld kXRows) and t the Upsilon row scales.
  constexpr int kXRows = 4 * 16 * TM;
  __shared__ __attribute__((aligned(16))) float t[kQlCols][kQlPad];
  __shared__ __attribute__((aligned(16))) uint8_t Lt[2][kQlCols][128];
  __shared__ uint8_t sc[2][kQlCols][4];
  __shared__ __attribute__((aligned(16))) float su[kQlCols / 2 * kXRows];
  static_assert(kXRows * 8 <= (int)sizeof(t), "row scales fit in t");
  const int tid = threadIdx.x, lane = tid & 63, l16 = lane & 15, lq = lane >> 4;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t j0 = (int64_t)blockIdx.x * kQlCols;
  const int64_t rows = q.rows, n = q.n;
  // an even number of steps: a padding step (odd step counts) reads clamped
  // rows, zeroes them, writes no R row and adds zero products
  const int nsteps = (int)((((rows + kQlRows - 1) / kQlRows) + 1) & ~1);
  const int spw = q.rwin ? (int)(q.wrows / kQlRows) : nsteps;  // steps per window
  // start staggered over 32 steps (4096 rows) at most, so the blocks in
  // flight read one or two 4096-column slices of Upsilon (which stay in L2)
  const int sspan = spw < 32 ? spw : 32;
  const int s0 = (int)(blockIdx.x % (unsigned)(sspan < nsteps ? sspan : nsteps));
  auto step_of = [&](int i) { return (i + s0 < nsteps) ? i + s0 : i + s0 - nsteps; };

  f4v acc[TM][4];
#pragma unroll
  for (int x = 0; x < TM; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y) acc[x][y] = f4v{0.f, 0.f, 0.f, 0.f};

  if (wave < 4) {
    // ---- tile waves
    auto load_a = [&](int st, d2v (&v)[NE]) {
      const int64_t gi = (int64_t)st * kQlRows + 2 * lane;  // even; rows is even
      const uint32_t boff = (uint32_t)((gi < rows ? gi : rows - 2) * 8);
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        const int64_t gj = j0 + wave + 4 * e;
        const char* col = reinterpret_cast<const char*>(q.A + (gj < n ? gj : n - 1) * q.lda);
        v[e] = __builtin_nontemporal_load(reinterpret_cast<const d2v*>(col + boff));
      }
    };
    auto step = [&](int i, d2v (&cur)[NE], d2v (&nxt)[NE]) {
      const int st = step_of(i), sn = (st + 1 < nsteps) ? st + 1 : 0;
      load_a(sn, nxt);
      __builtin_amdgcn_sched_barrier(0);
      const bool rval = (int64_t)st * kQlRows + 2 * lane < rows;
#pragma unroll
      for (int e = 0; e < NE; ++e) {
        const int jt = wave + 4 * e;
        const bool ok = rval && j0 + jt < n;
        const float2 f = ok ? float2{(float)cur[e][0], (float)cur[e][1]} : float2{0.f, 0.f};
        *reinterpret_cast<float2*>(&t[jt][2 * lane]) = f;
      }
      // the tile's LDS writes land before the barrier (lds_barrier waits for
      // nothing; __syncthreads would also wait for the next tile's loads)
      asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
      lds_barrier();  // 1: t complete
      const int buf = i & 1;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        const int idx = tid + 256 * p;
        {
          // L: half h of block b (rows 32 b + 16 h .. + 15) of column j;
          // lanes 2 c and 2 c + 1 hold the two halves of one block
          const int h = idx & 1, j = (idx >> 1) & (kQlCols - 1), bb = idx >> 7;
          float v[16];
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const f4v x = *reinterpret_cast<const f4v*>(&t[j][32 * bb + 16 * h + 4 * c]);
            v[4 * c] = x[0]; v[4 * c + 1] = x[1]; v[4 * c + 2] = x[2]; v[4 * c + 3] = x[3];
          }
          const int ex = block_exponent(__uint_as_float(pair_max_dpp(abs_max_bits(v))));
          const float inv = ldexpf(1.f, -ex);
          i4v w;
#pragma unroll
          for (int c = 0; c < 4; ++c)
            w[c] = (int)pack4_e4m3(v[4 * c] * inv, v[4 * c + 1] * inv, v[4 * c + 2] * inv, v[4 * c + 3] * inv);
          *reinterpret_cast<i4v*>(&Lt[buf][j][16 * ((2 * bb + h) ^ ((j >> 1) & 7))]) = w;
          sc[buf][j][bb] = (uint8_t)(127 + ex);  // both halves store the same byte
        }
        __builtin_amdgcn_sched_barrier(0);  // one half block at a time: fewer live registers
        {
          // R: half h of the 32-column block cb = p (columns 32 p + 16 h ..
          // + 15 of the tile) of row ri; lanes l and l + 32 hold the halves
          const int h = lane >> 5, ri = wave * 32 + (lane & 31);
          float v[16];
#pragma unroll
          for (int c = 0; c < 16; ++c) v[c] = t[32 * p + 16 * h + c][ri];
          const int ex = block_exponent(__uint_as_float(half_max_swap(abs_max_bits(v))));
          const float inv = ldexpf(1.f, -ex);
          i4v w;
#pragma unroll
          for (int c = 0; c < 4; ++c)
            w[c] = (int)pack4_e4m3(v[4 * c] * inv, v[4 * c + 1] * inv, v[4 * c + 2] * inv, v[4 * c + 3] * inv);
          const int64_t gr = (int64_t)st * kQlRows + ri, gc = j0 + 32 * p, bk = gc / 32;
          const bool ok = gr < rows && gc < n;
          // the row within its window (out of range: the sink row wrows of
          // window 0; a padding step may lie past the last window)
          const int wi = ok ? st / spw : 0;
          const int64_t lr = ok ? gr - (int64_t)wi * q.wrows : q.wrows;
          uint8_t* rw = q.R + wi * q.rwin;
          // R in rows (ld ldr) or in 64-byte column strips (a block's strip
          // is its 64 columns: whole cache lines per step)
          uint8_t* rp = q.rstrip ? rw + blockIdx.x * q.rstrip + lr * 64 + 32 * p + 16 * h
                                 : rw + lr * q.ldr + (ok ? gc + 16 * h : 0);
          uint8_t* sp = q.Rs + wi * q.swin + lr * q.ldrs + (ok ? (bk & 3) * (q.ldrs / 4) + (bk >> 2) : 0);
          *reinterpret_cast<i4v*>(rp) = w;
          *sp = (uint8_t)(127 + ex);  // both halves store the same byte
        }
        __builtin_amdgcn_sched_barrier(0);
      }
      asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
      lds_barrier();  // 2: L tile and scales complete; t free
    };
    d2v va[NE], vb[NE];
    load_a(step_of(0), va);
    // The stores of a step (two 16-byte R pieces and two scale bytes a
    // thread) sit between a tile's loads and the wait for them; the same
    // stores here, to the sink row, give the first step the same count, so
    // the compiler's wait (the minimum over the loop's entries) does not make
    // every step wait for the previous step's stores.
    {
      uint8_t* rp = q.rstrip ? q.R + blockIdx.x * q.rstrip + q.wrows * 64 : q.R + q.wrows * q.ldr;
      uint8_t* sp = q.Rs + q.wrows * q.ldrs;
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        *reinterpret_cast<i4v*>(rp + 16 * p) = i4v{0, 0, 0, 0};
        sp[p] = (uint8_t)127;
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    for (int i = 0; i < nsteps; i += 2) {
      step(i, va, vb);
      step(i + 1, vb, va);
    }
  } else {
    // ---- MFMA waves: the map holds 64 TM rows (zero rows past M), so no
    // row is clamped.  Each wave copies its fragment rows of the next step
    // into its own region of su by DMA (global_load_lds_dwordx4) as soon as
    // it has read this step's into registers: the copy has a whole step to
    // land (it queues behind the tile waves' loads), holds no registers, and
    // needs no barrier (a wave reads only what it copied).  Chunk c of row r
    // sits at position c ^ ((r >> 1) & 7), so the fragment reads of 16
    // consecutive rows at one chunk touch 16 distinct bank groups; one DMA
    // instruction fills 8 rows, lane l landing at row l / 8, position l % 8
    // and fetching the chunk that belongs there.
    const int m = wave - 4;
    constexpr int kURows = 16 * TM;
    static_assert(4 * kURows * 128 <= (int)sizeof(su), "fragment rows fit in su");
    uint8_t* ureg = reinterpret_cast<uint8_t*>(su) + m * kURows * 128;
    const uint32_t urow = (uint32_t)(((int64_t)m * kURows + (lane >> 3)) * q.ldu);
    const uint32_t uev = urow + 16 * ((lane & 7) ^ ((lane >> 4) & 7));
    const uint32_t uod = urow + 16 * ((lane & 7) ^ ((4 + (lane >> 4)) & 7));
    auto issue_u = [&](int st) {
      const uint8_t* base = static_cast<const uint8_t*>(g.ah) + g.koff_a + (int64_t)st * kQlRows;
#pragma unroll
      for (int d = 0; d < 2 * TM; ++d)
        __builtin_amdgcn_global_load_lds((const void*)(base + 8 * d * q.ldu + ((d & 1) ? uod : uev)),
                                         (SKEMA_LDS void*)(ureg + 1024 * d), 16, 0, 0);
    };
    // swizzled byte offsets of this lane's two chunks (lq and 4 + lq) in a
    // row of the L tile or of the fragment rows (rows 16 y + l16 all share
    // the swizzle (l16 >> 1) & 7)
    const int so0 = 16 * (lq ^ ((l16 >> 1) & 7)), so1 = so0 ^ 64;
    issue_u(step_of(0));
    for (int i = 0; i < nsteps; ++i) {
      const int buf = i & 1;
      lds_barrier();  // 1
      // While the tile waves quantise: this step's fragments into registers
      // (their copy was issued a step ago), then the next step's copy, whose
      // issue may wait behind the tile waves' loads; after barrier 2 only the
      // L tile reads and the MFMAs remain, short enough not to hold the tile
      // waves at the next barrier 1.
      wait_vm<0>();
      i8v u[TM];
#pragma unroll
      for (int x = 0; x < TM; ++x) {
        const i4v lo = *reinterpret_cast<const i4v*>(ureg + (16 * x + l16) * 128 + so0);
        const i4v hi = *reinterpret_cast<const i4v*>(ureg + (16 * x + l16) * 128 + so1);
        u[x] = i8v{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
      }
      // the fragments are in registers before the next copy lands in ureg
      asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
      issue_u(step_of(i + 1 < nsteps ? i + 1 : 0));
      lds_barrier();  // 2: the step's L tile is complete
#pragma unroll
      for (int y = 0; y < 4; ++y) {
        const int col = 16 * y + l16;
        const i4v lo = *reinterpret_cast<const i4v*>(&Lt[buf][col][so0]);
        const i4v hi = *reinterpret_cast<const i4v*>(&Lt[buf][col][so1]);
        const i8v bf = i8v{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
        const int sb = sc[buf][col][lq];
#pragma unroll
        for (int x = 0; x < TM; ++x)
          acc[x][y] = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(u[x], bf, acc[x][y], 0, 0, 0, 127, 0, sb);
      }
    }
    wait_vm<0>();  // the last (unused) copy lands before su is reused
  }

  // Epilogue: the block's columns of x are one contiguous range of HBM (x is
  // k x n with ld k = M), so the tile goes through LDS and back out in
  // coalesced read-modify-write passes by all waves, x = beta x + alpha
  // 2^e(row) acc (the row's Upsilon scale), two elements (16 bytes) a thread
  // and round; 32 columns a pass.
  __syncthreads();
  float* xs = su;
  double* rsc = reinterpret_cast<double*>(&t[0][0]);
  const int64_t M = g.M;
  for (int i = tid; i < M; i += kQlThreads) rsc[i] = ldexp(1.0, block_exponent(g.amax_a[i]));
  const int64_t ncols = n - j0 < kQlCols ? n - j0 : kQlCols;
  const double alpha = g.out.alpha1, beta = g.out.beta1;
  constexpr int64_t kStride = 2 * kQlThreads;
  constexpr int kEB = 8;
  const int64_t di = kStride % M, dj = kStride / M;
#pragma unroll
  for (int hc = 0; hc < 2; ++hc) {
    if (hc) __syncthreads();  // the first half's reads of xs are done
    // xs[j][i] = column 32 hc + j, ld kXRows: rows past M land in unused slots
    if (wave >= 4) {
#pragma unroll
      for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
          *reinterpret_cast<f4v*>(&xs[(16 * y + l16) * kXRows + (wave - 4) * 16 * TM + 16 * x + 4 * lq]) =
              acc[x][2 * hc + y];
    }
    __syncthreads();
    const int64_t len = (ncols - 32 * hc < 32 ? ncols - 32 * hc : 32) * M;  // elements (<= 0: none)
    double* xo = g.out.C1 + (j0 + 32 * hc) * g.out.ldc1;
    // element e = jl M + i of the range is xs[jl][i]; rounds of kEB pairs a
    // thread, the reads of x of a round all issued before the first write
    int64_t i = (2 * (int64_t)tid) % M, jl = (2 * (int64_t)tid) / M;
    for (int64_t e0 = 2 * (int64_t)tid; e0 + 1 < len; e0 += kEB * kStride) {
      d2v old[kEB];
#pragma unroll
      for (int b = 0; b < kEB; ++b) {
        const int64_t e = e0 + b * kStride;
        old[b] = (beta != 0.0 && e + 1 < len) ? *reinterpret_cast<const d2v*>(xo + e) : d2v{0.0, 0.0};
      }
#pragma unroll
      for (int b = 0; b < kEB; ++b) {
        const int64_t e = e0 + b * kStride;
        if (e + 1 < len) {
          // the pair's second element: next row, or row 0 of the next column
          const bool wrap = i + 1 == M;
          const double a0 = (double)xs[jl * kXRows + i], a1 = (double)xs[(jl + wrap) * kXRows + (wrap ? 0 : i + 1)];
          const double v0 = alpha * a0 * rsc[i], v1 = alpha * a1 * rsc[wrap ? 0 : i + 1];
          *reinterpret_cast<d2v*>(xo + e) = d2v{v0 + beta * old[b][0], v1 + beta * old[b][1]};
        }
        i += di;
        jl += dj;
        if (i >= M) { i -= M; ++jl; }
      }
    }
    if (len > 0 && (len & 1) && tid == 0) {
      const int64_t e = len - 1, ie = e % M;
      const double v = alpha * (double)xs[(e / M) * kXRows + ie] * rsc[ie];
      xo[e] = (beta != 0.0) ? v + beta * xo[e] : v;
    }
  }
}

// Sum of the FP32 partials of the K splits (in FP64, fixed order), then the
// epilogue.
// M a multiple of 4: one thread per four consecutive rows of a column
// (16-byte partial loads, one 32-byte read-modify-write of the state);
// otherwise one thread per element (consecutive threads, consecutive rows).
__global__ void k_nt_reduce(NtArgs g, int splits) {
  const int64_t total = g.M * g.N;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  if ((g.M & 3) == 0) {
    const int64_t qm = g.M / 4, quads = qm * g.N;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < quads; e += stride) {
      const int64_t j = e / qm, i0 = 4 * (e % qm);
      double v[4] = {0.0, 0.0, 0.0, 0.0};
      const float* p = g.part + j * g.M + i0;
      for (int z = 0; z < splits; ++z, p += total) {
        const float4 f = *reinterpret_cast<const float4*>(p);
        v[0] += (double)f.x; v[1] += (double)f.y; v[2] += (double)f.z; v[3] += (double)f.w;
      }
      store_nt4(g, i0, j, v);
    }
    return;
  }
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
    double v = 0.0;
    for (int z = 0; z < splits; ++z) v += (double)g.part[(int64_t)z * total + e];
    store_nt(g, e % g.M, e / g.M, v);
  }
}

// Tile and K-split of one NT product: measured choices for the three window
// products, otherwise a model over 64/128 x 64/128 tiles (a launch as rounds
// of `res` co-resident blocks per CU over 256 CUs, each round costing
// BM * BN * K/splits, plus FP32 partial traffic).
// tm > 0 selects the wide kernel k_ntx<form, tm, tn, ...> (BM = 64 tm,
// BN = 16 tn).
struct NtPlan {
  int bm, bn, splits;
  int tm = 0, tn = 0;
  int ns = 0;  // > 0: direct-to-LDS kernel k_mx (wave tile 16 tm x 16 tn, ns LDS slots)
  int wm = 2, wn = 2;  // k_mx waves along M and N (BM = 16 tm wm, BN = 16 tn wn)
};
NtPlan plan_ntx(int prec, int64_t M, int64_t N, int64_t K) {
  // wide-kernel tiles with the blocks per CU their registers admit
  struct C { int tm, tn, res; };
  const C cands[] = {{4, 8, 1}, {4, 4, 2}, {2, 8, 2}, {2, 4, 3}};
  const int kstep = prec == SKEMA_PREC_FP8 ? 128 : 64;
  NtPlan best{0, 0, 1, 2, 4};
  double best_cost = 1e300;
  for (C c : cands) {
    const int64_t bm = 64 * c.tm, bn = 16 * c.tn;
    const int64_t tiles = cdiv(M, bm) * cdiv(N, bn);
    for (int sp = 1; sp <= 8; sp *= 2) {
      if (sp > 1 && K / sp < 8 * kstep) break;
      const double rounds = (double)cdiv(tiles * sp, 256 * c.res);
      const double cost = rounds * (double)(bm * bn) * (double)cdiv(K, sp) * c.res +
                          (sp > 1 ? 8.0 * sp * (double)M * N : 0.0);
      if (cost < best_cost) { best_cost = cost; best = NtPlan{0, 0, sp, c.tm, c.tn}; }
    }
  }
  return best;
}

NtPlan plan_nt_2x2(int prec, int64_t M, int64_t N, int64_t K);
// Shapes of the window products:
//   right  [y | G] = A_w [Omega; Psi]^T : M = w, N = k+s, K = n (long)
//   left   x = Upsilon_w A_w            : M = k, N = n, K = w
//   core   z = Phi_w G                  : M = N = s, K = w
NtPlan plan_nt(int prec, int64_t M, int64_t N, int64_t K) {
  if (prec == SKEMA_PREC_FP8 || prec == SKEMA_PREC_BF16) {
    // measured per window product (w = 4096, n = 16384, k = 401, s = 803):
    // fp8 right and left on the 8-wave direct-to-LDS kernel (right 256 x 256
    // tile, 2 slots, K split 3: 0.111 ms + 0.018 ms split sum; left 448 x 128,
    // one tile over all k <= 448 rows, 2 slots, K split 2: 0.073 ms + its
    // sum, 2% per window ahead of 256 x 256 and of 448 x 128 on 4 x 2 waves;
    // 16 waves, K split 4, 128 x 256 tiles all slower), bf16
    // right on the 4-wave one (4x8 wave tiles, 3 slots, split 3; 0.26 ms) and
    // left on the wide kernel (0.15 ms), core (bf16) on the 4-wave direct-to-
    // LDS kernel (4x4, 4 slots, split 4; 0.032 ms); otherwise the round model
    if (M >= 2048 && N <= 2048 && K >= 8192) return NtPlan{0, 0, 3, 8, 4, 2, 2, 4};
    if (M <= 512 && N >= 4096 && K >= 2048)
      return prec == SKEMA_PREC_FP8 ? NtPlan{0, 0, 2, 14, 2, 2, 2, 4} : NtPlan{0, 0, 1, 4, 4, 3, 4, 2};
    if (M <= 1024 && N <= 1024 && K >= 2048) return NtPlan{0, 0, 4, 4, 4, 4};
    return plan_ntx(prec, M, N, K);
  }
  // split-BF16: the direct-to-LDS kernel with both planes in its ring
  // (128 x 128 tiles, 2 x 2 waves, 2 slots)
  if (prec == SKEMA_PREC_BF16X3) {
    if (M >= 2048 && N <= 2048 && K >= 8192) return NtPlan{0, 0, 3, 4, 4, 2};
    if (M <= 512 && N >= 4096 && K >= 2048) return NtPlan{0, 0, 2, 4, 4, 2};
    if (M <= 1024 && N <= 1024 && K >= 2048) return NtPlan{0, 0, 4, 4, 4, 2};
  }
  return plan_nt_2x2(prec, M, N, K);
}

// The 2 x 2-wave kernels (k_nt_*): every form, any alignment; the tile and
// split measured fastest for each window product at w = 4096, n = 16384,
// k = 401, s = 803, otherwise a round model.
NtPlan plan_nt_2x2(int prec, int64_t M, int64_t N, int64_t K) {
  const int ks = prec == SKEMA_PREC_FP8 ? 128 : (prec == SKEMA_PREC_BF16 ? 64 : 32);
  if (M >= 2048 && N <= 2048 && K >= 8192) {
    if (prec == SKEMA_PREC_FP8) return {128, 64, 1};
    if (prec == SKEMA_PREC_FP32) return {128, 128, 4};
    return {64, 64, 2};
  }
  if (M <= 512 && N >= 4096 && K >= 16 * ks) {
    if (prec == SKEMA_PREC_BF16) return {128, 64, 1};
    if (prec == SKEMA_PREC_FP8) return {64, 64, 1};
    return {64, 64, 2};
  }
  if (M <= 1024 && N <= 1024 && K >= 2048) {
    if (prec == SKEMA_PREC_BF16X3) return {64, 64, 2};
    return {128, 128, 4};
  }
  const NtPlan cands[] = {{128, 128, 1}, {128, 64, 1}, {64, 128, 1}, {64, 64, 1}};
  NtPlan best = cands[3];
  double best_cost = 1e300;
  const int kstep = prec == SKEMA_PREC_FP8 ? 128 : (prec == SKEMA_PREC_BF16 ? 64 : 32);
  for (NtPlan c : cands) {
    const int64_t tiles = cdiv(M, c.bm) * cdiv(N, c.bn);
    const int res = (c.bm * c.bn >= 128 * 128) ? 2 : 4;  // blocks per CU the LDS admits
    for (int sp = 1; sp <= 8; sp *= 2) {
      if (K / sp < 4 * kstep) break;
      const double rounds = (double)cdiv(tiles * sp, 256 * res);
      const double cost = rounds * c.bm * c.bn * (double)cdiv(K, sp) * res +
                          (sp > 1 ? 8.0 * sp * (double)M * N : 0.0);
      if (cost < best_cost) { best_cost = cost; best = c; best.splits = sp; }
    }
  }
  return best;
}

template <int BM, int BN>
void launch_nt_tile(hipStream_t st, int prec, const NtArgs& a, dim3 grid) {
  if (prec == SKEMA_PREC_FP8) hipLaunchKernelGGL((k_nt_fp8<BM, BN>), grid, dim3(kThreads), 0, st, a);
  else if (prec == SKEMA_PREC_FP32) hipLaunchKernelGGL((k_nt_f32<BM, BN>), grid, dim3(kThreads), 0, st, a);
  else if (prec == SKEMA_PREC_BF16X3) hipLaunchKernelGGL((k_nt_bf16<BM, BN, true>), grid, dim3(kThreads), 0, st, a);
  else hipLaunchKernelGGL((k_nt_bf16<BM, BN, false>), grid, dim3(kThreads), 0, st, a);
}
template <int TM, int TN>
void launch_ntx_tile(hipStream_t st, int prec, const NtArgs& a, dim3 grid) {
  if (prec == SKEMA_PREC_FP8) {
    if (a.asc) hipLaunchKernelGGL((k_ntx<kNxFp8, TM, TN, true, false>), grid, dim3(kThreads), 0, st, a);
    else if (a.bsc) hipLaunchKernelGGL((k_ntx<kNxFp8, TM, TN, false, true>), grid, dim3(kThreads), 0, st, a);
    else hipLaunchKernelGGL((k_ntx<kNxFp8, TM, TN, false, false>), grid, dim3(kThreads), 0, st, a);
  } else {
    hipLaunchKernelGGL((k_ntx<kNxBf16, TM, TN, false, false>), grid, dim3(kThreads), 0, st, a);
  }
}
template <int FORM, int TM, int TN, int NS, bool ASC, bool BSC, int WM = 2, int WN = 2, bool SEG = false>
void launch_mx_t(hipStream_t st, const NtArgs& a, dim3 grid) {
  hipLaunchKernelGGL((k_mx<FORM, WM, WN, TM, TN, NS, ASC, BSC, SEG>), grid, dim3(64 * WM * WN), 0, st, a);
}
// k_mx variants built (the ones the plans select; launch_mx returns false
// for any other combination):
//   fp8, A scales (right product):  2 x 4 waves of 8 x 4 (256 x 256), 2 slots
//   fp8, B scales (left product):   2 x 4 waves of 14 x 2 (448 x 128), 2 slots
//   bf16 right: 2 x 4 waves of 8 x 4, 2 slots; bf16 left: 4 x 2 waves of 4 x 4
//   (256 x 128), 3 slots; bf16 core: 2 x 2 waves of 4 x 4 (128 x 128), 4 slots
//   split-BF16 (both planes): 2 x 2 waves of 4 x 4 (128 x 128), 2 slots
bool launch_mx(hipStream_t st, int prec, const NtArgs& a, const NtPlan& pl, dim3 grid, bool dry = false) {
  const int key = pl.tm * 100 + pl.tn * 10 + pl.ns;
  const bool w24 = pl.wm == 2 && pl.wn == 4, w42 = pl.wm == 4 && pl.wn == 2, w22 = pl.wm == 2 && pl.wn == 2;
  if (prec == SKEMA_PREC_FP8 && a.asc && w24 && key == 842) {
    if (!dry) launch_mx_t<kNxFp8, 8, 4, 2, true, false, 2, 4>(st, a, grid);
    return true;
  }
  if (prec == SKEMA_PREC_FP8 && a.bsc && w24 && key == 1422) {
    if (!dry) launch_mx_t<kNxFp8, 14, 2, 2, false, true, 2, 4>(st, a, grid);
    return true;
  }
  if (prec == SKEMA_PREC_BF16X3 && !a.asc && !a.bsc && w22 && key == 442) {
    if (!dry) launch_mx_t<kNxBf16x3, 4, 4, 2, false, false, 2, 2>(st, a, grid);
    return true;
  }
  if (prec == SKEMA_PREC_BF16 && !a.asc && !a.bsc) {
    if (w24 && key == 842) {
      if (!dry) {
        if (a.seg_steps) launch_mx_t<kNxBf16, 8, 4, 2, false, false, 2, 4, true>(st, a, grid);
        else launch_mx_t<kNxBf16, 8, 4, 2, false, false, 2, 4>(st, a, grid);
      }
      return true;
    }
    if (w42 && key == 443) {
      if (!dry) launch_mx_t<kNxBf16, 4, 4, 3, false, false, 4, 2>(st, a, grid);
      return true;
    }
    if (w22 && key == 444) {
      if (!dry) launch_mx_t<kNxBf16, 4, 4, 4, false, false, 2, 2>(st, a, grid);
      return true;
    }
  }
  return false;
}
// The wide kernels need 16-byte aligned rows of both operands.
bool ntx_eligible(int prec, const NtArgs& a) {
  if (prec != SKEMA_PREC_FP8 && prec != SKEMA_PREC_BF16 && prec != SKEMA_PREC_BF16X3) return false;
  const int64_t es = prec == SKEMA_PREC_FP8 ? 1 : 2;
  const bool lo = prec != SKEMA_PREC_BF16X3 || (a.al && a.bl && ((uintptr_t)a.al % 16) == 0 && ((uintptr_t)a.bl % 16) == 0);
  return lo && ((uintptr_t)a.ah % 16) == 0 && ((uintptr_t)a.bh % 16) == 0 && (a.lda * es) % 16 == 0 &&
         (a.ldb * es) % 16 == 0 && (a.koff_a * es) % 16 == 0;
}
void launch_nt(hipStream_t st, int prec, const NtArgs& a, const NtPlan& pl, dim3 grid) {
  if (pl.ns > 0) {
    SKEMA_REQUIRE(launch_mx(st, prec, a, pl, grid), kErrUnsupported, "no direct-to-LDS kernel for this tile");
    return;
  }
  if (pl.tm > 0) {
    if (pl.tm == 4 && pl.tn == 8) launch_ntx_tile<4, 8>(st, prec, a, grid);
    else if (pl.tm == 4 && pl.tn == 4) launch_ntx_tile<4, 4>(st, prec, a, grid);
    else if (pl.tm == 2 && pl.tn == 8) launch_ntx_tile<2, 8>(st, prec, a, grid);
    else if (pl.tm == 2 && pl.tn == 4) launch_ntx_tile<2, 4>(st, prec, a, grid);
    else SKEMA_REQUIRE(false, kErrUnsupported, "no wide NT kernel for this tile");
    return;
  }
  if (pl.bm == 128 && pl.bn == 128) launch_nt_tile<128, 128>(st, prec, a, grid);
  else if (pl.bm == 128) launch_nt_tile<128, 64>(st, prec, a, grid);
  else if (pl.bn == 128) launch_nt_tile<64, 128>(st, prec, a, grid);
  else launch_nt_tile<64, 64>(st, prec, a, grid);
}

// Whether gemm_nt runs this product on the direct-to-LDS kernel k_mx (the
// one consumer of A in 64-byte column strips).
bool nt_picks_mx(int prec, const NtArgs& a) {
  NtPlan pl = plan_nt(prec, a.M, a.N, a.K);
  if ((a.asc || a.bsc) && pl.tm == 0) return false;
  const int kstep_mx = prec == SKEMA_PREC_FP8 ? 128 : 64;
  return pl.ns > 0 && ntx_eligible(prec, a) && a.K % kstep_mx == 0 && launch_mx(0, prec, a, pl, dim3(), true);
}

// t (optional): a timing scope closed right after the product kernel, so
// it covers that kernel and not the split-K reduction that follows.
// ws = nullptr: no K split.  fill: as many K splits as fill the 256 CUs
// once with the tiles (at least the plan's; for a long K over few tiles).
// force: this plan instead of plan_nt's.
void gemm_nt(hipStream_t st, int prec, NtArgs a, DevBuf* ws, Timed* t = nullptr, bool fill = false,
             const NtPlan* force = nullptr) {
  SKEMA_REQUIRE(!a.a_strip || nt_picks_mx(prec, a), kErrUnsupported, "A in column strips needs the k_mx kernel");
  NtPlan pl = force ? *force : plan_nt(prec, a.M, a.N, a.K);
  SKEMA_REQUIRE(!(a.asc && a.bsc), kErrUnsupported, "block scales on both NT operands");
  if ((a.asc || a.bsc) && pl.tm == 0) pl = plan_ntx(prec, a.M, a.N, a.K);
  SKEMA_REQUIRE(!(a.asc || a.bsc) || (ntx_eligible(prec, a) && (!a.asc || a.koff_a == 0)), kErrUnsupported,
                "block-scaled fp8 operands need 16-byte aligned rows (and no K offset on A)");
  // k_mx: 16-byte aligned rows, whole 128-byte K steps, a built variant;
  // otherwise the wide kernel, or the 2 x 2-wave one for split-BF16 (the
  // wide kernel has no lo planes)
  const int kstep_mx = prec == SKEMA_PREC_FP8 ? 128 : 64;
  if (pl.ns > 0 && !(ntx_eligible(prec, a) && a.K % kstep_mx == 0 && launch_mx(st, prec, a, pl, dim3(), true)))
    pl = prec == SKEMA_PREC_BF16X3 ? plan_nt_2x2(prec, a.M, a.N, a.K) : plan_ntx(prec, a.M, a.N, a.K);
  const int kstep = pl.ns > 0 ? kstep_mx : (prec == SKEMA_PREC_FP8 ? 128 : (prec == SKEMA_PREC_BF16 ? 64 : 32));
  if (pl.tm > 0 && !ntx_eligible(prec, a)) pl = NtPlan{64, 64, pl.splits};
  if (pl.ns > 0) { pl.bm = 16 * pl.tm * pl.wm; pl.bn = 16 * pl.tn * pl.wn; }
  else if (pl.tm > 0) { pl.bm = 64 * pl.tm; pl.bn = 16 * pl.tn; }
  int splits = ws ? pl.splits : 1;
  // a large output (the right product over a whole matrix) with enough
  // tiles to fill the CUs four times over: no K split (its FP32 partials
  // would be the output's size times the splits)
  if (a.M * a.N >= ((int64_t)1 << 24) && cdiv(a.M, pl.bm) * cdiv(a.N, pl.bn) >= 1024) splits = 1;
  if (ws && fill) {
    const int64_t tiles = cdiv(a.M, pl.bm) * cdiv(a.N, pl.bn);
    splits = (int)std::max<int64_t>(splits, std::min<int64_t>(256 / std::max<int64_t>(tiles, 1), a.K / (8 * kstep)));
  }
  // block-scaled operands: splits start on a 4-step boundary (one scale
  // dword), 16 for A scales (k_mx copies sixteen steps' bytes at once)
  const int64_t kgrain = a.asc ? 16 * kstep : (a.bsc ? 4 * kstep : kstep);
  a.k_per_split = cdiv(cdiv(a.K, splits), kgrain) * kgrain;
  splits = (int)cdiv(a.K, a.k_per_split);
  a.part = nullptr;
  if (splits > 1) {
    ws->reserve((size_t)splits * a.M * a.N * sizeof(float));
    a.part = ws->as<float>();
  }
  a.tiles_m = (int)cdiv(a.M, pl.bm);
  // wide kernels: keep the larger operand's rows within an XCD
  a.n_fast = (a.M >= a.N) ? 1 : 0;
  SKEMA_REQUIRE(!a.seg_steps || (prec == SKEMA_PREC_BF16 && pl.ns == 2 && pl.tm == 8 && pl.tn == 4 && pl.wm == 2 &&
                                 pl.wn == 4),
                kErrUnsupported, "K segments need the 256 x 256 bf16 k_mx");
  const dim3 grid((unsigned)(a.tiles_m * cdiv(a.N, pl.bn)), 1, (unsigned)splits);
  launch_nt(st, prec, a, pl, grid);
  check_launch("k_nt");
  if (t) t->stop();
  if (splits > 1) {
    const int64_t items = (a.M % 4 == 0) ? a.M / 4 * a.N : a.M * a.N;
    const int blocks = (int)std::min<int64_t>(cdiv(items, 256), 8192);
    hipLaunchKernelGGL(k_nt_reduce, dim3(blocks), dim3(256), 0, st, a, splits);
    check_launch("k_nt_reduce");
  }
}

unsigned grid1(int64_t e) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>(cdiv(e, 256), 65536)); }

}  // namespace

// Rows of the fp8 Upsilon: k rounded up to the rows the fused quantise +
// left kernel covers (4 waves of 16 TM rows), the extra rows zero; the row
// maxima (floats) follow at this byte offset.
static int64_t fp8_ups_rows(int64_t k) { return ql_tm(k) ? 64 * ql_tm(k) : k; }
static size_t fp8_ups_amax_offset(const Sketch& sk) { return align_up((size_t)fp8_ups_rows(sk.k) * sk.lp_ldups, 256); }

void prepare_lowp_maps(Sketch& sk) {
  const int prec = sk.p.precision;
  SKEMA_REQUIRE(prec == SKEMA_PREC_BF16 || prec == SKEMA_PREC_BF16X3 || prec == SKEMA_PREC_FP8 ||
                    prec == SKEMA_PREC_FP32,
                kErrUnsupported, "unknown reduced-precision form");
  hipStream_t st = sk.ctx->stream;
  const int64_t k = sk.k, s = sk.s, m = sk.p.m, n = sk.p.n;
  // Row-major reduced maps, K contiguous, leading dimension padded to 128.
  // two spare 128-element steps past the last column: the fused fp8
  // quantise + left product reads whole steps of Upsilon (and one padding
  // step) past the window's end
  sk.lp_ldups = cdiv(m, 128) * 128 + 256;
  sk.lp_ldomps = cdiv(n, 128) * 128;
  const size_t ue = (size_t)k * sk.lp_ldups, oe = (size_t)(k + s) * sk.lp_ldomps;
  if (prec == SKEMA_PREC_FP8) {
    // e4m3 rows followed by one float (row max |x|) per row; Upsilon's rows
    // padded with zero rows to the fused kernel's row count
    const size_t uo = fp8_ups_amax_offset(sk);
    sk.lp_ups.alloc(uo + (size_t)k * sizeof(float));
    sk.lp_omps.alloc(align_up(oe, 256) + (size_t)(k + s) * sizeof(float));
    if (uo > ue) fill_i32(st, (int*)(sk.lp_ups.as<uint8_t>() + ue), (int64_t)((uo - ue) / 4), 0);
    hipLaunchKernelGGL(k_convert_map_fp8, dim3((unsigned)k), dim3(kThreads), 0, st, sk.upsL, sk.ldups, k, m,
                       sk.lp_ups.as<uint8_t>(), sk.lp_ldups, (float*)(sk.lp_ups.as<uint8_t>() + uo));
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
  // Phi (s x m) row-major for the core product: fp32 for the fp32 form,
  // bf16 hi + lo for split-BF16, bf16 otherwise.
  if (prec == SKEMA_PREC_FP32) {
    sk.lp_phi.alloc((size_t)s * sk.lp_ldups * 4);
    hipLaunchKernelGGL(k_convert_map<float>, dim3(grid1(s * m)), dim3(256), 0, st, sk.phiL, sk.ldphi, s, m,
                       sk.lp_phi.as<float>(), (float*)nullptr, sk.lp_ldups);
  } else {
    const bool split = prec == SKEMA_PREC_BF16X3;
    const size_t pe = (size_t)s * sk.lp_ldups;
    sk.lp_phi.alloc(pe * 2 * (split ? 2 : 1));
    __bf16* ph = sk.lp_phi.as<__bf16>();
    hipLaunchKernelGGL(k_convert_map<__bf16>, dim3(grid1(s * m)), dim3(256), 0, st, sk.phiL, sk.ldphi, s, m, ph,
                       split ? ph + pe : nullptr, sk.lp_ldups);
  }
  check_launch("k_convert_map (Phi)");
  SKEMA_HIP(hipStreamSynchronize(st));
}

namespace {

// Window scratch layout: L (column-major, ld ldl = rows padded), R
// (row-major, ld ldr = n padded), for fp8 the row and column max |x| of the
// window (floats) or the E8M0 block scales, and G (rows x s, column-major,
// ld ldl) in the core product's element type.
struct WinLayout {
  bool fp8, split, f32, mx;
  int core_prec, gtype;
  int64_t ldl, ldr, ldls, ldrs;
  size_t esz, planes, gsz, le, re, ge;
  size_t o_l, o_r, o_rmax, o_cmax, o_g, o_ls, o_rs, total;
};
WinLayout win_layout(const Sketch& sk, int64_t row0, int64_t rows) {
  WinLayout w{};
  const int prec = sk.p.precision;
  const int64_t n = sk.p.n, s = sk.s;
  w.fp8 = prec == SKEMA_PREC_FP8;
  w.split = prec == SKEMA_PREC_BF16X3;
  w.f32 = prec == SKEMA_PREC_FP32;
  // fp8 windows that start on a 16-row boundary (every window of a multiple-
  // of-16 window size) take the block-scaled form: one read of A_w, E8M0
  // scales per 32-element block of L's columns and R's rows; others take the
  // two-pass form with one scale per row and per column.
  w.mx = w.fp8 && (row0 % 16) == 0;
  w.core_prec = w.f32 ? SKEMA_PREC_FP32 : (w.split ? SKEMA_PREC_BF16X3 : SKEMA_PREC_BF16);
  w.gtype = w.f32 ? 3 : (w.split ? 2 : 1);
  w.gsz = w.f32 ? 4 : 2;
  w.ldl = cdiv(rows, 128) * 128;
  w.ldr = cdiv(n, 128) * 128;
  w.esz = w.fp8 ? 1 : (w.f32 ? 4 : 2);
  w.planes = w.split ? 2 : 1;
  // scale rows: 4 groups of one byte per step of 128 elements, padded to
  // 4 bytes (L) or 16 (R: k_mx copies sixteen steps of R's scales at once)
  w.ldls = 4 * align_up((size_t)cdiv(rows, 128), 4);
  w.ldrs = 4 * align_up((size_t)cdiv(n, 128), 16);
  w.le = (size_t)n * w.ldl;
  w.re = (size_t)(rows + 1) * w.ldr;  // one spare row: the sink of k_quant_left_fp8's out-of-window stores
  w.ge = (size_t)s * w.ldl;
  w.o_l = 0;
  w.o_r = align_up(w.le * w.esz * w.planes, 256);
  w.o_rmax = w.o_r + align_up(w.re * w.esz * w.planes, 256);
  w.o_cmax = w.o_rmax + (w.fp8 ? align_up((size_t)rows * 4, 256) : 0);
  w.o_g = w.o_cmax + (w.fp8 ? align_up((size_t)n * 4, 256) : 0);
  w.total = align_up(w.o_g + align_up(w.ge * w.gsz * w.planes, 256), 256);
  // block scales (mx) in sk.lp_scales: L's at 0, R's at o_rs
  w.o_ls = 0;
  w.o_rs = w.mx ? align_up((size_t)n * w.ldls, 256) : 0;
  return w;
}

// Block scales of a window (the block-scaled fp8 form).  Unit scales (0x7f)
// wherever no block writes one (K padding past the window): every window of
// the same row count writes the same blocks, so the padding is refilled only
// when the row count (the layout) changes.
uint8_t* mx_scales(Sketch& sk, int64_t rows, const WinLayout& w, hipStream_t st) {
  const size_t sbytes = w.o_rs + align_up((size_t)(rows + 1) * w.ldrs, 256);  // + the sink row
  if (sk.lp_scales.bytes < sbytes || sk.lp_scales_rows != rows || sk.lp_scales_windows != 1) {
    sk.lp_scales.reserve(sbytes);
    fill_i32(st, sk.lp_scales.as<int>(), (int64_t)(sbytes / 4), 0x7f7f7f7f);
    sk.lp_scales_rows = rows;
    sk.lp_scales_windows = 1;
  }
  return sk.lp_scales.as<uint8_t>();
}

// The quantise pass of one window into `base` on stream st.
void quantize_window(Sketch& sk, const double* H, int64_t ldh, int64_t rows, const WinLayout& w, char* base,
                     hipStream_t st) {
  const int64_t n = sk.p.n;
  float* rmax = (float*)(base + w.o_rmax);
  float* cmax = (float*)(base + w.o_cmax);
  Timed t(sk.ctx, "window", st);
  dim3 grid((unsigned)cdiv(rows, kQ), (unsigned)cdiv(n, kQ));
  if (w.mx) {
    uint8_t* sc = mx_scales(sk, rows, w, st);
    dim3 gmx((unsigned)cdiv(rows, kMxR), (unsigned)cdiv(n, kMxC));
    hipLaunchKernelGGL(k_quant_fp8_mx, gmx, dim3(kThreads), 0, st, H, ldh, rows, n, (uint8_t*)(base + w.o_l), w.ldl,
                       sc + w.o_ls, w.ldls, (uint8_t*)(base + w.o_r), w.ldr, sc + w.o_rs, w.ldrs);
  } else if (w.fp8) {
    fill_i32(st, (int*)rmax, rows, 0);
    fill_i32(st, (int*)cmax, n, 0);
    hipLaunchKernelGGL(k_absmax_fp8, grid, dim3(kThreads), 0, st, H, ldh, rows, n, (unsigned*)rmax, (unsigned*)cmax);
    hipLaunchKernelGGL(k_quant_fp8, grid, dim3(kThreads), 0, st, H, ldh, rows, n, (const float*)rmax,
                       (const float*)cmax, (uint8_t*)(base + w.o_l), w.ldl, (uint8_t*)(base + w.o_r), w.ldr);
  } else if (w.f32) {
    hipLaunchKernelGGL((k_quant<float, false>), grid, dim3(kThreads), 0, st, H, ldh, rows, n, (float*)(base + w.o_l),
                       (float*)nullptr, w.ldl, (float*)(base + w.o_r), (float*)nullptr, w.ldr);
  } else if (w.split) {
    __bf16* L = (__bf16*)(base + w.o_l);
    __bf16* R = (__bf16*)(base + w.o_r);
    hipLaunchKernelGGL((k_quant<__bf16, true>), grid, dim3(kThreads), 0, st, H, ldh, rows, n, L, L + w.le, w.ldl, R,
                       R + w.re, w.ldr);
  } else {
    hipLaunchKernelGGL((k_quant<__bf16, false>), grid, dim3(kThreads), 0, st, H, ldh, rows, n, (__bf16*)(base + w.o_l),
                       (__bf16*)nullptr, w.ldl, (__bf16*)(base + w.o_r), (__bf16*)nullptr, w.ldr);
  }
  check_launch("k_quant");
}

// The three window products from the quantised copies in `base`, on the
// sketch's stream.
// The right product's NT arguments (A = R, B = [Omega; Psi]) without its
// outputs.
NtArgs right_args(const Sketch& sk, int64_t rows, const WinLayout& w, char* base, size_t scale_off = 0) {
  const int64_t k = sk.k, s = sk.s, n = sk.p.n;
  const size_t oe = (size_t)(k + s) * sk.lp_ldomps;
  NtArgs a{};
  a.M = rows; a.N = k + s; a.K = n;
  a.ah = base + w.o_r;
  a.al = w.split ? (const void*)((const __bf16*)(base + w.o_r) + w.re) : nullptr;
  a.amax_a = (w.fp8 && !w.mx) ? (const float*)(base + w.o_rmax) : nullptr;
  if (w.mx) { a.asc = sk.lp_scales.as<uint8_t>() + scale_off + w.o_rs; a.ldas = w.ldrs; }
  a.lda = w.ldr; a.koff_a = 0;
  a.bh = sk.lp_omps.p;
  a.bl = w.split ? (const void*)(sk.lp_omps.as<__bf16>() + oe) : nullptr;
  a.amax_b = w.fp8 ? (const float*)(sk.lp_omps.as<uint8_t>() + align_up(oe, 256)) : nullptr;
  a.ldb = sk.lp_ldomps;
  return a;
}

// Whether a bf16 product (a: K already the three segments' length) runs on
// the 256 x 256 k_mx, the one kernel that reads K segments.
bool seg_kernel_ok(const NtArgs& a, bool forced = false) {
  const NtPlan pl = plan_nt(SKEMA_PREC_BF16, a.M, a.N, a.K);
  return (forced || (pl.ns == 2 && pl.tm == 8 && pl.tn == 4 && pl.wm == 2 && pl.wn == 4)) && a.K % 64 == 0 &&
         ntx_eligible(SKEMA_PREC_BF16, a);
}
// The 256 x 256 bf16 k_mx with a K split of 2: split-BF16's whole-matrix
// left product as a segmented bf16 product (k x n x 3m: 2 row tiles of the
// 401 rows of x, 64 column tiles, 256 blocks)
constexpr NtPlan kSegLeftPlan{0, 0, 2, 8, 4, 2, 2, 4};

// left = false: the left product was done by k_quant_left_fp8; rstrip > 0:
// R in 64-column strips of rstrip bytes; scale_off: byte offset of this
// window's block scales in lp_scales.  whole: the window is the whole matrix
// of a nu = 1 update: X, Y and Z are written (not read), the core product's
// K split fills the CUs.
void window_products(Sketch& sk, int64_t row0, int64_t rows, const WinLayout& w, char* base, bool left = true,
                     int64_t rstrip = 0, size_t scale_off = 0, bool whole = false) {
  Ctx* c = sk.ctx;
  hipStream_t st = c->stream;
  const int prec = sk.p.precision;
  const int64_t k = sk.k, s = sk.s, m = sk.p.m, n = sk.p.n;
  const double eta = sk.p.eta, nu = whole ? 0.0 : sk.p.nu;
  const bool fp8 = w.fp8, split = w.split, mx = w.mx;
  float* cmax = (float*)(base + w.o_cmax);
  const size_t ue = (size_t)k * sk.lp_ldups;
  if (left) {
    Timed t(c, "sketch_left");
    NtArgs a{};
    a.M = k; a.N = n; a.K = rows;
    a.ah = sk.lp_ups.p;
    a.al = split ? (const void*)(sk.lp_ups.as<__bf16>() + ue) : nullptr;
    a.amax_a = fp8 ? (const float*)(sk.lp_ups.as<uint8_t>() + fp8_ups_amax_offset(sk)) : nullptr;
    a.lda = sk.lp_ldups; a.koff_a = row0;
    a.bh = base + w.o_l;
    a.bl = split ? (const void*)((const __bf16*)(base + w.o_l) + w.le) : nullptr;
    a.amax_b = (fp8 && !mx) ? cmax : nullptr;
    if (mx) { a.bsc = sk.lp_scales.as<uint8_t>() + w.o_ls; a.ldbs = w.ldls; }
    a.ldb = w.ldl;
    a.out.C1 = sk.X.as<double>(); a.out.ldc1 = k; a.out.alpha1 = eta; a.out.beta1 = nu; a.out.nsplit = n;
    a.out.C2 = a.out.C1; a.out.ldc2 = k;
    bool done = false;
    if (whole && split && rows % 64 == 0) {
      // split-BF16 over the whole matrix: [Uh | Uh | Ul] x [Lh | Ll | Lh] as
      // one bf16 product over K segments of the window's rows (see the right
      // product below), on the 256 x 256 kernel with a K split of 2
      NtArgs sgm = a;
      sgm.K = 3 * rows;
      sgm.seg_steps = (int)(rows / 64);
      sgm.seg_da0 = 0; sgm.seg_da1 = 0; sgm.seg_da2 = static_cast<const char*>(a.al) - static_cast<const char*>(a.ah);
      sgm.seg_db0 = 0; sgm.seg_db1 = static_cast<const char*>(a.bl) - static_cast<const char*>(a.bh); sgm.seg_db2 = 0;
      sgm.al = nullptr; sgm.bl = nullptr;
      if (seg_kernel_ok(sgm, true)) {
        gemm_nt(st, SKEMA_PREC_BF16, sgm, &c->gws.buf, &t, false, &kSegLeftPlan);
        done = true;
      }
    }
    if (!done) gemm_nt(st, prec, a, &c->gws.buf, &t);
  }
  {
    Timed t(c, "sketch_right");
    NtArgs a = right_args(sk, rows, w, base, scale_off);
    a.a_strip = rstrip;
    a.out.C1 = sk.Y.as<double>() + row0; a.out.ldc1 = m; a.out.alpha1 = eta; a.out.beta1 = nu; a.out.nsplit = k;
    a.gtype = w.gtype; a.gh = base + w.o_g; a.gl = base + w.o_g + w.ge * w.gsz; a.ldg = w.ldl;
    bool done = false;
    if (whole && split && n % 64 == 0) {
      // split-BF16 over the whole matrix: hi.hi + hi.lo + lo.hi as one bf16
      // product over K segments [Ah | Ah | Al] x [Bh | Bl | Bh] (whole
      // 64-element steps, no padding inside a segment), on the 256 x 256
      // kernel (no K split) instead of the two-plane 128 x 128 one
      NtArgs sgm = a;
      const int64_t segs = n / 64;
      sgm.K = 3 * segs * 64;
      sgm.seg_steps = (int)segs;
      sgm.seg_da0 = 0; sgm.seg_da1 = 0; sgm.seg_da2 = static_cast<const char*>(a.al) - static_cast<const char*>(a.ah);
      sgm.seg_db0 = 0; sgm.seg_db1 = static_cast<const char*>(a.bl) - static_cast<const char*>(a.bh); sgm.seg_db2 = 0;
      sgm.al = nullptr; sgm.bl = nullptr;
      if (seg_kernel_ok(sgm)) {
        gemm_nt(st, SKEMA_PREC_BF16, sgm, nullptr, &t);
        done = true;
      }
    }
    if (!done) gemm_nt(st, prec, a, &c->gws.buf, &t);
  }
  {
    Timed t(c, "sketch_core");
    const size_t pe = (size_t)s * sk.lp_ldups;
    NtArgs a{};
    a.M = s; a.N = s; a.K = rows;
    a.ah = sk.lp_phi.p;
    a.al = split ? (const void*)(sk.lp_phi.as<__bf16>() + pe) : nullptr;
    a.lda = sk.lp_ldups; a.koff_a = row0;
    a.bh = base + w.o_g;
    a.bl = split ? (const void*)(base + w.o_g + w.ge * w.gsz) : nullptr;
    a.ldb = w.ldl;
    a.out.C1 = sk.Z.as<double>(); a.out.ldc1 = s; a.out.alpha1 = eta; a.out.beta1 = nu; a.out.nsplit = s;
    a.out.C2 = a.out.C1; a.out.ldc2 = s;
    gemm_nt(st, w.core_prec, a, &c->gws.buf, &t, whole);
  }
}

void launch_quant_left(hipStream_t st, int64_t k, int64_t n, const NtArgs& a, const QlArgs& q) {
  const dim3 grid((unsigned)cdiv(n, kQlCols));
  switch (ql_tm(k)) {
    case 1: hipLaunchKernelGGL(k_quant_left_fp8<1>, grid, dim3(kQlThreads), 0, st, a, q); break;
    case 2: hipLaunchKernelGGL(k_quant_left_fp8<2>, grid, dim3(kQlThreads), 0, st, a, q); break;
    case 4: hipLaunchKernelGGL(k_quant_left_fp8<4>, grid, dim3(kQlThreads), 0, st, a, q); break;
    default: hipLaunchKernelGGL(k_quant_left_fp8<7>, grid, dim3(kQlThreads), 0, st, a, q); break;
  }
  check_launch("k_quant_left_fp8");
}

// Quantise pass and left product in one kernel (k_quant_left_fp8) for
// block-scaled fp8 windows with k <= 448, an even row count and 16-byte
// aligned column pairs; false (nothing launched) otherwise.
bool quant_left_fp8(Sketch& sk, const double* H, int64_t ldh, int64_t row0, int64_t rows, const WinLayout& w,
                    char* base, int64_t* rstrip_out) {
  const int64_t k = sk.k, n = sk.p.n;
  if (!w.mx || !ql_tm(k) || rows < 2 || (rows & 1) || (ldh & 1) || ((uintptr_t)H & 15) || n < 1) return false;
  // 32-bit offsets: the lane's rows of the window column, Upsilon rows
  if (rows * 8 >= ((int64_t)1 << 31) || k * sk.lp_ldups >= ((int64_t)1 << 31)) return false;
  Ctx* c = sk.ctx;
  hipStream_t st = c->stream;
  uint8_t* sc = mx_scales(sk, rows, w, st);
  Timed t(c, "sketch_left");
  NtArgs a{};
  a.M = k; a.N = n; a.K = rows;
  a.ah = sk.lp_ups.p;
  a.amax_a = (const float*)(sk.lp_ups.as<uint8_t>() + fp8_ups_amax_offset(sk));
  a.lda = sk.lp_ldups; a.koff_a = row0;
  // x is k x n with ld k: the kernel's epilogue takes a block's columns as
  // one contiguous range
  a.out.C1 = sk.X.as<double>(); a.out.ldc1 = k; a.out.alpha1 = sk.p.eta; a.out.beta1 = sk.p.nu; a.out.nsplit = n;
  a.out.C2 = a.out.C1; a.out.ldc2 = k;
  // R in 64-column strips when the right product runs on k_mx (whole
  // cache lines per block and step), otherwise in rows
  const int64_t rstrip = nt_picks_mx(sk.p.precision, right_args(sk, rows, w, base)) ? (rows + 1) * 64 : 0;
  *rstrip_out = rstrip;
  QlArgs q{H, ldh, rows, n, (uint8_t*)(base + w.o_r), w.ldr, sc + w.o_rs, w.ldrs, sk.lp_ldups, rstrip,
           rows, 0, 0};
  launch_quant_left(st, k, n, a, q);
  return true;
}


}  // namespace

void update_dense_window_lowp(Sketch& sk, const double* H, int64_t ldh, int64_t row0, int64_t rows) {
  const WinLayout w = win_layout(sk, row0, rows);
  sk.win.reserve(w.total);
  char* base = sk.win.as<char>();
  int64_t rstrip = 0;
  if (quant_left_fp8(sk, H, ldh, row0, rows, w, base, &rstrip)) {
    window_products(sk, row0, rows, w, base, false, rstrip);
    return;
  }
  quantize_window(sk, H, ldh, rows, w, base, sk.ctx->stream);
  window_products(sk, row0, rows, w, base);
}

// bf16 and split-BF16 at nu = 1: the quantisation is elementwise, so one
// quantise pass over the whole matrix and the three products over all rows
// give the windows' sums up to FP32 summation order.  Taken when the whole
// matrix's L, R and G copies fit in kWholeCopyBytes.  Not fp32: its left
// product would accumulate K = m terms in FP32 instead of K = w per window
// (summed in FP64), 4x its per-window error at 65536 x 16384.
constexpr size_t kWholeCopyBytes = (size_t)32 << 30;
bool update_dense_whole_lowp(Sketch& sk, const double* A, int64_t lda) {
  const int64_t m = sk.p.m, wr = sk.p.window;
  const int prec = sk.p.precision;
  if (prec != SKEMA_PREC_BF16 && prec != SKEMA_PREC_BF16X3) return false;
  if (sk.sparse_maps || sk.p.nu != 1.0 || wr <= 0 || wr >= m) return false;
  const WinLayout w = win_layout(sk, 0, m);
  if (w.total > kWholeCopyBytes) return false;
  sk.win.reserve(w.total);
  char* base = sk.win.as<char>();
  quantize_window(sk, A, lda, m, w, base, sk.ctx->stream);
  window_products(sk, 0, m, w, base, true, 0, 0, true);
  return true;
}

bool update_dense_all_lowp(Sketch& sk, const double* A, int64_t lda) {
  const int64_t m = sk.p.m, n = sk.p.n, k = sk.k, s = sk.s, wr = sk.p.window;
  const int prec = sk.p.precision;
  if (update_dense_whole_lowp(sk, A, lda)) return true;
  // nu = 1: the windows' products are plain sums over the windows (other nu
  // would weigh them in the FP32 accumulators, not exactly as the FP64 state
  // does)
  if (prec != SKEMA_PREC_FP8 || sk.sparse_maps || sk.p.nu != 1.0) return false;
  // several windows of whole 128-row steps (one window takes the per-window
  // path): the block scales of L (32-row blocks) and R (per row) are then
  // the same whether the blocks are counted from each window's first row or
  // from the matrix's
  if (wr <= 0 || wr % kQlRows != 0 || wr >= m) return false;
  if (!ql_tm(k) || (m & 1) || (lda & 1) || ((uintptr_t)A & 15) || n < 1) return false;
  // 32-bit offsets: the lane's rows of a column, Upsilon rows
  if (m * 8 >= ((int64_t)1 << 31) || k * sk.lp_ldups >= ((int64_t)1 << 31)) return false;
  Ctx* c = sk.ctx;
  hipStream_t st = c->stream;
  const double eta = sk.p.eta;
  // The whole matrix as one window: R (m rows + the pass's sink row), its
  // block scales and G (m x s); no L (the pass builds L in LDS).
  WinLayout w = win_layout(sk, 0, m);
  w.o_l = 0;
  w.o_r = 0;
  w.o_rs = 0;
  // R in 64-column strips (256-byte aligned) when the right product runs on
  // k_mx
  const int64_t rstrip = nt_picks_mx(prec, right_args(sk, m, w, nullptr)) ? (int64_t)align_up((size_t)(m + 1) * 64, 256) : 0;
  if (rstrip) w.re = (size_t)(w.ldr / 64) * rstrip;
  w.o_rmax = w.o_cmax = w.o_g = align_up(w.re * w.esz, 256);
  w.total = w.o_g + align_up(w.ge * w.gsz, 256);
  // R and G of the whole matrix beyond kWholeCopyBytes: the window loop
  if (w.total > kWholeCopyBytes) return false;
  sk.win.reserve(w.total);
  char* base = sk.win.as<char>();
  // unit scales (0x7f) in the K padding of R's rows; refilled when the
  // layout changes (lp_scales_windows 0 marks this one)
  const size_t sbytes = align_up((size_t)(m + 1) * w.ldrs, 256);
  if (sk.lp_scales.bytes < sbytes || sk.lp_scales_rows != m || sk.lp_scales_windows != 0) {
    sk.lp_scales.reserve(sbytes);
    fill_i32(st, sk.lp_scales.as<int>(), (int64_t)(sbytes / 4), 0x7f7f7f7f);
    sk.lp_scales_rows = m;
    sk.lp_scales_windows = 0;
  }
  NtArgs r = right_args(sk, m, w, base);
  {
    Timed t(c, "sketch_left");
    NtArgs a{};
    a.M = k; a.N = n; a.K = m;
    a.ah = sk.lp_ups.p;
    a.amax_a = (const float*)(sk.lp_ups.as<uint8_t>() + fp8_ups_amax_offset(sk));
    a.lda = sk.lp_ldups; a.koff_a = 0;
    // x (zero before the update, k x n with ld k) = eta Upsilon Q(A)
    // (nu = 1): written once, no read
    a.out.C1 = sk.X.as<double>(); a.out.ldc1 = k; a.out.alpha1 = eta; a.out.beta1 = 0.0; a.out.nsplit = n;
    a.out.C2 = a.out.C1; a.out.ldc2 = k;
    QlArgs q{A, lda, m, n, (uint8_t*)(base + w.o_r), w.ldr, sk.lp_scales.as<uint8_t>() + w.o_rs, w.ldrs,
             sk.lp_ldups, rstrip, m, 0, 0};
    launch_quant_left(st, k, n, a, q);
  }
  {
    // [y | G] = Q(A) [Omega; Psi]^T over all m rows: enough tiles to fill the
    // chip without a K split; y (zero before the update) written once
    Timed t(c, "sketch_right");
    r.a_strip = rstrip;
    r.out.C1 = sk.Y.as<double>(); r.out.ldc1 = m; r.out.alpha1 = eta; r.out.beta1 = 0.0; r.out.nsplit = k;
    r.gtype = w.gtype; r.gh = base + w.o_g; r.gl = nullptr; r.ldg = w.ldl;
    gemm_nt(st, prec, r, nullptr, &t);
  }
  {
    // z = eta Phi G (K = m), K split so the tiles times the splits fill
    // the CUs once
    Timed t(c, "sketch_core");
    NtArgs a{};
    a.M = s; a.N = s; a.K = m;
    a.ah = sk.lp_phi.p;
    a.lda = sk.lp_ldups; a.koff_a = 0;
    a.bh = base + w.o_g;
    a.ldb = w.ldl;
    a.out.C1 = sk.Z.as<double>(); a.out.ldc1 = s; a.out.alpha1 = eta; a.out.beta1 = 0.0; a.out.nsplit = s;
    a.out.C2 = a.out.C1; a.out.ldc2 = s;
    gemm_nt(st, w.core_prec, a, &c->gws.buf, &t, true);
  }
  return true;
}

void gemm_nt_bf16(hipStream_t st, int64_t M, int64_t N, int64_t K, const __bf16* A, int64_t lda, const __bf16* B,
                  int64_t ldb, double alpha, double* C, int64_t ldc, DevBuf* ws) {
  NtArgs a{};
  a.M = M; a.N = N; a.K = K;
  a.ah = A; a.lda = lda; a.koff_a = 0;
  a.bh = B; a.ldb = ldb;
  a.out.C1 = C; a.out.ldc1 = ldc; a.out.alpha1 = alpha; a.out.beta1 = 0.0; a.out.nsplit = N;
  a.out.C2 = C; a.out.ldc2 = ldc;
  gemm_nt(st, SKEMA_PREC_BF16, a, ws);
}

void lowp_core_bf16(Sketch& sk, const __bf16* G, int64_t ldg, int64_t row0, int64_t rows) {
  SKEMA_REQUIRE(sk.p.precision == SKEMA_PREC_BF16, kErrUnsupported, "bf16 core product needs the bf16 maps");
  const int64_t s = sk.s;
  NtArgs a{};
  a.M = s; a.N = s; a.K = rows;
  a.ah = sk.lp_phi.p;
  a.lda = sk.lp_ldups; a.koff_a = row0;
  a.bh = G;
  a.ldb = ldg;
  a.out.C1 = sk.Z.as<double>(); a.out.ldc1 = s; a.out.alpha1 = sk.p.eta; a.out.beta1 = sk.p.nu; a.out.nsplit = s;
  a.out.C2 = a.out.C1; a.out.ldc2 = s;
  gemm_nt(sk.ctx->stream, SKEMA_PREC_BF16, a, &sk.ctx->gws.buf);
}

}  // namespace skema
