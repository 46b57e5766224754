// BF16 form of the fused Gaussian-RBF window + sketch (BF16 Gram, FP32
// accumulation, FP32 exponent).  The point set is rounded to BF16 once per
// window update (row-major, features contiguous, padded to DP = 32, 64 or
// 128) together with the squared norms of the rounded points, so
// D_ij = |x~_i|^2 + |x~_j|^2 - 2 x~_i.x~_j is the squared distance of the
// rounded points up to FP32 accumulation.  The rounding errors of the d
// products are independent, so the Gram carries a relative error of about
// sqrt(d) 2^-8 |x|^2 / D, which gamma = O(1/d) turns into ~0.3% of K for the
// mixture data — the order of rounding K itself to BF16 for the sketch MFMAs.
//   left : X       <- nu X + eta * Upsilon[:, idx] K_w     (k x n)
//   right: [y | G]  =  K_w [Omega; Psi]^T                   (w x (k+s))
// then Z <- nu Z + eta Phi[:, idx] G on the BF16 NT GEMM.  K never leaves
// registers/LDS.
#include <cmath>

#include "dma.hpp"
#include "lowp.hpp"
#include "rbf_fused.hpp"
#include "rbf_lowp.hpp"
#include "util.hpp"

namespace skema {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef int i4v __attribute__((ext_vector_type(4)));

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

// exp(-gamma * max(D, 0)) in FP32, the window's own diagonal exactly 1
// (src/Skema_Kernel.cpp:49-74).
__device__ __forceinline__ float rbf_value_f(float ni, float nj, float g, float gamma, bool diag) {
  if (diag) return 1.f;
  float D = ni + nj - 2.f * g;
  D = D > 0.f ? D : 0.f;
  return __expf(-gamma * D);
}

struct RbfLpArgs {
  int64_t M = 0, N = 0, K = 0;   // product shape
  const __bf16* P = nullptr;     // rounded points, n x ldpb row-major
  int64_t ldpb = 0;
  const float* nrm = nullptr;    // |x~_j|^2
  float gamma = 0.f;
  int64_t row0 = 0;
  const __bf16* OP = nullptr;    // right: [Omega; Psi] row-major ((k+s) x ldop)
  int64_t ldop = 0;
  const __bf16* U = nullptr;     // left: Upsilon row-major (k x ldu)
  int64_t ldu = 0;
  GemmOut out;                   // right: columns < nsplit -> out.C1 (Y rows)
  __bf16* G = nullptr;           // right: columns >= nsplit as bf16, col-major ld ldg
  int64_t ldg = 0;
  int tiles_m = 0;
  int64_t k_per_split = 0;
  float* part = nullptr;         // right: FP32 partials [split][M x N]
};

__device__ __forceinline__ void store_right(const RbfLpArgs& g, int64_t i, int64_t j, double v) {
  if (j >= g.out.nsplit) g.G[i + (j - g.out.nsplit) * g.ldg] = (__bf16)(float)v;
  else store_out(g.out, i, j, v);
}

// 16 bytes (8 bf16) of a row-major bf16 matrix at element e, elements past
// `lim` (exclusive) read as zero.
__device__ __forceinline__ bf16x8 load8(const __bf16* row, int64_t e, int64_t lim) {
  const __bf16* src = row + e;
  if (e + 8 <= lim && (((uintptr_t)src) & 15) == 0) return *reinterpret_cast<const bf16x8*>(src);
  bf16x8 v;
#pragma unroll
  for (int q = 0; q < 8; ++q) v[q] = (e + q < lim) ? src[q] : (__bf16)0.f;
  return v;
}

// ---- points -> BF16 rows + norms of the rounded points --------------------------
__global__ void k_points_bf16(const double* X, int64_t ldp, int64_t n, int64_t d, int64_t ldpb, __bf16* P,
                              float* nrm) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  float s = 0.f;
  for (int64_t f0 = 0; f0 < ldpb; f0 += 8) {
    bf16x8 v;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int64_t f = f0 + q;
      const __bf16 h = (f < d) ? (__bf16)(float)X[j + f * ldp] : (__bf16)0.f;
      const float hf = (float)h;
      s += hf * hf;
      v[q] = h;
    }
    *reinterpret_cast<bf16x8*>(P + j * ldpb + f0) = v;
  }
  nrm[j] = s;
}

// ---- right -----------------------------------------------------------------------
// 512 threads own 64 window rows x 128*TN columns.  Per step of 32 points:
// wave w forms the Gram tile (row tile w%4, point tile w/4) over all DP
// features (its window rows' features in registers, the points' from LDS),
// exponentiates its four values per lane and writes K (bf16) to LDS; the
// sketch MFMAs of wave w cover all 64 rows x its 16*TN columns with the
// [Omega; Psi] operands loaded from global memory straight into registers.
// Pipelined one step apart (sketch of t, Gram + exp of t+1, points of t+2
// staged, t+3 loaded), double-buffered LDS, one barrier per step.
constexpr int kThreadsR = 512;

template <int DP, int TN>
__global__ __launch_bounds__(kThreadsR) void k_rbf_right_bf16(RbfLpArgs g) {
  constexpr int BM = 64, BK = 32, PS = DP + 8, CHUNKS = BK * DP / 8;
  __shared__ __attribute__((aligned(16))) __bf16 Xs[2][BK][PS];
  __shared__ __attribute__((aligned(16))) __bf16 Ks[2][BM][BK + 8];
  __shared__ float NJ[4][BK];
  __shared__ float NI[BM];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, lq = lane >> 4;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t m0 = (int64_t)(t % g.tiles_m) * BM;
  const int split = t / g.tiles_m;
  const int64_t cw0 = (int64_t)blockIdx.y * (128 * TN) + (int64_t)wave * (16 * TN);
  const int64_t kbeg = (int64_t)split * g.k_per_split;
  const int64_t kend = (g.K < kbeg + g.k_per_split) ? g.K : kbeg + g.k_per_split;
  const int64_t steps = (kend - kbeg + BK - 1) / BK;
  const int rt = wave & 3, pt = wave >> 2;

  bf16x8 xa[DP / 32];  // Gram A operand: window row m0 + 16 rt + l16, features 32c + 8 lq ..
  {
    const int64_t i = m0 + 16 * rt + l16;
#pragma unroll
    for (int c = 0; c < DP / 32; ++c)
      xa[c] = (i < g.M) ? load8(g.P + (g.row0 + i) * g.ldpb, 32 * c + 8 * lq, g.ldpb) : bf16x8{};
  }
  if (tid < BM) NI[tid] = (m0 + tid < g.M) ? g.nrm[g.row0 + m0 + tid] : 0.f;

  bf16x8 xr = {};
  float njr = 0.f;
  auto load_points = [&](int64_t u) {
    const int64_t j0 = kbeg + u * BK;
    if (tid < CHUNKS) {
      const int jj = tid / (DP / 8), fc = tid % (DP / 8);
      const int64_t j = j0 + jj;
      xr = (j < kend) ? load8(g.P + j * g.ldpb, 8 * fc, g.ldpb) : bf16x8{};
    }
    njr = (tid < BK && j0 + tid < kend) ? g.nrm[j0 + tid] : 0.f;
  };
  auto store_points = [&](int b, int64_t u) {
    if (tid < CHUNKS) *reinterpret_cast<bf16x8*>(&Xs[b][tid / (DP / 8)][8 * (tid % (DP / 8))]) = xr;
    if (tid < BK) NJ[u & 3][tid] = njr;
  };
  auto gram_exp = [&](int b, int64_t u) {  // Xs[b] (step u) -> Ks[b]
    f4v gr = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int c = 0; c < DP / 32; ++c) {
      const bf16x8 bx = *reinterpret_cast<const bf16x8*>(&Xs[b][16 * pt + l16][32 * c + 8 * lq]);
      gr = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa[c], bx, gr, 0, 0, 0);
    }
    const int jl = 16 * pt + l16;
    const int64_t j = kbeg + u * BK + jl;
    const float nj = NJ[u & 3][jl];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int il = 16 * rt + 4 * lq + v;
      const int64_t i = m0 + il;
      float kv = 0.f;
      if (i < g.M && j < kend) kv = rbf_value_f(NI[il], nj, gr[v], g.gamma, g.row0 + i == j);
      Ks[b][il][jl] = (__bf16)kv;
    }
  };
  bf16x8 op[TN], opn[TN];
  auto load_op = [&](bf16x8 (&dst)[TN], int64_t u) {
    const int64_t j = kbeg + u * BK + 8 * lq;
#pragma unroll
    for (int y = 0; y < TN; ++y) {
      const int64_t c = cw0 + 16 * y + l16;
      dst[y] = (c < g.N && j < kend) ? load8(g.OP + c * g.ldop, j, kend) : bf16x8{};
    }
  };

  f4v acc[4][TN];
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < TN; ++y) acc[x][y] = f4v{0.f, 0.f, 0.f, 0.f};

  load_points(0); store_points(0, 0);
  load_points(1); store_points(1, 1);
  load_op(op, 0);
  __syncthreads();
  gram_exp(0, 0);
  load_points(2);
  __syncthreads();
  for (int64_t u = 0; u < steps; ++u) {
    const int b = (int)(u & 1), nb = b ^ 1;
    if (u + 1 < steps) load_op(opn, u + 1);
#pragma unroll
    for (int x = 0; x < 4; ++x) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&Ks[b][16 * x + l16][8 * lq]);
#pragma unroll
      for (int y = 0; y < TN; ++y) acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, op[y], acc[x][y], 0, 0, 0);
    }
    if (u + 1 < steps) gram_exp(nb, u + 1);  // points of step u+1 sit in Xs[nb]
    store_points(b, u + 2);                  // points of step u+2 -> Xs[b]
    load_points(u + 3);
    __syncthreads();
#pragma unroll
    for (int y = 0; y < TN; ++y) op[y] = opn[y];
  }
  // 16x16 f32 result: lane l holds rows 4*(l/16)+v of column l%16, i.e. four
  // consecutive rows of one column-major column: stored as one vector when
  // all four rows exist.
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < TN; ++y) {
      const int64_t i0 = m0 + 16 * x + 4 * lq;
      const int64_t j = cw0 + 16 * y + l16;
      if (i0 >= g.M || j >= g.N) continue;
      if (g.part) {
        float* p = g.part + (int64_t)split * g.M * g.N + j * g.M;
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (i0 + v < g.M) p[i0 + v] = acc[x][y][v];
      } else if (i0 + 4 <= g.M && j >= g.out.nsplit) {
        __bf16* pg = g.G + i0 + (j - g.out.nsplit) * g.ldg;
        const bf16x4 h = bf16x4{(__bf16)acc[x][y][0], (__bf16)acc[x][y][1], (__bf16)acc[x][y][2],
                                (__bf16)acc[x][y][3]};
        if ((((uintptr_t)pg) & 7) == 0) *reinterpret_cast<bf16x4*>(pg) = h;
        else for (int v = 0; v < 4; ++v) pg[v] = h[v];
      } else {
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (i0 + v < g.M) store_right(g, i0 + v, j, (double)acc[x][y][v]);
      }
    }
}

// ---- right, tall tile ------------------------------------------------------------
// 512 threads own 256 window rows x 208 columns (13 column tiles) of
// [y | G]; k + s columns take ceil((k+s)/208) column blocks, each forming its
// own Gram tiles.  Wave w owns rows 32w .. 32w+31 and all 208 columns.
//
// K never leaves registers: per 32-point step the wave forms the transposed
// Gram tiles (points x its window rows) with the points as the A operand,
// read from LDS in the order pi(4a + t) = 8a + 4h + t for sub-tile h, so
// the f32 result (lane l: window row l%16, points pi(4 (l/16) + v)) holds
// exactly points 8 (l/16) .. +7 of row l%16 after the two sub-tiles: the
// A-operand layout of the sketch MFMA, in natural point order.  The
// [Omega; Psi] slice (208 x 32 points, 64-byte rows), the step's points
// (32 x DP) and their norms stream global -> LDS through global_load_lds
// into a kRingSlots-slot ring, so four steps of loads are in flight while
// one is consumed; one barrier per step.  16-byte chunks are XOR-swizzled
// within rows so the 16 rows a lane group reads at one chunk hit distinct
// banks.
// kTallXT 16-row tiles per wave: 4 (one wave per SIMD, accumulators in
// AGPRs) halves the LDS reads of the [Omega; Psi] operands per MFMA against 2.
constexpr int kTallRows = 256, kTallCols = 208, kRingSlots = 6, kTallXT = 4;

// chunk-position swizzle of a points row with CPR 16-byte chunks, for the
// rows 8a + 4h + t (a, t < 4) of either sub-tile
template <int CPR>
__device__ __forceinline__ int xs_swz(int r) {
  if constexpr (CPR == 16) return ((r >> 1) & 12) | (r & 3);
  else if constexpr (CPR == 8) return ((r >> 2) & 6) | ((r >> 1) & 1);
  else return (r >> 3) & 3;
}
// 64-byte rows of the [Omega; Psi] slice, read 16 consecutive rows at a time
__device__ __forceinline__ int os_swz(int r) { return (r >> 2) & 3; }

// exp(-gamma max(D, 0)) from u = -gamma log2(e) D = kappa g + (alpha_i + beta_j)
// with alpha = -gamma log2(e) |x_i|^2 (likewise beta) and kappa = 2 gamma
// log2(e): u <= 0 exactly when D >= 0.
__device__ __forceinline__ float rbf_exp2(float g, float ab, float kappa) {
  const float u = fmaf(g, kappa, ab);
  return __builtin_amdgcn_exp2f(u < 0.f ? u : 0.f);
}

template <int DP>
__global__ __launch_bounds__(64 * (kTallRows / (16 * kTallXT)), 1) void k_rbf_right_bf16_tall(RbfLpArgs g) {
  constexpr int XT = kTallXT;                        // 16-row tiles per wave
  constexpr int NW = kTallRows / (16 * XT);          // waves
  constexpr int BM = kTallRows, BN = kTallCols, BK = 32, TNC = BN / 16, NS = kRingSlots;
  constexpr int CPR = DP / 8;                       // 16-byte chunks per points row
  constexpr int RPI = 1024 / (2 * DP);              // points rows per 1 KB copy
  constexpr int OS_BYTES = BN * BK * 2;             // [Omega; Psi] slice
  constexpr int XS_BYTES = BK * DP * 2;             // points
  constexpr int NJ_OFF = OS_BYTES + XS_BYTES;       // norms: one 256-byte dword copy
  constexpr int SLOT = (NJ_OFF + 256 + 1023) / 1024 * 1024;
  constexpr int I_OS = OS_BYTES / 1024, I_XS = XS_BYTES / 1024, ITEMS = I_OS + I_XS + 1;
  constexpr int IPW = (ITEMS + NW - 1) / NW;        // copy items per wave, at most
  static_assert(OS_BYTES % 1024 == 0 && XS_BYTES % 1024 == 0 && (NS - 2) * IPW <= 62, "copy items");
  __shared__ __attribute__((aligned(1024))) uint8_t ring[NS * SLOT];
  // wave index as a scalar: the copy items, their counts and the vmcnt waits
  // below branch on it, and a per-lane value would turn every wave-uniform
  // branch into exec-masked code that runs each wait case in turn
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int l16 = lane & 15, lq = lane >> 4;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t m0 = (int64_t)(t % g.tiles_m) * BM;
  const int split = t / g.tiles_m;
  const int64_t c0 = (int64_t)blockIdx.y * BN;
  const int64_t kbeg = (int64_t)split * g.k_per_split;
  const int64_t kend = (g.K < kbeg + g.k_per_split) ? g.K : kbeg + g.k_per_split;
  const int steps = (int)((kend - kbeg + BK - 1) / BK);
  const int dw = (ITEMS - wave + NW - 1) / NW;  // copies this wave issues per step
  const float lg = 1.4426950408889634f * g.gamma;  // gamma log2(e)
  const float kappa = 2.f * lg;
  const int64_t wr0 = m0 + 16 * XT * wave;        // the wave's first row
  const bool rows_full = wr0 + 16 * XT <= g.M;

  // window rows of the wave: features (Gram B operand) and scaled norms
  bf16x8 xa[XT][DP / 32];
  float al[XT];
#pragma unroll
  for (int x = 0; x < XT; ++x) {
    const int64_t i = wr0 + 16 * x + l16;
#pragma unroll
    for (int c = 0; c < DP / 32; ++c)
      xa[x][c] = (i < g.M) ? load8(g.P + (g.row0 + i) * g.ldpb, 32 * c + 8 * lq, g.ldpb) : bf16x8{};
    al[x] = (i < g.M) ? -lg * g.nrm[g.row0 + i] : 0.f;
  }

  // Per-lane sources of the wave's copies at the first step; a step moves
  // each by BK points (2 BK bytes along an [Omega; Psi] row, BK points rows,
  // 4 BK bytes of norms).  Near the end of the split the points rows and
  // norms are clamped to the last point instead.
  const char* dsrc[IPW];
#pragma unroll
  for (int s = 0; s < IPW; ++s) {
    const int it = wave + NW * s;
    if (it < I_OS) {  // 16 rows of the [Omega; Psi] slice
      const int r = 16 * it + (lane >> 2), p = lane & 3;
      int64_t col = c0 + r;
      col = col < g.N ? col : g.N - 1;
      dsrc[s] = reinterpret_cast<const char*>(g.OP + col * g.ldop + kbeg + 8 * (p ^ os_swz(r)));
    } else if (it < I_OS + I_XS) {  // RPI points rows
      const int x = it - I_OS;
      const int r = RPI * x + lane / CPR, p = lane % CPR;
      dsrc[s] = reinterpret_cast<const char*>(g.P + (kbeg + r) * g.ldpb + 8 * (p ^ xs_swz<CPR>(r)));
    } else {  // norms of the step's points (lanes >= 32 duplicate, unused)
      dsrc[s] = reinterpret_cast<const char*>(g.nrm + kbeg + (lane & 31));
    }
  }
  auto issue = [&](int u) {
    uint8_t* slot = ring + (u % NS) * SLOT;
    const bool tail = kbeg + (int64_t)(u + 1) * BK > kend;
#pragma unroll
    for (int s = 0; s < IPW; ++s) {
      const int it = wave + NW * s;
      if (it < I_OS) {
        __builtin_amdgcn_global_load_lds((const void*)(dsrc[s] + (int64_t)u * (2 * BK)),
                                         (SKEMA_LDS void*)(slot + 1024 * it), 16, 0, 0);
      } else if (it < I_OS + I_XS) {
        const int x = it - I_OS;
        const char* src = dsrc[s] + (int64_t)u * BK * (2 * DP);
        if (tail) {
          const int r = RPI * x + lane / CPR, p = lane % CPR;
          int64_t j = kbeg + (int64_t)u * BK + r;
          j = j < kend ? j : kend - 1;
          src = reinterpret_cast<const char*>(g.P + j * g.ldpb + 8 * (p ^ xs_swz<CPR>(r)));
        }
        __builtin_amdgcn_global_load_lds((const void*)src, (SKEMA_LDS void*)(slot + OS_BYTES + 1024 * x), 16, 0, 0);
      } else if (it == I_OS + I_XS) {
        const char* src = dsrc[s] + (int64_t)u * (4 * BK);
        if (tail) {
          int64_t j = kbeg + (int64_t)u * BK + (lane & 31);
          j = j < kend ? j : kend - 1;
          src = reinterpret_cast<const char*>(g.nrm + j);
        }
        __builtin_amdgcn_global_load_lds((const void*)src, (SKEMA_LDS void*)(slot + NJ_OFF), 4, 0, 0);
      }
    }
  };

  // Gram tiles of step u (transposed: points x the wave's window rows)
  auto gram = [&](int u, f4v (&gr)[2][XT]) {
    const uint8_t* xs = ring + (u % NS) * SLOT + OS_BYTES;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int r = 8 * (l16 >> 2) + 4 * h + (l16 & 3);  // points row read by this lane
#pragma unroll
      for (int x = 0; x < XT; ++x) gr[h][x] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < DP / 32; ++c) {
        const int pos = (4 * c + lq) ^ xs_swz<CPR>(r);
        const bf16x8 bx = *reinterpret_cast<const bf16x8*>(xs + r * (2 * DP) + 16 * pos);
#pragma unroll
        for (int x = 0; x < XT; ++x) gr[h][x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bx, xa[x][c], gr[h][x], 0, 0, 0);
      }
    }
  };
  // K of step u from its Gram tiles, in registers (A operands of the
  // sketch), without checks: interior steps (every row and point in range,
  // the window's diagonal not among them)
  auto exp_fast = [&](int u, const f4v (&gr)[2][XT], bf16x8 (&kv)[XT]) {
    const float* nj = reinterpret_cast<const float*>(ring + (u % NS) * SLOT + NJ_OFF);
    f4v be[2];
    be[0] = *reinterpret_cast<const f4v*>(nj + 8 * lq);
    be[1] = *reinterpret_cast<const f4v*>(nj + 8 * lq + 4);
#pragma unroll
    for (int x = 0; x < XT; ++x)
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int v = 0; v < 4; ++v) kv[x][4 * h + v] = (__bf16)rbf_exp2(gr[h][x][v], al[x] - lg * be[h][v], kappa);
  };
  auto interior = [&](int u) {
    const int64_t j0 = kbeg + (int64_t)u * BK, gi0 = g.row0 + wr0;
    return rows_full && j0 + BK <= kend && (j0 + BK <= gi0 || j0 >= gi0 + 16 * XT);
  };
  // the same with the range and diagonal checks
  auto exp_edge = [&](int u, const f4v (&gr)[2][XT], bf16x8 (&kv)[XT]) {
    const float* nj = reinterpret_cast<const float*>(ring + (u % NS) * SLOT + NJ_OFF);
    const int64_t j0 = kbeg + (int64_t)u * BK;
#pragma unroll
    for (int x = 0; x < XT; ++x) {
      const int64_t i = wr0 + 16 * x + l16;
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int64_t j = j0 + 8 * lq + 4 * h + v;
          float e = 0.f;
          if (i < g.M && j < kend)
            e = (g.row0 + i == j) ? 1.f : rbf_exp2(gr[h][x][v], al[x] - lg * nj[8 * lq + 4 * h + v], kappa);
          kv[x][4 * h + v] = (__bf16)e;
        }
    }
  };

  f4v acc[XT][TNC];
#pragma unroll
  for (int x = 0; x < XT; ++x)
#pragma unroll
    for (int y = 0; y < TNC; ++y) acc[x][y] = f4v{0.f, 0.f, 0.f, 0.f};

  bf16x8 kv[XT];
#pragma unroll
  for (int p = 0; p < NS - 1; ++p)
    if (p < steps) issue(p);
  if (steps > 0) {
    wait_vm_upto(((NS - 2) < (steps - 1) ? (NS - 2) : (steps - 1)) * dw);  // step 0 complete
    lds_barrier();
    f4v gr[2][XT];
    gram(0, gr);
    if (interior(0)) exp_fast(0, gr, kv);
    else exp_edge(0, gr, kv);
  }
  for (int u = 0; u < steps; ++u) {
    if (u + 1 < steps) {  // step u + 1 complete: at most steps u+2 .. min(u+NS-2, steps-1) in flight
      const int last = (u + NS - 2 < steps - 1) ? u + NS - 2 : steps - 1;
      wait_vm_upto((last - (u + 1)) * dw);
    }
    lds_barrier();  // every wave's copies of step u+1 landed; slot (u-1) % NS is free
    if (u + NS - 1 < steps) issue(u + NS - 1);
    const int un = (u + 1 < steps) ? u + 1 : u;  // the last step forms a Gram it does not use
    f4v gr[2][XT];
    gram(un, gr);
    // the sketch MFMAs of step u, column tile by column tile, each followed
    // by a share of the exponentials of step u+1 (which need only the Gram
    // above): the matrix cores run one tile while the VALU forms the next
    // K values (sched_barrier keeps the compiler from regrouping them)
    const uint8_t* os = ring + (u % NS) * SLOT;
    const float* nj = reinterpret_cast<const float*>(ring + (un % NS) * SLOT + NJ_OFF);
    f4v be[2];
    be[0] = *reinterpret_cast<const f4v*>(nj + 8 * lq);
    be[1] = *reinterpret_cast<const f4v*>(nj + 8 * lq + 4);
    bf16x8 kn[XT];
    bf16x8 o = *reinterpret_cast<const bf16x8*>(os + 64 * l16 + 16 * (lq ^ os_swz(l16)));
#pragma unroll
    for (int y = 0; y < TNC; ++y) {
      bf16x8 on = o;
      if (y + 1 < TNC) {
        const int r = 16 * (y + 1) + l16;
        on = *reinterpret_cast<const bf16x8*>(os + 64 * r + 16 * (lq ^ os_swz(r)));
      }
#pragma unroll
      for (int x = 0; x < XT; ++x) acc[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(kv[x], o, acc[x][y], 0, 0, 0);
      // K values e in [32 y / TNC, 32 (y + 1) / TNC) of this lane's 32 (x, h, v)
#pragma unroll
      for (int e = (8 * XT * y) / TNC; e < (8 * XT * (y + 1)) / TNC; ++e) {
        const int x = e / 8, h = (e / 4) % 2, v = e % 4;
        float k = rbf_exp2(gr[h][x][v], al[x] - lg * be[h][v], kappa);
        asm volatile("" : "+v"(k));  // formed here, not sunk to its use
        kn[x][4 * h + v] = (__bf16)k;
      }
      __builtin_amdgcn_sched_barrier(0);
      o = on;
    }
    if (!interior(un)) exp_edge(un, gr, kn);
#pragma unroll
    for (int x = 0; x < XT; ++x) kv[x] = kn[x];
  }
#pragma unroll
  for (int x = 0; x < XT; ++x)
#pragma unroll
    for (int y = 0; y < TNC; ++y) {
      const int64_t i0 = wr0 + 16 * x + 4 * lq;
      const int64_t j = c0 + 16 * y + l16;
      if (i0 >= g.M || j >= g.N) continue;
      if (g.part) {
        float* p = g.part + (int64_t)split * g.M * g.N + j * g.M;
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (i0 + v < g.M) p[i0 + v] = acc[x][y][v];
      } else if (i0 + 4 <= g.M && j >= g.out.nsplit) {
        __bf16* pg = g.G + i0 + (j - g.out.nsplit) * g.ldg;
        const bf16x4 hh = bf16x4{(__bf16)acc[x][y][0], (__bf16)acc[x][y][1], (__bf16)acc[x][y][2],
                                 (__bf16)acc[x][y][3]};
        if ((((uintptr_t)pg) & 7) == 0) *reinterpret_cast<bf16x4*>(pg) = hh;
        else for (int v = 0; v < 4; ++v) pg[v] = hh[v];
      } else {
#pragma unroll
        for (int v = 0; v < 4; ++v)
          if (i0 + v < g.M) store_right(g, i0 + v, j, (double)acc[x][y][v]);
      }
    }
}

// ---- left, ring ------------------------------------------------------------------
// X <- nu X + eta Upsilon[:, idx] K_w with kLeftRingWaves waves per block:
// wave w owns kLeftJT 16-point tiles of the block's points (features in
// registers) and all 16 TP map rows, so each Upsilon fragment read from LDS
// feeds kLeftJT MFMAs.  Per step of 32 window rows each wave forms its Gram
// tiles (window rows x its points) with the window rows as the A operand
// read in the order pi(4a + t) = 8a + 4x + t for sub-tile x, so the f32
// results hold window rows 8 (l/16) .. +7 of point l%16: the B operand of
// the sketch MFMA in natural row order, matched by one 16-byte read of the
// Upsilon slice.  The step's window rows (32 x DP), their norms and the
// Upsilon slice (16 TP rows x 32 window rows, 64-byte rows) stream global ->
// LDS through global_load_lds into a kRingSlots-slot ring shared by all
// waves, one barrier per step.  Needs row0 a multiple of 8 (16-byte aligned
// slices).
constexpr int kLeftRingWaves = 4, kLeftJT = 4;

template <int DP, int TP>
__global__ __launch_bounds__(64 * kLeftRingWaves, 1) void k_rbf_left_bf16_ring(RbfLpArgs g) {
  constexpr int NW = kLeftRingWaves, JT = kLeftJT, BK = 32, NS = kRingSlots;
  constexpr int PR = 16 * TP;                       // map rows of the block
  constexpr int CPR = DP / 8;
  constexpr int RPI = 1024 / (2 * DP);
  constexpr int US_BYTES = PR * BK * 2;             // Upsilon slice
  constexpr int XS_BYTES = BK * DP * 2;             // window rows
  constexpr int NI_OFF = US_BYTES + XS_BYTES;
  constexpr int SLOT = (NI_OFF + 256 + 1023) / 1024 * 1024;
  constexpr int I_US = US_BYTES / 1024, I_XS = XS_BYTES / 1024, ITEMS = I_US + I_XS + 1;
  constexpr int IPW = (ITEMS + NW - 1) / NW;
  static_assert(US_BYTES % 1024 == 0 && XS_BYTES % 1024 == 0 && (NS - 2) * IPW <= 62, "copy items");
  __shared__ __attribute__((aligned(1024))) uint8_t ring[NS * SLOT];
  // wave index as a scalar: the copy items, their counts and the vmcnt waits
  // below branch on it, and a per-lane value would turn every wave-uniform
  // branch into exec-masked code that runs each wait case in turn
  const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int l16 = lane & 15, lq = lane >> 4;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t wj0 = (int64_t)t * 16 * JT * NW + 16 * JT * wave;  // the wave's first point
  const int64_t p0 = (int64_t)blockIdx.y * PR;
  const int steps = (int)((g.K + BK - 1) / BK);
  const int dw = (ITEMS - wave + NW - 1) / NW;  // copies this wave issues per step
  const float lg = 1.4426950408889634f * g.gamma;
  const float kappa = 2.f * lg;
  const bool pts_full = wj0 + 16 * JT <= g.N;

  bf16x8 xb[JT][DP / 32];
  float be[JT];
#pragma unroll
  for (int y = 0; y < JT; ++y) {
    const int64_t jp = wj0 + 16 * y + l16;
#pragma unroll
    for (int c = 0; c < DP / 32; ++c)
      xb[y][c] = (jp < g.N) ? load8(g.P + jp * g.ldpb, 32 * c + 8 * lq, g.ldpb) : bf16x8{};
    be[y] = (jp < g.N) ? -lg * g.nrm[jp] : 0.f;
  }

  // Per-lane sources of the wave's copies at the first step; a step moves
  // each by BK window rows (2 BK bytes along an Upsilon row, BK points rows,
  // 4 BK bytes of norms).  Past the window's end the sources are clamped.
  const char* dsrc[IPW];
#pragma unroll
  for (int s = 0; s < IPW; ++s) {
    const int it = wave + NW * s;
    if (it < I_US) {  // 16 map rows of the Upsilon slice
      const int r = 16 * it + (lane >> 2), p = lane & 3;
      int64_t row = p0 + r;
      row = row < g.M ? row : g.M - 1;
      dsrc[s] = reinterpret_cast<const char*>(g.U + row * g.ldu + g.row0 + 8 * (p ^ os_swz(r)));
    } else if (it < I_US + I_XS) {  // RPI window rows
      const int x = it - I_US;
      const int r = RPI * x + lane / CPR, p = lane % CPR;
      dsrc[s] = reinterpret_cast<const char*>(g.P + (g.row0 + r) * g.ldpb + 8 * (p ^ xs_swz<CPR>(r)));
    } else {  // norms of the step's window rows (lanes >= 32 duplicate, unused)
      dsrc[s] = reinterpret_cast<const char*>(g.nrm + g.row0 + (lane & 31));
    }
  }
  auto issue = [&](int u) {
    uint8_t* slot = ring + (u % NS) * SLOT;
    const int64_t i0 = (int64_t)u * BK;
    const bool tail = i0 + BK > g.K;
#pragma unroll
    for (int s = 0; s < IPW; ++s) {
      const int it = wave + NW * s;
      if (it < I_US) {
        const char* src = dsrc[s] + i0 * 2;
        if (tail) {  // a chunk starting past the window reads the row's first chunk (finite, multiplied by zero)
          const int r = 16 * it + (lane >> 2), p = lane & 3;
          if (i0 + 8 * (p ^ os_swz(r)) >= g.K) src = dsrc[s] - 16 * (p ^ os_swz(r));
        }
        __builtin_amdgcn_global_load_lds((const void*)src, (SKEMA_LDS void*)(slot + 1024 * it), 16, 0, 0);
      } else if (it < I_US + I_XS) {
        const int x = it - I_US;
        const char* src = dsrc[s] + i0 * (2 * DP);
        if (tail) {
          const int r = RPI * x + lane / CPR, p = lane % CPR;
          int64_t i = i0 + r;
          i = i < g.K ? i : g.K - 1;
          src = reinterpret_cast<const char*>(g.P + (g.row0 + i) * g.ldpb + 8 * (p ^ xs_swz<CPR>(r)));
        }
        __builtin_amdgcn_global_load_lds((const void*)src, (SKEMA_LDS void*)(slot + US_BYTES + 1024 * x), 16, 0, 0);
      } else if (it == I_US + I_XS) {
        const char* src = dsrc[s] + i0 * 4;
        if (tail) {
          int64_t i = i0 + (lane & 31);
          i = i < g.K ? i : g.K - 1;
          src = reinterpret_cast<const char*>(g.nrm + g.row0 + i);
        }
        __builtin_amdgcn_global_load_lds((const void*)src, (SKEMA_LDS void*)(slot + NI_OFF), 4, 0, 0);
      }
    }
  };

  // Gram tiles of step u (window rows x the wave's points)
  auto gram = [&](int u, f4v (&gr)[2][JT]) {
    const uint8_t* xs = ring + (u % NS) * SLOT + US_BYTES;
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      const int r = 8 * (l16 >> 2) + 4 * x + (l16 & 3);
#pragma unroll
      for (int y = 0; y < JT; ++y) gr[x][y] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < DP / 32; ++c) {
        const int pos = (4 * c + lq) ^ xs_swz<CPR>(r);
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(xs + r * (2 * DP) + 16 * pos);
#pragma unroll
        for (int y = 0; y < JT; ++y) gr[x][y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, xb[y][c], gr[x][y], 0, 0, 0);
      }
    }
  };
  auto interior = [&](int u) {
    const int64_t i0 = (int64_t)u * BK, gi0 = g.row0 + i0;  // the step's window rows, as point indices
    return pts_full && i0 + BK <= g.K && (gi0 + BK <= wj0 || gi0 >= wj0 + 16 * JT);
  };
  // K of step u (B operands: lane l holds point l%16 of tile y, window rows
  // 8 (l/16) .. +7 of the step) with the range and diagonal checks
  auto exp_edge = [&](int u, const f4v (&gr)[2][JT], bf16x8 (&kb)[JT]) {
    const float* ni = reinterpret_cast<const float*>(ring + (u % NS) * SLOT + NI_OFF);
    const int64_t i0 = (int64_t)u * BK;
#pragma unroll
    for (int y = 0; y < JT; ++y) {
      const int64_t jp = wj0 + 16 * y + l16;
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int64_t i = i0 + 8 * lq + 4 * x + v;
          float e = 0.f;
          if (i < g.K && jp < g.N)
            e = (g.row0 + i == jp) ? 1.f : rbf_exp2(gr[x][y][v], be[y] - lg * ni[8 * lq + 4 * x + v], kappa);
          kb[y][4 * x + v] = (__bf16)e;
        }
    }
  };
  auto exp_fast = [&](int u, const f4v (&gr)[2][JT], bf16x8 (&kb)[JT]) {
    const float* ni = reinterpret_cast<const float*>(ring + (u % NS) * SLOT + NI_OFF);
    const f4v m0 = *reinterpret_cast<const f4v*>(ni + 8 * lq);
    const f4v m1 = *reinterpret_cast<const f4v*>(ni + 8 * lq + 4);
#pragma unroll
    for (int y = 0; y < JT; ++y)
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int v = 0; v < 4; ++v)
          kb[y][4 * x + v] = (__bf16)rbf_exp2(gr[x][y][v], be[y] - lg * (x ? m1[v] : m0[v]), kappa);
  };

  f4v acc[JT][TP];
#pragma unroll
  for (int y = 0; y < JT; ++y)
#pragma unroll
    for (int z = 0; z < TP; ++z) acc[y][z] = f4v{0.f, 0.f, 0.f, 0.f};

  bf16x8 kb[JT];
#pragma unroll
  for (int p = 0; p < NS - 1; ++p)
    if (p < steps) issue(p);
  if (steps > 0) {
    wait_vm_upto(((NS - 2) < (steps - 1) ? (NS - 2) : (steps - 1)) * dw);  // step 0 complete
    lds_barrier();
    f4v gr[2][JT];
    gram(0, gr);
    if (interior(0)) exp_fast(0, gr, kb);
    else exp_edge(0, gr, kb);
  }
  for (int u = 0; u < steps; ++u) {
    if (u + 1 < steps) {  // step u + 1 complete: at most steps u+2 .. min(u+NS-2, steps-1) in flight
      const int last = (u + NS - 2 < steps - 1) ? u + NS - 2 : steps - 1;
      wait_vm_upto((last - (u + 1)) * dw);
    }
    lds_barrier();  // every wave's copies of step u+1 landed; slot (u-1) % NS is free
    if (u + NS - 1 < steps) issue(u + NS - 1);
    const int un = (u + 1 < steps) ? u + 1 : u;  // the last step forms a Gram it does not use
    f4v gr[2][JT];
    gram(un, gr);
    // the sketch MFMAs of step u, map-row tile by map-row tile, each
    // followed by a share of the exponentials of step u+1
    const uint8_t* us = ring + (u % NS) * SLOT;
    const float* ni = reinterpret_cast<const float*>(ring + (un % NS) * SLOT + NI_OFF);
    f4v mi[2];
    mi[0] = *reinterpret_cast<const f4v*>(ni + 8 * lq);
    mi[1] = *reinterpret_cast<const f4v*>(ni + 8 * lq + 4);
    bf16x8 kn[JT];
    bf16x8 a = *reinterpret_cast<const bf16x8*>(us + 64 * l16 + 16 * (lq ^ os_swz(l16)));
#pragma unroll
    for (int z = 0; z < TP; ++z) {
      bf16x8 an = a;
      if (z + 1 < TP) {
        const int r = 16 * (z + 1) + l16;
        an = *reinterpret_cast<const bf16x8*>(us + 64 * r + 16 * (lq ^ os_swz(r)));
      }
#pragma unroll
      for (int y = 0; y < JT; ++y) acc[y][z] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, kb[y], acc[y][z], 0, 0, 0);
#pragma unroll
      for (int e = (8 * JT * z) / TP; e < (8 * JT * (z + 1)) / TP; ++e) {
        const int y = e / 8, x = (e / 4) % 2, v = e % 4;
        float k = rbf_exp2(gr[x][y][v], be[y] - lg * mi[x][v], kappa);
        asm volatile("" : "+v"(k));  // formed here, not sunk to its use
        kn[y][4 * x + v] = (__bf16)k;
      }
      __builtin_amdgcn_sched_barrier(0);
      a = an;
    }
    if (!interior(un)) exp_edge(un, gr, kn);
#pragma unroll
    for (int y = 0; y < JT; ++y) kb[y] = kn[y];
  }
#pragma unroll
  for (int y = 0; y < JT; ++y) {
    const int64_t jp = wj0 + 16 * y + l16;
#pragma unroll
    for (int z = 0; z < TP; ++z)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t p = p0 + 16 * z + 4 * lq + v;
        if (p < g.M && jp < g.N) store_out(g.out, p, jp, (double)acc[y][z][v]);
      }
  }
}

// Sum of the FP32 partials of the point splits (fixed order), then the
// epilogue; consecutive threads own consecutive rows, so the stores coalesce.
__global__ void k_rbf_right_bf16_reduce(RbfLpArgs g, int splits) {
  const int64_t total = g.M * g.N;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    double v = 0.0;
    for (int z = 0; z < splits; ++z) v += (double)g.part[(int64_t)z * total + e];
    store_right(g, e % g.M, e / g.M, v);
  }
}

// ---- left ------------------------------------------------------------------------
// NW waves own 16 NW points (wave w: points 16w..16w+15, features in
// registers) and 16*TP map rows; the window rows and the Upsilon slice of a
// step are staged once per block, so more waves per block mean fewer L2
// bytes per flop.  Per step of 32 window rows each wave forms
// its two 16x16 Gram tiles (rows 0-15 and 16-31 of the step), exponentiates
// them in registers and packs the 8 values of each lane into the B operand of
// one sketch MFMA: K-slot q of lane l stands for window row
// 16*(q/4) + 4*(l/16) + q%4, which is exactly where the f32 result layout put
// them, and the Upsilon A operand is read from LDS in the same row order.
template <int DP, int TP, int NW = 4>
__global__ __launch_bounds__(64 * NW) void k_rbf_left_bf16(RbfLpArgs g) {
  constexpr int NT = 64 * NW;
  constexpr int BK = 32, PS = DP + 8, PR = 16 * TP, US = BK + 8;
  constexpr int XCH = BK * DP / 8, UCH = PR * BK / 8;  // 16-byte chunks per step
  constexpr int XR = (XCH + NT - 1) / NT, UR = (UCH + NT - 1) / NT;
  __shared__ __attribute__((aligned(16))) __bf16 Xw[2][BK][PS];
  __shared__ __attribute__((aligned(16))) __bf16 Us[2][PR][US];
  __shared__ float NI[2][BK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, lq = lane >> 4;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t n0 = (int64_t)t * 16 * NW;
  const int64_t p0 = (int64_t)blockIdx.y * PR;
  const int64_t steps = (g.K + BK - 1) / BK;
  const int64_t jp = n0 + 16 * wave + l16;

  bf16x8 xb[DP / 32];
#pragma unroll
  for (int c = 0; c < DP / 32; ++c) xb[c] = (jp < g.N) ? load8(g.P + jp * g.ldpb, 32 * c + 8 * lq, g.ldpb) : bf16x8{};
  const float nj = (jp < g.N) ? g.nrm[jp] : 0.f;

  bf16x8 xr[XR], ur[UR];
  float nir = 0.f;
  auto load_step = [&](int64_t u) {
    const int64_t i0 = u * BK;
#pragma unroll
    for (int q = 0; q < XR; ++q) {
      const int c = tid + NT * q;
      const int ii = c / (DP / 8), fc = c % (DP / 8);
      const int64_t i = i0 + ii;
      xr[q] = (c < XCH && i < g.K) ? load8(g.P + (g.row0 + i) * g.ldpb, 8 * fc, g.ldpb) : bf16x8{};
    }
#pragma unroll
    for (int q = 0; q < UR; ++q) {
      const int c = tid + NT * q;
      const int pp = c / (BK / 8), ic = c % (BK / 8);
      const int64_t p = p0 + pp;
      ur[q] = (c < UCH && p < g.M) ? load8(g.U + p * g.ldu + g.row0, i0 + 8 * ic, g.K) : bf16x8{};
    }
    nir = (tid < BK && i0 + tid < g.K) ? g.nrm[g.row0 + i0 + tid] : 0.f;
  };
  auto store_step = [&](int b) {
#pragma unroll
    for (int q = 0; q < XR; ++q) {
      const int c = tid + NT * q;
      if (c < XCH) *reinterpret_cast<bf16x8*>(&Xw[b][c / (DP / 8)][8 * (c % (DP / 8))]) = xr[q];
    }
#pragma unroll
    for (int q = 0; q < UR; ++q) {
      const int c = tid + NT * q;
      if (c < UCH) *reinterpret_cast<bf16x8*>(&Us[b][c / (BK / 8)][8 * (c % (BK / 8))]) = ur[q];
    }
    if (tid < BK) NI[b][tid] = nir;
  };

  f4v acc[TP];
#pragma unroll
  for (int y = 0; y < TP; ++y) acc[y] = f4v{0.f, 0.f, 0.f, 0.f};

  load_step(0);
  store_step(0);
  load_step(1);
  __syncthreads();
  for (int64_t u = 0; u < steps; ++u) {
    const int b = (int)(u & 1);
    const int64_t i0 = u * BK;
    f4v gr[2];
#pragma unroll
    for (int x = 0; x < 2; ++x) {
      gr[x] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int c = 0; c < DP / 32; ++c) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(&Xw[b][16 * x + l16][32 * c + 8 * lq]);
        gr[x] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, xb[c], gr[x], 0, 0, 0);
      }
    }
    bf16x8 kb;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int il = 16 * (q >> 2) + 4 * lq + (q & 3);
      const int64_t i = i0 + il;
      float kv = 0.f;
      if (i < g.K && jp < g.N) kv = rbf_value_f(NI[b][il], nj, gr[q >> 2][q & 3], g.gamma, g.row0 + i == jp);
      kb[q] = (__bf16)kv;
    }
#pragma unroll
    for (int y = 0; y < TP; ++y) {
      const __bf16* row = &Us[b][16 * y + l16][0];
      const bf16x4 lo = *reinterpret_cast<const bf16x4*>(row + 4 * lq);
      const bf16x4 hi = *reinterpret_cast<const bf16x4*>(row + 16 + 4 * lq);
      const bf16x8 a = bf16x8{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
      acc[y] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, kb, acc[y], 0, 0, 0);
    }
    store_step(b ^ 1);
    load_step(u + 2);
    __syncthreads();
  }
#pragma unroll
  for (int y = 0; y < TP; ++y)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t p = p0 + 16 * y + 4 * lq + v;
      if (p < g.M && jp < g.N) store_out(g.out, p, jp, (double)acc[y][v]);
    }
}

constexpr int kTnChoices[] = {1, 2, 3, 4, 5};
constexpr int kTpChoices[] = {1, 2, 4, 8, 13};

template <size_t L>
int pick(const int (&choices)[L], int64_t need) {
  for (int c : choices)
    if (c >= need) return c;
  return choices[L - 1];
}

template <int DP>
void launch_right(hipStream_t st, const RbfLpArgs& a, dim3 grid, int tn) {
  switch (tn) {
    case 1: hipLaunchKernelGGL((k_rbf_right_bf16<DP, 1>), grid, dim3(kThreadsR), 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_rbf_right_bf16<DP, 2>), grid, dim3(kThreadsR), 0, st, a); break;
    case 3: hipLaunchKernelGGL((k_rbf_right_bf16<DP, 3>), grid, dim3(kThreadsR), 0, st, a); break;
    case 4: hipLaunchKernelGGL((k_rbf_right_bf16<DP, 4>), grid, dim3(kThreadsR), 0, st, a); break;
    default: hipLaunchKernelGGL((k_rbf_right_bf16<DP, 5>), grid, dim3(kThreadsR), 0, st, a); break;
  }
}
// Blocks of 8 waves (128 points) for the long map-row tiles, where two
// waves per SIMD is the register-bound occupancy anyway; 4 waves otherwise.
constexpr int kLeftWavesWide = 8;
template <int DP>
void launch_left(hipStream_t st, const RbfLpArgs& a, dim3 grid, int tp) {
  switch (tp) {
    case 1: hipLaunchKernelGGL((k_rbf_left_bf16<DP, 1>), grid, dim3(256), 0, st, a); break;
    case 2: hipLaunchKernelGGL((k_rbf_left_bf16<DP, 2>), grid, dim3(256), 0, st, a); break;
    case 4: hipLaunchKernelGGL((k_rbf_left_bf16<DP, 4>), grid, dim3(256), 0, st, a); break;
    case 8: hipLaunchKernelGGL((k_rbf_left_bf16<DP, 8, kLeftWavesWide>), grid, dim3(64 * kLeftWavesWide), 0, st, a); break;
    default: hipLaunchKernelGGL((k_rbf_left_bf16<DP, 13, kLeftWavesWide>), grid, dim3(64 * kLeftWavesWide), 0, st, a); break;
  }
}

}  // namespace

bool rbf_bf16_supported(const Sketch& sk) {
  return !sk.sparse_maps && sk.p.precision == SKEMA_PREC_BF16 && sk.p.nfeat >= 1 && sk.p.nfeat <= kRbfMaxFeat;
}

void rbf_points_bf16(hipStream_t st, const double* points, int64_t ldp, int64_t n, int64_t d, int dp, __bf16* P,
                     float* nrm) {
  hipLaunchKernelGGL(k_points_bf16, dim3((unsigned)cdiv(n, 256)), dim3(256), 0, st, points, ldp, n, d, (int64_t)dp,
                     P, nrm);
  check_launch("k_points_bf16");
}

void rbf_right_bf16(Ctx* c, const __bf16* P, int dp, const float* nrm, float gamma, int64_t npts, int64_t row0,
                    int64_t rows, const __bf16* OP, int64_t ldop, int64_t k, int64_t s, const GemmOut& yout,
                    __bf16* G, int64_t ldg) {
  hipStream_t st = c->stream;
  RbfLpArgs a;
  a.P = P; a.ldpb = dp; a.nrm = nrm; a.gamma = gamma; a.row0 = row0;
  a.M = rows; a.N = k + s; a.K = npts;
  a.OP = OP; a.ldop = ldop;
  a.out = yout; a.out.nsplit = k;
  a.G = G; a.ldg = ldg;
  // tall tiles (256 rows x 208 columns) once the window fills them;
  // otherwise 64 rows x all columns
  const bool tall = a.M >= 128;
  const int tn = pick(kTnChoices, cdiv(cdiv(a.N, 16), 8));
  const int64_t col_blocks = tall ? cdiv(a.N, kTallCols) : cdiv(a.N, 128 * tn);
  a.tiles_m = (int)cdiv(a.M, tall ? kTallRows : 64);
  // one 512-thread block per CU: choose the split of the points that
  // minimises rounds x points per split plus the partial traffic
  const int64_t per_split = a.tiles_m * col_blocks;
  int64_t splits = 1, best = INT64_MAX;
  const int64_t max_splits = std::max<int64_t>(1, std::min<int64_t>(64, a.K / 1024));
  for (int64_t sp = 1; sp <= max_splits; ++sp) {
    const int64_t cost = cdiv(per_split * sp, 256) * cdiv(a.K, sp) + (sp > 1 ? sp * (a.M * a.N / 65536 + 1) : 0);
    if (cost < best) { best = cost; splits = sp; }
  }
  a.k_per_split = cdiv(cdiv(a.K, splits), 32) * 32;
  splits = cdiv(a.K, a.k_per_split);
  a.part = nullptr;
  if (splits > 1) {
    c->gws.buf.reserve((size_t)splits * a.M * a.N * sizeof(float));
    a.part = c->gws.buf.as<float>();
  }
  const dim3 grid((unsigned)(a.tiles_m * splits), (unsigned)col_blocks);
  if (tall) {
    const dim3 tb(64 * (kTallRows / (16 * kTallXT)));
    if (dp == 32) hipLaunchKernelGGL(k_rbf_right_bf16_tall<32>, grid, tb, 0, st, a);
    else if (dp == 64) hipLaunchKernelGGL(k_rbf_right_bf16_tall<64>, grid, tb, 0, st, a);
    else hipLaunchKernelGGL(k_rbf_right_bf16_tall<128>, grid, tb, 0, st, a);
  } else if (dp == 32) launch_right<32>(st, a, grid, tn);
  else if (dp == 64) launch_right<64>(st, a, grid, tn);
  else launch_right<128>(st, a, grid, tn);
  check_launch("k_rbf_right_bf16");
  if (splits > 1) {
    const int blocks = (int)std::min<int64_t>(cdiv(a.M * a.N, 256), 8192);
    hipLaunchKernelGGL(k_rbf_right_bf16_reduce, dim3(blocks), dim3(256), 0, st, a, (int)splits);
    check_launch("k_rbf_right_bf16_reduce");
  }
}

void update_rbf_window_bf16(Sketch& sk, const double* points, int64_t ldp, int64_t row0, int64_t rows) {
  Ctx* c = sk.ctx;
  hipStream_t st = c->stream;
  const int64_t k = sk.k, s = sk.s, m = sk.p.m, n = sk.p.n, d = sk.p.nfeat;
  const double eta = sk.p.eta, nu = sk.p.nu;
  const int dp = rbf_bf16_padded_features(d);
  const int64_t ldg = cdiv(rows, 128) * 128;
  const size_t o_p = 0, o_n = align_up((size_t)n * dp * 2, 256), o_g = o_n + align_up((size_t)n * 4, 256),
               total = o_g + align_up((size_t)s * ldg * 2, 256);
  sk.win.reserve(total);
  char* base = sk.win.as<char>();
  __bf16* P = (__bf16*)(base + o_p);
  float* nrm = (float*)(base + o_n);
  __bf16* G = (__bf16*)(base + o_g);
  {
    Timed t(c, "window");
    rbf_points_bf16(st, points, ldp, n, d, dp, P, nrm);
  }
  {
    Timed t(c, "rbf_left");
    RbfLpArgs a;
    a.P = P; a.ldpb = dp; a.nrm = nrm; a.gamma = (float)sk.p.gamma; a.row0 = row0;
    a.M = k; a.N = n; a.K = rows;
    a.U = sk.lp_ups.as<__bf16>(); a.ldu = sk.lp_ldups;
    a.out.C1 = sk.X.as<double>(); a.out.ldc1 = k; a.out.alpha1 = eta; a.out.beta1 = nu; a.out.nsplit = n;
    a.out.C2 = a.out.C1; a.out.ldc2 = k;
    const int tp = pick(kTpChoices, cdiv(k, 16));
    if (tp >= 8 && row0 % 8 == 0) {  // long map-row tiles on aligned windows: the ring kernel
      const dim3 grid((unsigned)cdiv(n, 16 * kLeftJT * kLeftRingWaves), (unsigned)cdiv(k, 16 * tp));
      const dim3 block(64 * kLeftRingWaves);
#define SKEMA_LEFT_RING(DP, TP) hipLaunchKernelGGL((k_rbf_left_bf16_ring<DP, TP>), grid, block, 0, st, a)
      if (tp == 8) {
        if (dp == 32) SKEMA_LEFT_RING(32, 8);
        else if (dp == 64) SKEMA_LEFT_RING(64, 8);
        else SKEMA_LEFT_RING(128, 8);
      } else {
        if (dp == 32) SKEMA_LEFT_RING(32, 13);
        else if (dp == 64) SKEMA_LEFT_RING(64, 13);
        else SKEMA_LEFT_RING(128, 13);
      }
#undef SKEMA_LEFT_RING
      check_launch("k_rbf_left_bf16_ring");
    } else {
      const int64_t pts = tp >= 8 ? 16 * kLeftWavesWide : 64;  // points per block
      const dim3 grid((unsigned)cdiv(n, pts), (unsigned)cdiv(k, 16 * tp));
      if (dp == 32) launch_left<32>(st, a, grid, tp);
      else if (dp == 64) launch_left<64>(st, a, grid, tp);
      else launch_left<128>(st, a, grid, tp);
      check_launch("k_rbf_left_bf16");
    }
  }
  {
    Timed t(c, "rbf_right");
    GemmOut y;
    y.C1 = sk.Y.as<double>() + row0; y.ldc1 = m; y.alpha1 = eta; y.beta1 = nu;
    rbf_right_bf16(c, P, dp, nrm, (float)sk.p.gamma, n, row0, rows, sk.lp_omps.as<__bf16>(), sk.lp_ldomps, k, s, y,
                   G, ldg);
  }
  {
    Timed t(c, "sketch_core");
    lowp_core_bf16(sk, G, ldg, row0, rows);
  }
}

}  // namespace skema
