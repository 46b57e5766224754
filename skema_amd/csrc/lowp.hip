// Reduced-precision forms of the two large sketch contractions of a dense
// window (A_w is w x n, FP64, column-major in HBM):
//   left : x        = Upsilon[:, idx] * A_w          (k x n, reduce over w)
//   right: [y | G]  = A_w * [Omega; Psi]^T           (w x (k+s), reduce over n)
// The maps are converted once (sketch creation) to the form's storage,
// row-major with K contiguous.  A_w is converted while it is staged into LDS,
// so no reduced copy of the window is ever written to HBM.  Products
// accumulate in FP32 on the matrix cores; the epilogue widens to FP64 and
// applies the nu/eta update into the FP64 state.  The small core product
// Z += Phi[:, idx] * G stays on the FP64 path.
//
// Forms:
//   bf16   : v_mfma_f32_16x16x32_bf16, one term
//   bf16x3 : same instruction, x = hi + lo (both bf16), a*b ~ hi*hi + hi*lo + lo*hi
//   fp32   : v_mfma_f32_16x16x4_f32
//   fp8    : v_mfma_scale_f32_16x16x128_f8f6f4 with e4m3 operands and one
//            power-of-two (E8M0) scale per 32 consecutive K elements of every
//            row / column, chosen from the block's max |x| as it is staged.
#include <cmath>

#include "lowp.hpp"
#include "util.hpp"

namespace skema {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef int i8v __attribute__((ext_vector_type(8)));

constexpr int kThreads = 256;

struct LpArgs {
  int64_t M, N, K;
  const void* map_hi;  // reduced map, row-major, K contiguous
  const void* map_lo;  // split-BF16 low part (bf16x3 only)
  const float* map_scale;  // unused by the bf16 forms
  int64_t ldmap;
  int64_t map_koff;    // K offset into the map's columns (window row0 for left)
  const double* data;  // A_w
  int64_t ldd;
  GemmOut out;
  int tiles_m;
};

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

// ---------------------------------------------------------------------------
// BF16 / split-BF16 GEMM.  LEFT: A operand = map (M = k rows), B = A_w
// (K = w rows, contiguous).  RIGHT: A operand = A_w (M = w rows, contiguous),
// B = map (N = k+s rows).  LDS tiles are K-major: [row][BK + 8] bf16.
// ---------------------------------------------------------------------------
template <int BM, int BN, bool SPLIT, bool LEFT>
__global__ __launch_bounds__(kThreads) void k_gemm_bf16(LpArgs g) {
  constexpr int BK = 64, LDK = BK + 8;
  constexpr int TM = BM / 32, TN = BN / 32;
  constexpr int NP = SPLIT ? 2 : 1;
  constexpr int MAPR = LEFT ? BM : BN;   // rows of the map tile
  constexpr int DATR = LEFT ? BN : BM;   // rows (K-major) of the data tile
  constexpr int MAPC = MAPR * BK / 8 / kThreads;  // 16-byte map chunks per thread
  constexpr int DATC = DATR * BK / 8 / kThreads;  // 8-element data chunks per thread
  __shared__ __attribute__((aligned(16))) __bf16 sA[NP][BM][LDK];
  __shared__ __attribute__((aligned(16))) __bf16 sB[NP][BN][LDK];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t m0 = (int64_t)(t % g.tiles_m) * BM;
  const int64_t n0 = (int64_t)(t / g.tiles_m) * BN;
  const int64_t r0_map = LEFT ? m0 : n0;
  const int64_t r0_dat = LEFT ? n0 : m0;
  const int64_t nmap = LEFT ? g.M : g.N;
  const int64_t ndat = LEFT ? g.N : g.M;
  const __bf16* mh = static_cast<const __bf16*>(g.map_hi);
  const __bf16* ml = static_cast<const __bf16*>(g.map_lo);

  f4v acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = f4v{0.f, 0.f, 0.f, 0.f};

  bf16x8 rm[NP][MAPC];
  double rd[DATC][8];

  auto gload = [&](int64_t k0) {
    // map chunks: chunk c -> row c / (BK/8), k-chunk c % (BK/8)
#pragma unroll
    for (int e = 0; e < MAPC; ++e) {
      const int c = tid + e * kThreads;
      const int row = c / (BK / 8), kc = (c % (BK / 8)) * 8;
      const int64_t gr = r0_map + row, gk = k0 + kc;
#pragma unroll
      for (int h = 0; h < NP; ++h) {
        const __bf16* src = (h == 0 ? mh : ml) + gr * g.ldmap + g.map_koff + gk;
        bf16x8 v;
        if (gr < nmap && gk + 8 <= g.K && (((uintptr_t)src) & 15) == 0) {
          v = *reinterpret_cast<const bf16x8*>(src);
        } else {
#pragma unroll
          for (int q = 0; q < 8; ++q) v[q] = (gr < nmap && gk + q < g.K) ? src[q] : (__bf16)0.0f;
        }
        rm[h][e] = v;
      }
    }
    // data chunks
#pragma unroll
    for (int e = 0; e < DATC; ++e) {
      const int c = tid + e * kThreads;
      int row, kc;
      if (LEFT) { row = c / (BK / 8); kc = (c % (BK / 8)) * 8; }   // K-contiguous columns of A_w
      else      { row = c % BM;       kc = (c / BM) * 8; }          // M-contiguous rows of A_w
      const int64_t gr = r0_dat + row, gk = k0 + kc;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        double v = 0.0;
        if (gr < ndat && gk + q < g.K) v = LEFT ? g.data[(gk + q) + gr * g.ldd] : g.data[gr + (gk + q) * g.ldd];
        rd[e][q] = v;
      }
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int e = 0; e < MAPC; ++e) {
      const int c = tid + e * kThreads;
      const int row = c / (BK / 8), kc = (c % (BK / 8)) * 8;
#pragma unroll
      for (int h = 0; h < NP; ++h) {
        __bf16* dst = LEFT ? &sA[h][row][kc] : &sB[h][row][kc];
        *reinterpret_cast<bf16x8*>(dst) = rm[h][e];
      }
    }
#pragma unroll
    for (int e = 0; e < DATC; ++e) {
      const int c = tid + e * kThreads;
      int row, kc;
      if (LEFT) { row = c / (BK / 8); kc = (c % (BK / 8)) * 8; }
      else      { row = c % BM;       kc = (c / BM) * 8; }
      bf16x8 hi, lo;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const float x = (float)rd[e][q];
        const __bf16 h = (__bf16)x;
        hi[q] = h;
        if (SPLIT) lo[q] = (__bf16)(x - (float)h);
      }
      __bf16* dh = LEFT ? &sB[0][row][kc] : &sA[0][row][kc];
      *reinterpret_cast<bf16x8*>(dh) = hi;
      if (SPLIT) {
        __bf16* dl = LEFT ? &sB[NP - 1][row][kc] : &sA[NP - 1][row][kc];
        *reinterpret_cast<bf16x8*>(dl) = lo;
      }
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
  // D layout of the 16x16 f32 MFMAs: lane l holds rows 4*(l/16)+v of column l%16.
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

// Map conversion: src logical (rows x cols) column-major FP64 (ld) ->
// row-major bf16 hi (and lo) with leading dimension ldo.
__global__ void k_convert_map_bf16(const double* src, int64_t ld, int64_t rows, int64_t cols, __bf16* hi, __bf16* lo,
                                   int64_t ldo) {
  const int64_t total = rows * cols;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t r = e % rows, c = e / rows;
    const float x = (float)src[r + c * ld];
    const __bf16 h = (__bf16)x;
    hi[r * ldo + c] = h;
    if (lo) lo[r * ldo + c] = (__bf16)(x - (float)h);
  }
}

template <int BM, int BN, bool SPLIT, bool LEFT>
void launch_bf16(hipStream_t st, const LpArgs& a, int tiles) {
  hipLaunchKernelGGL((k_gemm_bf16<BM, BN, SPLIT, LEFT>), dim3(tiles), dim3(kThreads), 0, st, a);
}

template <bool SPLIT, bool LEFT>
void gemm_bf16(hipStream_t st, LpArgs a) {
  // Same tile rule as the FP64 GEMM: minimise the busiest CU's share.
  struct Cfg { int bm, bn; };
  const Cfg cfgs[] = {{128, 128}, {128, 64}, {64, 128}, {64, 64}};
  Cfg best = cfgs[3];
  int64_t best_cost = INT64_MAX;
  for (const Cfg& c : cfgs) {
    const int64_t tiles = cdiv(a.M, c.bm) * cdiv(a.N, c.bn);
    const int64_t cost = cdiv(tiles, 256) * c.bm * c.bn;
    if (cost < best_cost) { best_cost = cost; best = c; }
  }
  a.tiles_m = (int)cdiv(a.M, best.bm);
  const int tiles = (int)(a.tiles_m * cdiv(a.N, best.bn));
  if (best.bm == 128 && best.bn == 128) launch_bf16<128, 128, SPLIT, LEFT>(st, a, tiles);
  else if (best.bm == 128) launch_bf16<128, 64, SPLIT, LEFT>(st, a, tiles);
  else if (best.bn == 128) launch_bf16<64, 128, SPLIT, LEFT>(st, a, tiles);
  else launch_bf16<64, 64, SPLIT, LEFT>(st, a, tiles);
  check_launch("k_gemm_bf16");
}

}  // namespace

void prepare_lowp_maps(Sketch& sk) {
  const int prec = sk.p.precision;
  SKEMA_REQUIRE(prec == SKEMA_PREC_BF16 || prec == SKEMA_PREC_BF16X3, kErrUnsupported,
                "this reduced-precision form is not built yet");
  hipStream_t st = sk.ctx->stream;
  const int64_t k = sk.k, s = sk.s, m = sk.p.m, n = sk.p.n;
  const bool split = prec == SKEMA_PREC_BF16X3;
  // Upsilon: k x m, row-major, ld padded to a multiple of 8 elements.
  sk.lp_ldups = (m + 7) / 8 * 8;
  sk.lp_ldomps = (n + 7) / 8 * 8;
  const size_t ups_elems = (size_t)k * sk.lp_ldups, omps_elems = (size_t)(k + s) * sk.lp_ldomps;
  sk.lp_ups.alloc(ups_elems * 2 * (split ? 2 : 1));
  sk.lp_omps.alloc(omps_elems * 2 * (split ? 2 : 1));
  __bf16* uh = sk.lp_ups.as<__bf16>();
  __bf16* oh = sk.lp_omps.as<__bf16>();
  auto grid = [](int64_t e) { return (unsigned)std::min<int64_t>(cdiv(e, 256), 65536); };
  hipLaunchKernelGGL(k_convert_map_bf16, dim3(grid(k * m)), dim3(256), 0, st, sk.upsL, sk.ldups, k, m, uh,
                     split ? uh + ups_elems : nullptr, sk.lp_ldups);
  hipLaunchKernelGGL(k_convert_map_bf16, dim3(grid((k + s) * n)), dim3(256), 0, st, sk.omps.as<double>(), k + s,
                     k + s, n, oh, split ? oh + omps_elems : nullptr, sk.lp_ldomps);
  check_launch("k_convert_map_bf16");
  SKEMA_HIP(hipStreamSynchronize(st));
}

void update_dense_window_lowp(Sketch& sk, const double* H, int64_t ldh, int64_t row0, int64_t rows) {
  Ctx* c = sk.ctx;
  hipStream_t st = c->stream;
  const int64_t k = sk.k, s = sk.s, m = sk.p.m, n = sk.p.n;
  const double eta = sk.p.eta, nu = sk.p.nu;
  const bool split = sk.p.precision == SKEMA_PREC_BF16X3;
  double* G = sk.G.as<double>();
  const size_t ups_elems = (size_t)k * sk.lp_ldups, omps_elems = (size_t)(k + s) * sk.lp_ldomps;
  {
    Timed t(c, "sketch_left");
    LpArgs a{};
    a.M = k; a.N = n; a.K = rows;
    a.map_hi = sk.lp_ups.as<__bf16>();
    a.map_lo = split ? sk.lp_ups.as<__bf16>() + ups_elems : nullptr;
    a.ldmap = sk.lp_ldups;
    a.map_koff = row0;
    a.data = H; a.ldd = ldh;
    a.out.C1 = sk.X.as<double>(); a.out.ldc1 = k; a.out.alpha1 = eta; a.out.beta1 = nu; a.out.nsplit = n;
    a.out.C2 = a.out.C1; a.out.ldc2 = k;
    if (split) gemm_bf16<true, true>(st, a); else gemm_bf16<false, true>(st, a);
  }
  {
    Timed t(c, "sketch_right");
    LpArgs a{};
    a.M = rows; a.N = k + s; a.K = n;
    a.map_hi = sk.lp_omps.as<__bf16>();
    a.map_lo = split ? sk.lp_omps.as<__bf16>() + omps_elems : nullptr;
    a.ldmap = sk.lp_ldomps;
    a.map_koff = 0;
    a.data = H; a.ldd = ldh;
    a.out.C1 = sk.Y.as<double>() + row0; a.out.ldc1 = m; a.out.alpha1 = eta; a.out.beta1 = nu; a.out.nsplit = k;
    a.out.C2 = G; a.out.ldc2 = rows; a.out.alpha2 = 1.0; a.out.beta2 = 0.0;
    if (split) gemm_bf16<true, false>(st, a); else gemm_bf16<false, false>(st, a);
  }
  {
    Timed t(c, "sketch_core");
    gemm_f64(st, 'N', 'N', s, s, rows, eta, sk.phiL + row0 * sk.ldphi, sk.ldphi, G, rows, nu, sk.Z.as<double>(), s,
             &c->gws);
  }
}

}  // namespace skema
