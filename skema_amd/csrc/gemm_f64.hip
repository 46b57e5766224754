// FP64 GEMM on v_mfma_f64_16x16x4_f64 for gfx950.
//
// Block tile BM x BN (4 waves in a 2x2 grid, each wave owning a
// (BM/2) x (BN/2) patch of 16x16 MFMA blocks), K staged through LDS in slices
// of 16 with a one-slice register prefetch.  Operands are column-major with
// arbitrary leading dimension; transposition is a template parameter so every
// global load stays coalesced along the contiguous index.
//
// MFMA 16x16x4 f64 lane layout (64 lanes, l = lane):
//   A operand  A[i][k]  with i = l % 16, k = l / 16
//   B operand  B[k][j]  with j = l % 16, k = l / 16
//   D result   D[i][j]  with j = l % 16, i = l / 16 + 4 * v, v = 0..3
//              (rows interleave by 4 across the four accumulator registers)
#include "gemm.hpp"

namespace skema {
namespace {

constexpr int kBK = 16;
constexpr int kThreads = 256;

struct KArgs {
  int64_t M, N, K;
  const double* A;
  int64_t lda;
  const double* B;
  int64_t ldb;
  GemmOut out;
  int64_t k_per_split;
  double* partial;  // split-K slices [z][M x N], nullptr -> direct epilogue
  int tiles_m, tiles_n;
};

// Workgroup ids are dealt round-robin to the 8 XCDs; give each XCD a
// contiguous range of tiles so neighbouring tiles share that XCD's L2.
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

template <int BM, int BN, bool TA, bool TB>
__global__ __launch_bounds__(kThreads) void k_gemm_f64(KArgs g) {
  constexpr int WTM = BM / 2, WTN = BN / 2, TM = WTM / 16, TN = WTN / 16;
  constexpr int NA = BM * kBK / kThreads, NB = BN * kBK / kThreads;
  __shared__ double As[kBK][BM + 2];
  __shared__ double Bs[kBK][BN + 2];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t m0 = (int64_t)(t % g.tiles_m) * BM;
  const int64_t n0 = (int64_t)(t / g.tiles_m) * BN;
  const int64_t kbeg = (int64_t)blockIdx.z * g.k_per_split;
  const int64_t kend = (g.K < kbeg + g.k_per_split) ? g.K : kbeg + g.k_per_split;

  d4 acc[TM][TN];
#pragma unroll
  for (int a = 0; a < TM; ++a)
#pragma unroll
    for (int b = 0; b < TN; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

  double ra[NA], rb[NB];

  auto a_idx = [&](int e, int& i, int& p) {
    if (!TA) { i = tid % BM; p = tid / BM + e * (kThreads / BM); }
    else     { p = tid % kBK; i = tid / kBK + e * (kThreads / kBK); }
  };
  auto b_idx = [&](int e, int& p, int& j) {
    if (!TB) { p = tid % kBK; j = tid / kBK + e * (kThreads / kBK); }
    else     { j = tid % BN; p = tid / BN + e * (kThreads / BN); }
  };
  auto gload = [&](int64_t k0) {
#pragma unroll
    for (int e = 0; e < NA; ++e) {
      int i, p;
      a_idx(e, i, p);
      const int64_t gi = m0 + i, gp = k0 + p;
      double v = 0.0;
      if (gi < g.M && gp < kend) v = TA ? g.A[gp + gi * g.lda] : g.A[gi + gp * g.lda];
      ra[e] = v;
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) {
      int p, j;
      b_idx(e, p, j);
      const int64_t gj = n0 + j, gp = k0 + p;
      double v = 0.0;
      if (gj < g.N && gp < kend) v = TB ? g.B[gj + gp * g.ldb] : g.B[gp + gj * g.ldb];
      rb[e] = v;
    }
  };
  auto sstore = [&]() {
#pragma unroll
    for (int e = 0; e < NA; ++e) {
      int i, p;
      a_idx(e, i, p);
      As[p][i] = ra[e];
    }
#pragma unroll
    for (int e = 0; e < NB; ++e) {
      int p, j;
      b_idx(e, p, j);
      Bs[p][j] = rb[e];
    }
  };

  if (kbeg < kend) gload(kbeg);
  for (int64_t k0 = kbeg; k0 < kend; k0 += kBK) {
    __syncthreads();
    sstore();
    __syncthreads();
    if (k0 + kBK < kend) gload(k0 + kBK);
#pragma unroll
    for (int kk = 0; kk < kBK; kk += 4) {
      double a[TM], b[TN];
#pragma unroll
      for (int x = 0; x < TM; ++x) a[x] = As[kk + (lane >> 4)][wm * WTM + x * 16 + (lane & 15)];
#pragma unroll
      for (int y = 0; y < TN; ++y) b[y] = Bs[kk + (lane >> 4)][wn * WTN + y * 16 + (lane & 15)];
#pragma unroll
      for (int x = 0; x < TM; ++x)
#pragma unroll
        for (int y = 0; y < TN; ++y)
          acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], acc[x][y], 0, 0, 0);
    }
  }

#pragma unroll
  for (int x = 0; x < TM; ++x)
#pragma unroll
    for (int y = 0; y < TN; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t i = m0 + wm * WTM + x * 16 + (lane >> 4) + 4 * v;
        const int64_t j = n0 + wn * WTN + y * 16 + (lane & 15);
        if (i < g.M && j < g.N) {
          if (g.partial)
            g.partial[(int64_t)blockIdx.z * g.M * g.N + i + j * g.M] = acc[x][y][v];
          else
            store_out(g.out, i, j, acc[x][y][v]);
        }
      }
}

// Deterministic split-K reduction: slices are summed in z order.
__global__ void k_splitk_reduce(const double* part, int splits, int64_t M, int64_t N, GemmOut o) {
  const int64_t total = M * N;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int z = 0; z < splits; ++z) s += part[(int64_t)z * total + e];
    store_out(o, e % M, e / M, s);
  }
}

template <int BM, int BN>
void launch_cfg(hipStream_t st, bool ta, bool tb, const KArgs& a, dim3 grid) {
  if (!ta && !tb) hipLaunchKernelGGL((k_gemm_f64<BM, BN, false, false>), grid, dim3(kThreads), 0, st, a);
  else if (!ta && tb) hipLaunchKernelGGL((k_gemm_f64<BM, BN, false, true>), grid, dim3(kThreads), 0, st, a);
  else if (ta && !tb) hipLaunchKernelGGL((k_gemm_f64<BM, BN, true, false>), grid, dim3(kThreads), 0, st, a);
  else hipLaunchKernelGGL((k_gemm_f64<BM, BN, true, true>), grid, dim3(kThreads), 0, st, a);
}

constexpr int kCUs = 256;

}  // namespace

void gemm_f64(hipStream_t stream, char transA, char transB, int64_t M, int64_t N, int64_t K,
              const double* A, int64_t lda, const double* B, int64_t ldb, const GemmOut& out,
              GemmWorkspace* ws) {
  SKEMA_REQUIRE(M >= 0 && N >= 0 && K >= 0, kErrShape, "gemm_f64: negative dimension");
  if (M == 0 || N == 0) return;
  const bool ta = (transA == 'T' || transA == 't');
  const bool tb = (transB == 'T' || transB == 't');
  SKEMA_REQUIRE(lda >= (ta ? (K > 0 ? K : 1) : M), kErrShape, "gemm_f64: lda too small");
  SKEMA_REQUIRE(ldb >= (tb ? N : (K > 0 ? K : 1)), kErrShape, "gemm_f64: ldb too small");
  SKEMA_REQUIRE(out.nsplit >= 0 && out.nsplit <= N, kErrArg, "gemm_f64: bad column split");
  SKEMA_REQUIRE(out.nsplit == 0 || out.ldc1 >= M, kErrShape, "gemm_f64: ldc too small");
  SKEMA_REQUIRE(out.nsplit == N || out.ldc2 >= M, kErrShape, "gemm_f64: ldc too small");
  SKEMA_REQUIRE(M * N < (int64_t)1 << 40, kErrShape, "gemm_f64: output too large");

  // Tile choice: minimise the busiest CU's share (tiles are spread evenly,
  // one tile's time ~ BM*BN*K); larger tiles win ties (less L2 traffic).
  struct Cfg { int bm, bn; };
  const Cfg cfgs[] = {{128, 128}, {128, 64}, {64, 128}, {64, 64}};
  Cfg best = cfgs[3];
  int64_t best_cost = INT64_MAX;
  for (const Cfg& c : cfgs) {
    const int64_t tiles = cdiv(M, c.bm) * cdiv(N, c.bn);
    const int64_t cost = cdiv(tiles, kCUs) * c.bm * c.bn;
    if (cost < best_cost) { best_cost = cost; best = c; }
  }
  const int tiles_m = (int)cdiv(M, best.bm), tiles_n = (int)cdiv(N, best.bn);
  const int64_t tiles = (int64_t)tiles_m * tiles_n;
  SKEMA_REQUIRE(tiles < (int64_t)1 << 31, kErrShape, "gemm_f64: too many tiles");

  // Split K when the output alone cannot fill the chip.
  int splits = 1;
  if (tiles < kCUs && K >= 4 * 256 && ws != nullptr) {
    splits = (int)cdiv(2 * kCUs, tiles);
    const int64_t max_by_k = K / 256;
    if (splits > max_by_k) splits = (int)max_by_k;
    if (splits > 64) splits = 64;
    if (splits < 1) splits = 1;
  }
  KArgs a;
  a.M = M; a.N = N; a.K = K; a.A = A; a.lda = lda; a.B = B; a.ldb = ldb; a.out = out;
  a.tiles_m = tiles_m; a.tiles_n = tiles_n;
  a.partial = nullptr;
  a.k_per_split = K;
  if (splits > 1) {
    a.k_per_split = cdiv(cdiv(K, splits), kBK) * kBK;
    splits = (int)cdiv(K, a.k_per_split);
    ws->buf.reserve((size_t)splits * M * N * sizeof(double));
    a.partial = ws->buf.as<double>();
  }
  if (K == 0) a.k_per_split = 0;
  dim3 grid((unsigned)tiles, 1, (unsigned)splits);
  if (best.bm == 128 && best.bn == 128) launch_cfg<128, 128>(stream, ta, tb, a, grid);
  else if (best.bm == 128) launch_cfg<128, 64>(stream, ta, tb, a, grid);
  else if (best.bn == 128) launch_cfg<64, 128>(stream, ta, tb, a, grid);
  else launch_cfg<64, 64>(stream, ta, tb, a, grid);
  check_launch("k_gemm_f64");
  if (splits > 1) {
    const int64_t total = M * N;
    const int blocks = (int)std::min<int64_t>(cdiv(total, 256), 4096);
    hipLaunchKernelGGL(k_splitk_reduce, dim3(blocks), dim3(256), 0, stream, a.partial, splits, M, N, out);
    check_launch("k_splitk_reduce");
  }
}

void gemm_f64(hipStream_t stream, char transA, char transB, int64_t M, int64_t N, int64_t K,
              double alpha, const double* A, int64_t lda, const double* B, int64_t ldb,
              double beta, double* C, int64_t ldc, GemmWorkspace* ws) {
  GemmOut o;
  o.C1 = C; o.ldc1 = ldc; o.alpha1 = alpha; o.beta1 = beta; o.nsplit = N;
  o.C2 = C; o.ldc2 = ldc;
  gemm_f64(stream, transA, transB, M, N, K, A, lda, B, ldb, o, ws);
}

}  // namespace skema
