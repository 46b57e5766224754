// Fused Gaussian-RBF window + sketch (FP64).  K_w[i, j] =
// exp(-gamma * ||x_{row0+i} - x_j||^2) is formed tile by tile on the matrix
// cores as ||x_i||^2 + ||x_j||^2 - 2 x_i.x_j (the Gram tile is an MFMA
// product over the d features), exponentiated in registers, handed through
// LDS to the sketch MFMAs, and never written to HBM.
//   left : X       <- nu X + eta * Upsilon[:, idx] K_w     (k x n)
//   right: [y | G]  =  K_w [Omega; Psi]^T                   (w x (k+s))
// The point block a tile needs for its Gram is held in registers for the
// whole K loop (d <= 128); the other side is staged per K step.
#include "rbf.hpp"
#include "rbf_fused.hpp"
#include "util.hpp"

namespace skema {
namespace {

constexpr int kThreads = 256;

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

// exp(-gamma * max(D, 0)) with the window's own diagonal exactly 1
// (src/Skema_Kernel.cpp:49-74).
__device__ __forceinline__ double rbf_value(double ni, double nj, double g, double gamma, bool diag) {
  if (diag) return 1.0;
  double D = ni + nj - 2.0 * g;
  D = D > 0.0 ? D : 0.0;
  return exp(-gamma * D);
}

// ---- right: [y | G](i, c) = sum_j K(i, j) OP(c, j); M = rows, N = k+s, K = n.
// Tile 64 x 128; Gram rows (64 window rows) live in registers, one 16-row
// slab per wave; the 16 points of a K step are staged in LDS.
template <int DMAX>
__global__ __launch_bounds__(kThreads) void k_rbf_right_f64(RbfArgs g) {
  constexpr int BM = 64, BN = 128, BK = 16, TF = DMAX / 4;
  __shared__ double As[BK][BM + 2];
  __shared__ double Bs[BK][BN + 2];
  __shared__ double Xj[DMAX][BK + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t m0 = (int64_t)(t % g.tiles_m) * BM;
  const int64_t n0 = (int64_t)(t / g.tiles_m) * BN;
  const int64_t kbeg = (int64_t)blockIdx.z * g.k_per_split;
  const int64_t kend = (g.K < kbeg + g.k_per_split) ? g.K : kbeg + g.k_per_split;

  // Gram A fragments: window row (m0 + 16*wave + lane%16), features 4t + lane/16.
  double xa[TF];
  {
    const int64_t i = m0 + 16 * wave + (lane & 15);
#pragma unroll
    for (int f = 0; f < TF; ++f) {
      const int64_t ff = 4 * f + (lane >> 4);
      xa[f] = (i < g.M && ff < g.d) ? g.X[(g.row0 + i) + ff * g.ldp] : 0.0;
    }
  }
  // norms of the four Gram output rows this lane produces (f64 D layout)
  double ni[4];
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int64_t i = m0 + 16 * wave + (lane >> 4) + 4 * v;
    ni[v] = (i < g.M) ? g.nrm[g.row0 + i] : 0.0;
  }
  d4 acc[2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

  for (int64_t j0 = kbeg; j0 < kend; j0 += BK) {
    __syncthreads();
    // stage the 16 points' features: Xj[f][jj]
    for (int e = tid; e < DMAX * BK; e += kThreads) {
      const int jj = e % BK, f = e / BK;
      const int64_t j = j0 + jj;
      Xj[f][jj] = (j < kend && f < g.d) ? g.X[j + (int64_t)f * g.ldp] : 0.0;
    }
    // stage OP tile: Bs[jj][c] = OP(c, j0 + jj)
    for (int e = tid; e < BN * BK; e += kThreads) {
      const int c = e % BN, jj = e / BN;
      const int64_t cc = n0 + c, j = j0 + jj;
      Bs[jj][c] = (cc < g.N && j < kend) ? g.OP[cc + j * g.ldop] : 0.0;
    }
    __syncthreads();
    // Gram slab (16 rows x 16 points) per wave, then the RBF values into As
    d4 gr = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int f = 0; f < TF; ++f) {
      if (4 * f < g.d) gr = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[f], Xj[4 * f + (lane >> 4)][lane & 15], gr, 0, 0, 0);
    }
    {
      const int jj = lane & 15;
      const int64_t j = j0 + jj;
      const double nj = (j < kend) ? g.nrm[j] : 0.0;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int il = 16 * wave + (lane >> 4) + 4 * v;
        const int64_t i = m0 + il;
        double kv = 0.0;
        if (i < g.M && j < kend) kv = rbf_value(ni[v], nj, gr[v], g.gamma, g.row0 + i == j);
        As[jj][il] = kv;
      }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double a[2], b[4];
#pragma unroll
      for (int x = 0; x < 2; ++x) a[x] = As[kk + (lane >> 4)][wm * 32 + x * 16 + (lane & 15)];
#pragma unroll
      for (int y = 0; y < 4; ++y) b[y] = Bs[kk + (lane >> 4)][wn * 64 + y * 16 + (lane & 15)];
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], acc[x][y], 0, 0, 0);
    }
  }
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t i = m0 + wm * 32 + x * 16 + (lane >> 4) + 4 * v;
        const int64_t j = n0 + wn * 64 + y * 16 + (lane & 15);
        if (i < g.M && j < g.N) {
          if (g.partial) g.partial[(int64_t)blockIdx.z * g.M * g.N + i + j * g.M] = acc[x][y][v];
          else store_out(g.out, i, j, acc[x][y][v]);
        }
      }
}

// ---- left: X(p, j) = nu X + eta * sum_i U(p, row0+i) K(i, j); M = k (<= 256
// per tile row), N = n, K = rows.  Tile 256 x 64: wave w owns rows 64w..64w+63;
// the 64 points of the tile are the Gram B side, held in registers.
template <int DMAX>
__global__ __launch_bounds__(kThreads) void k_rbf_left_f64(RbfArgs g) {
  constexpr int BM = 256, BN = 64, BK = 16, TF = DMAX / 4;
  __shared__ double As[BK][BM + 2];
  __shared__ double Bs[BK][BN + 2];
  __shared__ double Xi[DMAX][BK + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t m0 = (int64_t)(t % g.tiles_m) * BM;
  const int64_t n0 = (int64_t)(t / g.tiles_m) * BN;

  // Gram B fragments: point (n0 + 16*wave + lane%16), features 4t + lane/16.
  double xb[TF];
  const int64_t jp = n0 + 16 * wave + (lane & 15);
#pragma unroll
  for (int f = 0; f < TF; ++f) {
    const int64_t ff = 4 * f + (lane >> 4);
    xb[f] = (jp < g.N && ff < g.d) ? g.X[jp + ff * g.ldp] : 0.0;
  }
  const double nj = (jp < g.N) ? g.nrm[jp] : 0.0;

  d4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

  for (int64_t i0 = 0; i0 < g.K; i0 += BK) {
    __syncthreads();
    for (int e = tid; e < DMAX * BK; e += kThreads) {
      const int ii = e % BK, f = e / BK;
      const int64_t i = i0 + ii;
      Xi[f][ii] = (i < g.K && f < g.d) ? g.X[(g.row0 + i) + (int64_t)f * g.ldp] : 0.0;
    }
    for (int e = tid; e < BM * BK; e += kThreads) {
      const int p = e % BM, kk = e / BM;
      const int64_t pp = m0 + p, i = i0 + kk;
      As[kk][p] = (pp < g.M && i < g.K) ? g.U[pp + (g.row0 + i) * g.ldu] : 0.0;
    }
    __syncthreads();
    d4 gr = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int f = 0; f < TF; ++f) {
      if (4 * f < g.d) gr = __builtin_amdgcn_mfma_f64_16x16x4f64(Xi[4 * f + (lane >> 4)][lane & 15], xb[f], gr, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int ii = (lane >> 4) + 4 * v;
      const int64_t i = i0 + ii;
      double kv = 0.0;
      if (i < g.K && jp < g.N) kv = rbf_value(g.nrm[g.row0 + i], nj, gr[v], g.gamma, g.row0 + i == jp);
      Bs[ii][16 * wave + (lane & 15)] = kv;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      double a[4], b[4];
#pragma unroll
      for (int x = 0; x < 4; ++x) a[x] = As[kk + (lane >> 4)][wave * 64 + x * 16 + (lane & 15)];
#pragma unroll
      for (int y = 0; y < 4; ++y) b[y] = Bs[kk + (lane >> 4)][y * 16 + (lane & 15)];
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], acc[x][y], 0, 0, 0);
    }
  }
#pragma unroll
  for (int x = 0; x < 4; ++x)
#pragma unroll
    for (int y = 0; y < 4; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t p = m0 + wave * 64 + x * 16 + (lane >> 4) + 4 * v;
        const int64_t j = n0 + y * 16 + (lane & 15);
        if (p < g.M && j < g.N) store_out(g.out, p, j, acc[x][y][v]);
      }
}

__global__ void k_splitk_reduce_rbf(const double* part, int splits, int64_t M, int64_t N, GemmOut o) {
  const int64_t total = M * N;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int z = 0; z < splits; ++z) s += part[(int64_t)z * total + e];
    store_out(o, e % M, e / M, s);
  }
}

template <int DMAX>
void launch_right(hipStream_t st, const RbfArgs& a, dim3 grid) {
  hipLaunchKernelGGL(k_rbf_right_f64<DMAX>, grid, dim3(kThreads), 0, st, a);
}
template <int DMAX>
void launch_left(hipStream_t st, const RbfArgs& a, dim3 grid) {
  hipLaunchKernelGGL(k_rbf_left_f64<DMAX>, grid, dim3(kThreads), 0, st, a);
}

}  // namespace

void rbf_fused_right_f64(hipStream_t st, RbfArgs a, DevBuf* ws) {
  SKEMA_REQUIRE(a.d >= 1 && a.d <= kRbfMaxFeat, kErrUnsupported, "fused RBF: d above 128");
  a.tiles_m = (int)cdiv(a.M, 64);
  const int64_t tiles = (int64_t)a.tiles_m * cdiv(a.N, 128);
  int splits = 1;
  if (tiles < 512 && ws) {
    splits = (int)cdiv(512, tiles);
    const int64_t max_by_k = a.K / 1024;
    if (splits > max_by_k) splits = (int)max_by_k;
    if (splits < 1) splits = 1;
  }
  a.k_per_split = cdiv(cdiv(a.K, splits), 16) * 16;
  splits = (int)cdiv(a.K, a.k_per_split);
  a.partial = nullptr;
  if (splits > 1) {
    ws->reserve((size_t)splits * a.M * a.N * sizeof(double));
    a.partial = ws->as<double>();
  }
  dim3 grid((unsigned)tiles, 1, (unsigned)splits);
  if (a.d <= 32) launch_right<32>(st, a, grid);
  else if (a.d <= 64) launch_right<64>(st, a, grid);
  else launch_right<128>(st, a, grid);
  check_launch("k_rbf_right_f64");
  if (splits > 1) {
    const int blocks = (int)std::min<int64_t>(cdiv(a.M * a.N, 256), 4096);
    hipLaunchKernelGGL(k_splitk_reduce_rbf, dim3(blocks), dim3(256), 0, st, a.partial, splits, a.M, a.N, a.out);
    check_launch("k_splitk_reduce_rbf");
  }
}

void rbf_fused_left_f64(hipStream_t st, RbfArgs a) {
  SKEMA_REQUIRE(a.d >= 1 && a.d <= kRbfMaxFeat, kErrUnsupported, "fused RBF: d above 128");
  a.tiles_m = (int)cdiv(a.M, 256);
  const int64_t tiles = (int64_t)a.tiles_m * cdiv(a.N, 64);
  a.partial = nullptr;
  dim3 grid((unsigned)tiles);
  if (a.d <= 32) launch_left<32>(st, a, grid);
  else if (a.d <= 64) launch_left<64>(st, a, grid);
  else launch_left<128>(st, a, grid);
  check_launch("k_rbf_left_f64");
}

}  // namespace skema

namespace skema {

bool rbf_fused_supported(const Sketch& sk) {
  return !sk.sparse_maps && sk.p.precision == SKEMA_PREC_FP64 && sk.p.nfeat >= 1 && sk.p.nfeat <= kRbfMaxFeat;
}

void update_rbf_window_fused(Sketch& sk, const double* points, int64_t ldp, int64_t row0, int64_t rows) {
  Ctx* c = sk.ctx;
  hipStream_t st = c->stream;
  const int64_t k = sk.k, s = sk.s, m = sk.p.m, n = sk.p.n;
  const double eta = sk.p.eta, nu = sk.p.nu;
  sk.G.reserve((size_t)rows * s * sizeof(double));
  sk.win.reserve((size_t)n * sizeof(double));
  double* nrm = sk.win.as<double>();
  double* G = sk.G.as<double>();
  {
    Timed t(c, "window");
    row_norms2(st, points, n, sk.p.nfeat, ldp, nrm);
  }
  RbfArgs base;
  base.X = points;
  base.ldp = ldp;
  base.d = sk.p.nfeat;
  base.nrm = nrm;
  base.gamma = sk.p.gamma;
  base.row0 = row0;
  {
    Timed t(c, "rbf_left");
    RbfArgs a = base;
    a.M = k; a.N = n; a.K = rows;
    a.U = sk.upsL; a.ldu = sk.ldups;
    a.out.C1 = sk.X.as<double>(); a.out.ldc1 = k; a.out.alpha1 = eta; a.out.beta1 = nu; a.out.nsplit = n;
    a.out.C2 = a.out.C1; a.out.ldc2 = k;
    rbf_fused_left_f64(st, a);
  }
  {
    Timed t(c, "rbf_right");
    RbfArgs a = base;
    a.M = rows; a.N = k + s; a.K = n;
    a.OP = sk.omps.as<double>(); a.ldop = k + s;
    a.out.C1 = sk.Y.as<double>() + row0; a.out.ldc1 = m; a.out.alpha1 = eta; a.out.beta1 = nu; a.out.nsplit = k;
    a.out.C2 = G; a.out.ldc2 = rows; a.out.alpha2 = 1.0; a.out.beta2 = 0.0;
    rbf_fused_right_f64(st, a, &c->gws.buf);
  }
  {
    Timed t(c, "sketch_core");
    gemm_f64(st, 'N', 'N', s, s, rows, eta, sk.phiL + row0 * sk.ldphi, sk.ldphi, G, rows, nu, sk.Z.as<double>(), s,
             &c->gws);
  }
}

}  // namespace skema
