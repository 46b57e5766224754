// Fused Gaussian-RBF window + sketch (FP64).  K_w[i, j] =
// exp(-gamma * ||x_{row0+i} - x_j||^2) is formed tile by tile on the matrix
// cores as ||x_i||^2 + ||x_j||^2 - 2 x_i.x_j (the Gram tile is an MFMA
// product over the d features), exponentiated in registers, handed through
// LDS to the sketch MFMAs, and never written to HBM.
//   left : X       <- nu X + eta * Upsilon[:, idx] K_w     (k x n)
//   right: [y | G]  =  K_w [Omega; Psi]^T                   (w x (k+s))
// The point block a tile needs for its Gram is held in registers for the
// whole K loop (d <= 128); the other side is staged per K step in LDS.
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
//
// One 512-thread block owns 32 window rows and 128*TN columns (all of k+s in
// practice), so every Gram tile is formed once.  Per step of 16 points:
//   Gram   : the 32 x 16 tile as 16x16 MFMA products, wave w taking row tile
//            w%2 over feature quarter w/2 (window-row features in registers),
//   exp    : the four feature quarters summed from LDS, K = exp(-gamma D) -> LDS,
//   sketch : wave w accumulates rows 0..31 x its 16*TN columns; the OP
//            operands come straight from global memory into registers.
// The three stages are software-pipelined one and two steps apart (sketch of
// step t, exp of t+1, Gram of t+2, points of t+3 staged, t+4 loaded), with
// double-buffered LDS and one barrier per step.  Two waves share each SIMD,
// so one wave's loads hide behind the other's MFMAs.
constexpr int kThreadsR = 512;

template <int DMAX, int TN>
__global__ __launch_bounds__(kThreadsR) void k_rbf_right_f64(RbfArgs g) {
  constexpr int BM = 32, BK = 16, FQ = DMAX / 4, TF = FQ / 4;
  constexpr int XS = BK * DMAX / kThreadsR;  // staged point features per thread
  __shared__ double Xs[2][DMAX][BK + 1];
  __shared__ double Gp[2][4][BK][BM + 1];  // [buf][feature quarter][point][row]
  __shared__ double Ks[2][BK][BM + 2];     // [buf][point][row]
  __shared__ double NJ[4][BK];             // point norms of step u in slot u%4
  __shared__ double NI[BM];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, lq = lane >> 4;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t m0 = (int64_t)(t % g.tiles_m) * BM;
  const int split = t / g.tiles_m;
  const int64_t cw0 = (int64_t)blockIdx.y * (128 * TN) + (int64_t)wave * (16 * TN);
  const int64_t kbeg = (int64_t)split * g.k_per_split;
  const int64_t kend = (g.K < kbeg + g.k_per_split) ? g.K : kbeg + g.k_per_split;
  const int64_t steps = (kend - kbeg + BK - 1) / BK;

  // Gram A operand: window row m0 + 16*(wave%2) + l16, features FQ*(wave/2) + 4f + lq.
  const int rt = wave & 1, fq = wave >> 1;
  double xa[TF];
  {
    const int64_t i = m0 + 16 * rt + l16;
#pragma unroll
    for (int f = 0; f < TF; ++f) {
      const int64_t ff = FQ * fq + 4 * f + lq;
      xa[f] = (i < g.M && ff < g.d) ? g.X[(g.row0 + i) + ff * g.ldp] : 0.0;
    }
  }
  if (tid < BM) NI[tid] = (m0 + tid < g.M) ? g.nrm[g.row0 + m0 + tid] : 0.0;

  double xr[XS];
  double njr = 0.0;
  auto load_points = [&](int64_t u) {  // features and norms of the 16 points of step u
    const int64_t j0 = kbeg + u * BK;
#pragma unroll
    for (int q = 0; q < XS; ++q) {
      const int e = tid + q * kThreadsR, jj = e & (BK - 1), f = e / BK;
      const int64_t j = j0 + jj;
      xr[q] = (j < kend && f < g.d) ? g.X[j + (int64_t)f * g.ldp] : 0.0;
    }
    const int64_t j = j0 + tid;
    njr = (tid < BK && j < kend) ? g.nrm[j] : 0.0;
  };
  auto store_points = [&](int b, int64_t u) {  // staged step u -> Xs[b], NJ[u % 4]
#pragma unroll
    for (int q = 0; q < XS; ++q) {
      const int e = tid + q * kThreadsR;
      Xs[b][e / BK][e & (BK - 1)] = xr[q];
    }
    if (tid < BK) NJ[u & 3][tid] = njr;
  };
  auto gram = [&](int b) {  // partial Gram of buffer b -> Gp[b][fq]
    d4 gr = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int f = 0; f < TF; ++f)
      gr = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[f], Xs[b][FQ * fq + 4 * f + lq][l16], gr, 0, 0, 0);
#pragma unroll
    for (int v = 0; v < 4; ++v) Gp[b][fq][l16][16 * rt + lq + 4 * v] = gr[v];
  };
  auto rbf_exp = [&](int b, int64_t u) {  // Gp[b] -> Ks[b] for step u (one element per thread)
    const int jj = tid / BM, ii = tid % BM;
    const int64_t i = m0 + ii, j = kbeg + u * BK + jj;
    double kv = 0.0;
    if (i < g.M && j < kend) {
      const double gs = (Gp[b][0][jj][ii] + Gp[b][1][jj][ii]) + (Gp[b][2][jj][ii] + Gp[b][3][jj][ii]);
      kv = rbf_value(NI[ii], NJ[u & 3][jj], gs, g.gamma, g.row0 + i == j);
    }
    Ks[b][jj][ii] = kv;
  };
  double op[TN][4];
  auto load_op = [&](double (&dst)[TN][4], int64_t u) {
    const int64_t j0 = kbeg + u * BK + lq;
    const double* base = g.OP + (cw0 + l16) + j0 * g.ldop;
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
#pragma unroll
      for (int y = 0; y < TN; ++y) {
        const bool ok = (cw0 + l16 + 16 * y < g.N) && (j0 + 4 * kk < kend);
        dst[y][kk] = ok ? base[16 * y + 4 * kk * g.ldop] : 0.0;
      }
  };

  d4 acc[2][TN];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < TN; ++y) acc[x][y] = d4{0.0, 0.0, 0.0, 0.0};

  // prologue: points of steps 0, 1 staged; Gram 0, 1; exp 0; points 2 staged.
  load_points(0); store_points(0, 0);
  load_points(1); store_points(1, 1);
  load_op(op, 0);
  __syncthreads();
  gram(0);
  gram(1);
  load_points(2);
  __syncthreads();
  rbf_exp(0, 0);
  store_points(0, 2);
  load_points(3);
  __syncthreads();

  for (int64_t u = 0; u < steps; ++u) {
    const int b = (int)(u & 1), nb = b ^ 1;
#pragma unroll
    for (int kk = 0; kk < 4; ++kk) {
      double a[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) a[x] = Ks[b][4 * kk + lq][16 * x + l16];
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < TN; ++y) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], op[y][kk], acc[x][y], 0, 0, 0);
    }
    // the next step's OP operands load while this wave forms K and the
    // block's other wave on the SIMD keeps the matrix core busy
    if (u + 1 < steps) load_op(op, u + 1);
    if (u + 1 < steps) rbf_exp(nb, u + 1);
    if (u + 2 < steps) gram(b);           // points of step u+2 sit in Xs[b]
    store_points(nb, u + 3);              // points of step u+3 -> Xs[nb]
    load_points(u + 4);
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < TN; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t i = m0 + 16 * x + lq + 4 * v;
        const int64_t j = cw0 + 16 * y + l16;
        if (i < g.M && j < g.N) {
          if (g.partial) g.partial[(int64_t)split * g.M * g.N + i + j * g.M] = acc[x][y][v];
          else store_out(g.out, i, j, acc[x][y][v]);
        }
      }
}

// ---- left: X(p, j) = nu X + eta * sum_i U(p, row0+i) K(i, j); M = k, N = n,
// K = rows.  One block owns 64 points (wave w: points 16w..16w+15) and 16*TP
// map rows.  Per step of 16 window rows each wave forms its 16x16 Gram tile
// over all features (its points' features held in registers), exponentiates
// it in registers, and uses it directly as the B operand of the sketch MFMAs:
// the f64 16x16x4 result layout (lane l: column l%16, rows l/16 + 4v) is the
// B-operand layout of K-substep v (lane l: column l%16, K index 4v + l/16), so
// K never leaves the registers.  Window-row features and the Upsilon slice
// are double-buffered in LDS with register prefetch, one barrier per step.
template <int DMAX, int TP>
__global__ __launch_bounds__(kThreads, 1) void k_rbf_left_f64(RbfArgs g) {
  constexpr int BN = 64, BK = 16, TF = DMAX / 4;
  constexpr int XS = BK * DMAX / kThreads;
  constexpr int PR = 16 * TP;                                  // map rows per block
  constexpr int US = (PR * BK + kThreads - 1) / kThreads;      // staged U per thread
  __shared__ double Xs[2][DMAX][BK + 1];
  __shared__ double Us[2][BK][PR + 2];
  __shared__ double NI[2][BK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, lq = lane >> 4;
  const int t = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t n0 = (int64_t)t * BN;
  const int64_t p0 = (int64_t)blockIdx.y * PR;
  const int64_t steps = (g.K + BK - 1) / BK;

  const int64_t jp = n0 + 16 * wave + l16;  // this lane's point (B-operand column)
  double xb[TF];
#pragma unroll
  for (int f = 0; f < TF; ++f) {
    const int64_t ff = 4 * f + lq;
    xb[f] = (jp < g.N && ff < g.d) ? g.X[jp + ff * g.ldp] : 0.0;
  }
  const double nj = (jp < g.N) ? g.nrm[jp] : 0.0;

  double xr[XS], ur[US], nir = 0.0;
  auto load_step = [&](int64_t u) {
    const int64_t i0 = u * BK;
#pragma unroll
    for (int q = 0; q < XS; ++q) {
      const int e = tid + q * kThreads, ii = e & (BK - 1), f = e / BK;
      const int64_t i = i0 + ii;
      xr[q] = (i < g.K && f < g.d) ? g.X[(g.row0 + i) + (int64_t)f * g.ldp] : 0.0;
    }
#pragma unroll
    for (int q = 0; q < US; ++q) {
      const int e = tid + q * kThreads, pp = e % PR, ii = e / PR;
      const int64_t p = p0 + pp, i = i0 + ii;
      ur[q] = (e < PR * BK && p < g.M && i < g.K) ? g.U[p + (g.row0 + i) * g.ldu] : 0.0;
    }
    const int64_t i = i0 + tid;
    nir = (tid < BK && i < g.K) ? g.nrm[g.row0 + i] : 0.0;
  };
  auto store_step = [&](int b) {
#pragma unroll
    for (int q = 0; q < XS; ++q) {
      const int e = tid + q * kThreads;
      Xs[b][e / BK][e & (BK - 1)] = xr[q];
    }
#pragma unroll
    for (int q = 0; q < US; ++q) {
      const int e = tid + q * kThreads;
      if (e < PR * BK) Us[b][e / PR][e % PR] = ur[q];
    }
    if (tid < BK) NI[b][tid] = nir;
  };

  d4 acc[TP];
#pragma unroll
  for (int x = 0; x < TP; ++x) acc[x] = d4{0.0, 0.0, 0.0, 0.0};

  load_step(0);
  store_step(0);
  load_step(1);
  __syncthreads();
  for (int64_t u = 0; u < steps; ++u) {
    const int b = (int)(u & 1);
    const int64_t i0 = u * BK;
    d4 gr = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int f = 0; f < TF; ++f) gr = __builtin_amdgcn_mfma_f64_16x16x4f64(Xs[b][4 * f + lq][l16], xb[f], gr, 0, 0, 0);
    double kv[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int ii = lq + 4 * v;
      const int64_t i = i0 + ii;
      kv[v] = (i < g.K && jp < g.N) ? rbf_value(NI[b][ii], nj, gr[v], g.gamma, g.row0 + i == jp) : 0.0;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int x = 0; x < TP; ++x)
        acc[x] = __builtin_amdgcn_mfma_f64_16x16x4f64(Us[b][4 * v + lq][16 * x + l16], kv[v], acc[x], 0, 0, 0);
    store_step(b ^ 1);
    load_step(u + 2);
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < TP; ++x)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t p = p0 + 16 * x + lq + 4 * v;
      if (p < g.M && jp < g.N) store_out(g.out, p, jp, acc[x][v]);
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

// Column tiles per wave of the right kernel / map-row tiles of the left one.
constexpr int kTnChoices[] = {1, 2, 3, 4, 5};
constexpr int kTpChoices[] = {1, 2, 4, 8, 13};

template <int DMAX, int TN>
void launch_right_tn(hipStream_t st, const RbfArgs& a, dim3 grid) {
  hipLaunchKernelGGL((k_rbf_right_f64<DMAX, TN>), grid, dim3(kThreadsR), 0, st, a);
}
template <int DMAX>
void launch_right(hipStream_t st, const RbfArgs& a, dim3 grid, int tn) {
  switch (tn) {
    case 1: launch_right_tn<DMAX, 1>(st, a, grid); break;
    case 2: launch_right_tn<DMAX, 2>(st, a, grid); break;
    case 3: launch_right_tn<DMAX, 3>(st, a, grid); break;
    case 4: launch_right_tn<DMAX, 4>(st, a, grid); break;
    default: launch_right_tn<DMAX, 5>(st, a, grid); break;
  }
}
template <int DMAX, int TP>
void launch_left_tp(hipStream_t st, const RbfArgs& a, dim3 grid) {
  hipLaunchKernelGGL((k_rbf_left_f64<DMAX, TP>), grid, dim3(kThreads), 0, st, a);
}
template <int DMAX>
void launch_left(hipStream_t st, const RbfArgs& a, dim3 grid, int tp) {
  switch (tp) {
    case 1: launch_left_tp<DMAX, 1>(st, a, grid); break;
    case 2: launch_left_tp<DMAX, 2>(st, a, grid); break;
    case 4: launch_left_tp<DMAX, 4>(st, a, grid); break;
    case 8: launch_left_tp<DMAX, 8>(st, a, grid); break;
    default: launch_left_tp<DMAX, 13>(st, a, grid); break;
  }
}

template <size_t L>
int pick_tiles(const int (&choices)[L], int64_t need) {
  for (int c : choices)
    if (c >= need) return c;
  return choices[L - 1];
}

}  // namespace

void rbf_fused_right_f64(hipStream_t st, RbfArgs a, DevBuf* ws) {
  SKEMA_REQUIRE(a.d >= 1 && a.d <= kRbfMaxFeat, kErrUnsupported, "fused RBF: d above 128");
  // columns: 8 waves x 16*TN each per block, more blocks along y if needed
  const int tn = pick_tiles(kTnChoices, cdiv(cdiv(a.N, 16), 8));
  const int64_t col_blocks = cdiv(a.N, 128 * tn);
  a.tiles_m = (int)cdiv(a.M, 32);
  // Split the points over blocks: one 512-thread block fills a CU, so the
  // time is about (rounds of 256 blocks) x (points per split), plus the
  // partial-sum traffic of each extra split (M x N doubles written and read,
  // worth about M N / 65536 points of block time).  Take the split count
  // that minimises it, splits at least 1024 points long.
  const int64_t per_split = a.tiles_m * col_blocks;
  int64_t splits = 1;
  if (ws) {
    int64_t best = INT64_MAX;
    const int64_t max_splits = std::max<int64_t>(1, std::min<int64_t>(64, a.K / 1024));
    for (int64_t sp = 1; sp <= max_splits; ++sp) {
      const int64_t cost = cdiv(per_split * sp, 256) * cdiv(a.K, sp) + (sp > 1 ? sp * (a.M * a.N / 65536 + 1) : 0);
      if (cost < best) { best = cost; splits = sp; }
    }
  }
  a.k_per_split = cdiv(cdiv(a.K, splits), 16) * 16;
  splits = cdiv(a.K, a.k_per_split);
  a.partial = nullptr;
  if (splits > 1) {
    ws->reserve((size_t)splits * a.M * a.N * sizeof(double));
    a.partial = ws->as<double>();
  }
  const dim3 grid((unsigned)(a.tiles_m * splits), (unsigned)col_blocks);
  if (a.d <= 32) launch_right<32>(st, a, grid, tn);
  else if (a.d <= 64) launch_right<64>(st, a, grid, tn);
  else launch_right<128>(st, a, grid, tn);
  check_launch("k_rbf_right_f64");
  if (splits > 1) {
    const int blocks = (int)std::min<int64_t>(cdiv(a.M * a.N, 256), 4096);
    hipLaunchKernelGGL(k_splitk_reduce_rbf, dim3(blocks), dim3(256), 0, st, a.partial, (int)splits, a.M, a.N, a.out);
    check_launch("k_splitk_reduce_rbf");
  }
}

void rbf_fused_left_f64(hipStream_t st, RbfArgs a) {
  SKEMA_REQUIRE(a.d >= 1 && a.d <= kRbfMaxFeat, kErrUnsupported, "fused RBF: d above 128");
  const int tp = pick_tiles(kTpChoices, cdiv(a.M, 16));
  const dim3 grid((unsigned)cdiv(a.N, 64), (unsigned)cdiv(a.M, 16 * tp));
  a.partial = nullptr;
  if (a.d <= 32) launch_left<32>(st, a, grid, tp);
  else if (a.d <= 64) launch_left<64>(st, a, grid, tp);
  else launch_left<128>(st, a, grid, tp);
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
