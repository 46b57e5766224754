// Adjoint of the fused RBF window + Gauss sketch.
//
// Loss over the n points X (n x d), K = exp(-gamma D), D_ij = ||x_i - x_j||^2:
//   L = ||K Omega^T||_F^2 + ||Phi K Psi^T||_F^2,   Y = K Omega^T, W = Phi K Psi^T.
// dL/dK = Gbar = 2 (Y Omega + Phi^T W Psi) = 2 B OP with B = [Y | Phi^T W]
// (n x h) and OP = [Omega; Psi] (h x n), h = k + s.  With H = (Gbar + Gbar^T)
// o K (elementwise):
//   dL/dx_a   = -2 gamma (sum_j H_aj x_a - sum_j H_aj x_j)
//   dL/dgamma = -1/2 sum_ij H_ij D_ij.
// Gbar + Gbar^T = 2 F R^T with F = [B | OP^T] (n x 2h) and R the same rows
// with the two column halves swapped, so a tile (I, J) of H is
// 2 (F_I R_J^T) o K_IJ: one GEMM tile with inner dimension 2h, the RBF tile
// recomputed from its Gram on the matrix cores, and the product H X_J folded
// into the same tile loop.  Neither K nor H is ever written out.
#include <vector>

#include "adjoint.hpp"
#include "linalg.hpp"
#include "rbf.hpp"
#include "rbf_fused.hpp"
#include "util.hpp"

namespace skema {
namespace {

constexpr int kThreads = 512;

struct AdjArgs {
  const double* X;
  int64_t ldp, d, n;
  const double* nrm;
  double gamma;
  const double* F;  // n x 2h, ld = n
  int64_t h;
  int64_t jblocks_per_split;
  double* part_hx;  // [split][n x d]
  double* part_rs;  // [split][n]
  double* part_dg;  // [block]
};

// Tile pass over (I, J) pairs of 64 rows x 64 points (512 threads).  The block owns I and
// walks the J tiles of its split; everything is formed transposed (rows j,
// columns i) because the f64 16x16x4 result layout (lane l: column l%16,
// rows l/16 + 4v) then serves directly as the B operand of the H X product
// with K = j.  Per J tile:
//   S^T = R_J F_I^T over the 2h columns of F, streamed through
//         double-buffered LDS chunks of 16 (one barrier per chunk); wave w
//         owns j tile w%4 and i tiles 2(w/4), 2(w/4)+1,
//   the Gram^T tile over the d features (X_I in registers, X_J in LDS),
//   H^T = 2 S^T o K^T with its row sums and sum H o D in registers,
//   H^T through LDS, and (H X)^T += X_J^T H^T with wave w owning feature
//   tiles w%4 (+4, ...) and the same two i tiles, accumulated over all J.
// Neither K nor H is ever written to HBM.
template <int DMAX>
__global__ __launch_bounds__(kThreads) void k_rbf_adjoint_f64(AdjArgs g) {
  constexpr int BM = 64, BJ = 64, BKC = 16, TF = DMAX / 4, FT = DMAX / 16;
  constexpr int FTW = (FT + 3) / 4;  // feature tiles per wave in the H X product
  __shared__ double Rc[2][BKC][BJ + 2];
  __shared__ double Fc[2][BKC][BM + 2];
  __shared__ double XJ[DMAX][BJ + 1];
  __shared__ double Hs[BJ][BM + 2];
  __shared__ double red[BM];
  __shared__ double redg[8];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int l16 = lane & 15, lq = lane >> 4;
  const int jt = wave & 3, ih = wave >> 2;
  const int64_t i0 = (int64_t)blockIdx.x * BM;
  const int split = blockIdx.y;
  const int64_t nblk = cdiv(g.n, BJ);
  const int64_t jb0 = (int64_t)split * g.jblocks_per_split;
  const int64_t jb1 = (jb0 + g.jblocks_per_split < nblk) ? jb0 + g.jblocks_per_split : nblk;
  const int64_t two_h = 2 * g.h;
  const int nch = (int)((two_h + BKC - 1) / BKC);

  // Gram B operand: X_I rows i = i0 + 16 it + l16 (it = 2 ih, 2 ih + 1),
  // features 4t + lq; held in registers up to 64 features, re-read from
  // global memory (cache-resident, once per J tile) above that.
  constexpr int TFR = DMAX <= 64 ? TF : 1;
  double xi[2][TFR];
  double ni[2];
  int64_t irow[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int64_t i = i0 + 16 * (2 * ih + u) + l16;
    irow[u] = i;
    if constexpr (DMAX <= 64) {
#pragma unroll
      for (int t = 0; t < TF; ++t) {
        const int64_t f = 4 * t + lq;
        xi[u][t] = (i < g.n && f < g.d) ? g.X[i + f * g.ldp] : 0.0;
      }
    }
    ni[u] = (i < g.n) ? g.nrm[i] : 0.0;
  }
  auto xi_at = [&](int u, int t) -> double {
    if constexpr (DMAX <= 64) {
      return xi[u][t];
    } else {
      const int64_t f = 4 * t + lq;
      return (irow[u] < g.n && f < g.d) ? g.X[irow[u] + f * g.ldp] : 0.0;
    }
  };
  d4 hx[FTW][2];
#pragma unroll
  for (int q = 0; q < FTW; ++q)
#pragma unroll
    for (int u = 0; u < 2; ++u) hx[q][u] = d4{0.0, 0.0, 0.0, 0.0};
  double rs[2] = {0.0, 0.0};
  double dg = 0.0;

  // chunk c of R_J (columns of F with halves swapped) and F_I: 4 doubles per thread
  double cr[2], cf[2];
  auto load_chunk = [&](int64_t j0, int t) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int e = tid + q * kThreads;  // 0 .. 1023
      const int cc = e / 64, r = e % 64;
      const int64_t c = (int64_t)t * BKC + cc;
      const int64_t crs = (c < g.h) ? c + g.h : c - g.h;
      const int64_t j = j0 + r, i = i0 + r;
      cr[q] = (c < two_h && j < g.n) ? g.F[j + crs * g.n] : 0.0;
      cf[q] = (c < two_h && i < g.n) ? g.F[i + c * g.n] : 0.0;
    }
  };
  auto store_chunk = [&](int b) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int e = tid + q * kThreads;
      Rc[b][e / 64][e % 64] = cr[q];
      Fc[b][e / 64][e % 64] = cf[q];
    }
  };

  for (int64_t jb = jb0; jb < jb1; ++jb) {
    const int64_t j0 = jb * BJ;
    __syncthreads();  // previous tile's readers of XJ / Hs / Rc / Fc are done
    for (int e = tid; e < DMAX * BJ; e += kThreads) {
      const int jj = e % BJ, f = e / BJ;
      const int64_t j = j0 + jj;
      XJ[f][jj] = (j < g.n && f < g.d) ? g.X[j + (int64_t)f * g.ldp] : 0.0;
    }
    load_chunk(j0, 0);
    store_chunk(0);
    if (nch > 1) load_chunk(j0, 1);
    __syncthreads();

    d4 sa[2] = {d4{0.0, 0.0, 0.0, 0.0}, d4{0.0, 0.0, 0.0, 0.0}};
    for (int t = 0; t < nch; ++t) {
      const int b = t & 1;
#pragma unroll
      for (int kk = 0; kk < BKC; kk += 4) {
        const double a = Rc[b][kk + lq][16 * jt + l16];
#pragma unroll
        for (int u = 0; u < 2; ++u)
          sa[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Fc[b][kk + lq][16 * (2 * ih + u) + l16], sa[u], 0, 0, 0);
      }
      if (t + 1 < nch) {
        store_chunk(b ^ 1);
        if (t + 2 < nch) load_chunk(j0, t + 2);
      }
      __syncthreads();
    }

    // Gram^T (rows j of tile jt, columns i of tiles 2ih, 2ih+1)
    d4 gr[2] = {d4{0.0, 0.0, 0.0, 0.0}, d4{0.0, 0.0, 0.0, 0.0}};
#pragma unroll
    for (int t = 0; t < TF; ++t) {
      if (4 * t < g.d) {
        const double a = XJ[4 * t + lq][16 * jt + l16];
#pragma unroll
        for (int u = 0; u < 2; ++u) gr[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, xi_at(u, t), gr[u], 0, 0, 0);
      }
    }
    // H^T = 2 S^T o K^T, row sums over j, sum H o D; H^T -> LDS
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int il = 16 * (2 * ih + u) + l16;
      const int64_t i = i0 + il;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int jl = 16 * jt + lq + 4 * v;
        const int64_t j = j0 + jl;
        double hv = 0.0;
        if (i < g.n && j < g.n) {
          double D = ni[u] + g.nrm[j] - 2.0 * gr[u][v];
          D = (D > 0.0 && i != j) ? D : 0.0;
          const double kv = (i == j) ? 1.0 : exp(-g.gamma * D);
          hv = 2.0 * sa[u][v] * kv;
          dg += hv * D;
        }
        rs[u] += hv;
        Hs[jl][il] = hv;
      }
    }
    __syncthreads();
    // (H X)^T(f, i) += sum_j X_J(j, f) H^T(j, i)
#pragma unroll
    for (int kk = 0; kk < BJ; kk += 4) {
      double bh[2];
#pragma unroll
      for (int u = 0; u < 2; ++u) bh[u] = Hs[kk + lq][16 * (2 * ih + u) + l16];
#pragma unroll
      for (int q = 0; q < FTW; ++q) {
        const int ft = jt + 4 * q;
        if (ft < FT) {
          const double a = XJ[16 * ft + l16][kk + lq];
#pragma unroll
          for (int u = 0; u < 2; ++u) hx[q][u] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bh[u], hx[q][u], 0, 0, 0);
        }
      }
    }
  }

  // (H X)^T result: lane l holds column i (l%16) and feature rows 16 ft + l/16 + 4v
#pragma unroll
  for (int q = 0; q < FTW; ++q) {
    const int ft = jt + 4 * q;
    if (ft >= FT) continue;
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t i = i0 + 16 * (2 * ih + u) + l16;
        const int64_t f = 16 * ft + lq + 4 * v;
        if (i < g.n && f < g.d) g.part_hx[(int64_t)split * g.n * g.d + i + f * g.n] = hx[q][u][v];
      }
  }
  // row sums: over the 4 lane groups of a column, then over the 4 j-tile waves
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    double r = rs[u];
    r += __shfl_xor(r, 16);
    r += __shfl_xor(r, 32);
    rs[u] = r;
  }
  __syncthreads();
  if (jt == 0 && lq == 0) {
#pragma unroll
    for (int u = 0; u < 2; ++u) red[16 * (2 * ih + u) + l16] = 0.0;
  }
  __syncthreads();
  for (int w = 0; w < 4; ++w) {  // fixed order over the j-tile waves
    if (jt == w && lq == 0) {
#pragma unroll
      for (int u = 0; u < 2; ++u) red[16 * (2 * ih + u) + l16] += rs[u];
    }
    __syncthreads();
  }
  if (tid < BM && i0 + tid < g.n) g.part_rs[(int64_t)split * g.n + i0 + tid] = red[tid];
  // sum H o D over the block
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) dg += __shfl_xor(dg, o);
  if (lane == 0) redg[wave] = dg;
  __syncthreads();
  if (tid == 0) {
    double sdg = 0.0;
    for (int w = 0; w < 8; ++w) sdg += redg[w];
    g.part_dg[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = sdg;
  }
}

// dX = -2 gamma (rowsum o X - HX), deterministic sums over the splits.
__global__ void k_adjoint_finish(const double* X, int64_t ldp, int64_t n, int64_t d, int splits,
                                 const double* part_hx, const double* part_rs, double gamma, double* dX, int64_t ldg) {
  const int64_t total = n * d;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % n, f = e / n;
    double hx = 0.0, rs = 0.0;
    for (int s = 0; s < splits; ++s) {
      hx += part_hx[(int64_t)s * total + e];
      rs += part_rs[(int64_t)s * n + i];
    }
    dX[i + f * ldg] = -2.0 * gamma * (rs * X[i + f * ldp] - hx);
  }
}

// Logical (rows x cols) view of a dense map into dst (ld = ldd).
void logical_copy(hipStream_t st, const Map* m, double* dst, int64_t ldd) {
  if (!m->transposed) copy_matrix(st, m->nrow, m->ncol, m->dense.as<double>(), m->nrow, dst, ldd);
  else transpose(st, m->nrow, m->ncol, m->dense.as<double>(), m->nrow, dst, ldd);
}

}  // namespace

void rbf_sketch_loss_grad(Ctx* c, const double* X, int64_t n, int64_t d, int64_t ldp, double gamma, const Map* omega,
                          const Map* phi, const Map* psi, int64_t window, int precision, double* loss, double* dX,
                          int64_t ldg, double* dgamma) {
  SKEMA_REQUIRE(precision == SKEMA_PREC_FP64, kErrUnsupported, "rbf_loss_grad: only the FP64 form is built");
  SKEMA_REQUIRE(!omega->sparse && !phi->sparse && !psi->sparse, kErrUnsupported, "rbf_loss_grad: dense maps only");
  SKEMA_REQUIRE(d <= kRbfMaxFeat, kErrUnsupported, "rbf_loss_grad: d above 128");
  const int64_t k = omega->lrows(), s = phi->lrows();
  SKEMA_REQUIRE(omega->lcols() == n && phi->lcols() == n && psi->lcols() == n && psi->lrows() == s, kErrShape,
                "rbf_loss_grad: maps must be Omega (k x n), Phi (s x n), Psi (s x n)");
  hipStream_t st = c->stream;
  const int64_t h = k + s;
  DevBuf nrm(n * sizeof(double)), OP((size_t)h * n * sizeof(double)), F((size_t)n * 2 * h * sizeof(double)),
      PhiL((size_t)s * n * sizeof(double)), W((size_t)s * s * sizeof(double)), sc(2 * sizeof(double)), red;
  row_norms2(st, X, n, d, ldp, nrm.as<double>());
  logical_copy(st, omega, OP.as<double>(), h);
  logical_copy(st, psi, OP.as<double>() + k, h);
  logical_copy(st, phi, PhiL.as<double>(), s);
  double* Fp = F.as<double>();
  double* Y = Fp;           // F[:, 0:k]   = Y = K Omega^T
  double* G = Fp + k * n;   // F[:, k:h]   = G = K Psi^T, then Phi^T W
  // forward: [Y | G] window by window through the fused generator
  for (int64_t r0 = 0; r0 < n; r0 += window) {
    const int64_t rows = std::min(window, n - r0);
    RbfArgs a;
    a.X = X; a.ldp = ldp; a.d = d; a.nrm = nrm.as<double>(); a.gamma = gamma; a.row0 = r0;
    a.M = rows; a.N = h; a.K = n; a.OP = OP.as<double>(); a.ldop = h;
    a.out.C1 = Y + r0; a.out.ldc1 = n; a.out.alpha1 = 1.0; a.out.beta1 = 0.0; a.out.nsplit = k;
    a.out.C2 = G + r0; a.out.ldc2 = n; a.out.alpha2 = 1.0; a.out.beta2 = 0.0;
    rbf_fused_right_f64(st, a, &c->gws.buf);
  }
  // W = Phi G (s x s); loss = ||Y||^2 + ||W||^2
  gemm_f64(st, 'N', 'N', s, s, n, 1.0, PhiL.as<double>(), s, G, n, 0.0, W.as<double>(), s, &c->gws);
  sum_squares(st, n, k, Y, n, sc.as<double>(), &red);
  sum_squares(st, s, s, W.as<double>(), s, sc.as<double>() + 1, &red);
  double hs[2];
  SKEMA_HIP(hipMemcpyAsync(hs, sc.p, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
  // B = [Y | Phi^T W] in place of G; F[:, h:2h] = OP^T
  gemm_f64(st, 'T', 'N', n, s, s, 1.0, PhiL.as<double>(), s, W.as<double>(), s, 0.0, G, n, &c->gws);
  transpose(st, h, n, OP.as<double>(), h, Fp + h * n, n);
  SKEMA_HIP(hipStreamSynchronize(st));
  *loss = hs[0] + hs[1];
  if (!dX) {
    *dgamma = 0.0;
    return;
  }
  // tile pass
  // one 512-thread block per CU: split the J tiles until the grid covers
  // the 256 CUs
  const int64_t iblocks = cdiv(n, 64), jblocks = cdiv(n, 64);
  int splits = (int)std::max<int64_t>(1, std::min<int64_t>(jblocks, cdiv(256, iblocks)));
  const int64_t per = cdiv(jblocks, splits);
  splits = (int)cdiv(jblocks, per);
  DevBuf phx((size_t)splits * n * d * sizeof(double)), prs((size_t)splits * n * sizeof(double)),
      pdg((size_t)splits * iblocks * sizeof(double));
  AdjArgs g;
  g.X = X; g.ldp = ldp; g.d = d; g.n = n; g.nrm = nrm.as<double>(); g.gamma = gamma; g.F = Fp; g.h = h;
  g.jblocks_per_split = per; g.part_hx = phx.as<double>(); g.part_rs = prs.as<double>(); g.part_dg = pdg.as<double>();
  dim3 grid((unsigned)iblocks, (unsigned)splits);
  {
    Timed t(c, "rbf_adjoint");
    if (d <= 32) hipLaunchKernelGGL(k_rbf_adjoint_f64<32>, grid, dim3(kThreads), 0, st, g);
    else if (d <= 64) hipLaunchKernelGGL(k_rbf_adjoint_f64<64>, grid, dim3(kThreads), 0, st, g);
    else hipLaunchKernelGGL(k_rbf_adjoint_f64<128>, grid, dim3(kThreads), 0, st, g);
    check_launch("k_rbf_adjoint_f64");
  }
  const int nb = (int)std::min<int64_t>(cdiv(n * d, 256), 4096);
  hipLaunchKernelGGL(k_adjoint_finish, dim3(nb), dim3(256), 0, st, X, ldp, n, d, splits, phx.as<double>(),
                     prs.as<double>(), gamma, dX, ldg);
  check_launch("k_adjoint_finish");
  std::vector<double> hdg((size_t)splits * iblocks);
  SKEMA_HIP(hipMemcpyAsync(hdg.data(), pdg.p, hdg.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  SKEMA_HIP(hipStreamSynchronize(st));
  double sdg = 0.0;
  for (double v : hdg) sdg += v;
  *dgamma = -0.5 * sdg;
}

}  // namespace skema
