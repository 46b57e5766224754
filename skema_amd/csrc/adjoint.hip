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

constexpr int kThreads = 256;

struct AdjArgs {
  const double* X;
  int64_t ldp, d, n;
  const double* nrm;
  double gamma;
  const double* F;  // n x 2h, ld = n
  int64_t h;
  int64_t jblocks_per_split;
  double* part_hx;  // [split][n x d]
  double* part_rs;  // [split][2][n]
  double* part_dg;  // [block]
};

// Tile 64 x 64 (rows I of H, points J), 2 x 2 waves of 32 x 32.
template <int DMAX>
__global__ __launch_bounds__(kThreads) void k_rbf_adjoint_f64(AdjArgs g) {
  constexpr int BM = 64, BN = 64, BKC = 16, TF = DMAX / 4, ND = DMAX / 32;
  __shared__ double Fa[BKC][BM + 2];
  __shared__ double Fb[BKC][BN + 2];
  __shared__ double Xj[DMAX][BN + 1];
  __shared__ double Hs[BN][BM + 2];
  __shared__ double red[4];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int64_t m0 = (int64_t)blockIdx.x * BM;
  const int split = blockIdx.y;
  const int64_t nblk = cdiv(g.n, BN);
  const int64_t jb0 = (int64_t)split * g.jblocks_per_split;
  const int64_t jb1 = (jb0 + g.jblocks_per_split < nblk) ? jb0 + g.jblocks_per_split : nblk;
  const int64_t two_h = 2 * g.h;

  // Gram A fragments of the block's rows, and their norms.
  double xa[2][TF];
  double ni[2][4];
#pragma unroll
  for (int x = 0; x < 2; ++x) {
    const int64_t i = m0 + wm * 32 + x * 16 + (lane & 15);
#pragma unroll
    for (int t = 0; t < TF; ++t) {
      const int64_t f = 4 * t + (lane >> 4);
      xa[x][t] = (i < g.n && f < g.d) ? g.X[i + f * g.ldp] : 0.0;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t r = m0 + wm * 32 + x * 16 + (lane >> 4) + 4 * v;
      ni[x][v] = (r < g.n) ? g.nrm[r] : 0.0;
    }
  }
  d4 hx[2][ND];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < ND; ++y) hx[x][y] = d4{0.0, 0.0, 0.0, 0.0};
  double rs[2][4] = {};
  double dg = 0.0;

  for (int64_t jb = jb0; jb < jb1; ++jb) {
    const int64_t n0 = jb * BN;
    __syncthreads();
    for (int e = tid; e < DMAX * BN; e += kThreads) {
      const int jj = e % BN, f = e / BN;
      const int64_t j = n0 + jj;
      Xj[f][jj] = (j < g.n && f < g.d) ? g.X[j + (int64_t)f * g.ldp] : 0.0;
    }
    __syncthreads();
    // Gram and S = F_I R_J^T on the same 2 x 2 wave tiling.
    d4 gr[2][2], sa[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
      for (int y = 0; y < 2; ++y) { gr[x][y] = d4{0.0, 0.0, 0.0, 0.0}; sa[x][y] = gr[x][y]; }
#pragma unroll
    for (int t = 0; t < TF; ++t) {
      if (4 * t < g.d) {
#pragma unroll
        for (int y = 0; y < 2; ++y) {
          const double b = Xj[4 * t + (lane >> 4)][wn * 32 + y * 16 + (lane & 15)];
#pragma unroll
          for (int x = 0; x < 2; ++x) gr[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(xa[x][t], b, gr[x][y], 0, 0, 0);
        }
      }
    }
    for (int64_t c0 = 0; c0 < two_h; c0 += BKC) {
      __syncthreads();
      for (int e = tid; e < BKC * BM; e += kThreads) {
        const int i = e % BM, cc = e / BM;
        const int64_t gi = m0 + i, c = c0 + cc;
        Fa[cc][i] = (gi < g.n && c < two_h) ? g.F[gi + c * g.n] : 0.0;
        const int64_t gj = n0 + i;
        const int64_t cr = (c < g.h) ? c + g.h : c - g.h;
        Fb[cc][i] = (gj < g.n && c < two_h) ? g.F[gj + cr * g.n] : 0.0;
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < BKC; kk += 4) {
        double a[2], b[2];
#pragma unroll
        for (int x = 0; x < 2; ++x) a[x] = Fa[kk + (lane >> 4)][wm * 32 + x * 16 + (lane & 15)];
#pragma unroll
        for (int y = 0; y < 2; ++y) b[y] = Fb[kk + (lane >> 4)][wn * 32 + y * 16 + (lane & 15)];
#pragma unroll
        for (int x = 0; x < 2; ++x)
#pragma unroll
          for (int y = 0; y < 2; ++y) sa[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], sa[x][y], 0, 0, 0);
      }
    }
    // H = 2 S o K, its row sums and sum H o D, and H into LDS for H X_J.
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      const int jl = wn * 32 + y * 16 + (lane & 15);
      const int64_t j = n0 + jl;
      const double nj = (j < g.n) ? g.nrm[j] : 0.0;
#pragma unroll
      for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int il = wm * 32 + x * 16 + (lane >> 4) + 4 * v;
          const int64_t i = m0 + il;
          double hv = 0.0;
          if (i < g.n && j < g.n) {
            double D = ni[x][v] + nj - 2.0 * gr[x][y][v];
            D = (D > 0.0 && i != j) ? D : 0.0;
            const double kv = (i == j) ? 1.0 : exp(-g.gamma * D);
            hv = 2.0 * sa[x][y][v] * kv;
            dg += hv * D;
          }
          rs[x][v] += hv;
          Hs[jl][il] = hv;
        }
    }
    __syncthreads();
    // HX_I += H (64 x 64) X_J (64 x d)
#pragma unroll
    for (int kk = 0; kk < BN; kk += 4) {
      double a[2];
#pragma unroll
      for (int x = 0; x < 2; ++x) a[x] = Hs[kk + (lane >> 4)][wm * 32 + x * 16 + (lane & 15)];
#pragma unroll
      for (int y = 0; y < ND; ++y) {
        const int f = wn * (DMAX / 2) + y * 16 + (lane & 15);
        const double b = Xj[f][kk + (lane >> 4)];
#pragma unroll
        for (int x = 0; x < 2; ++x) hx[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b, hx[x][y], 0, 0, 0);
      }
    }
  }
  // partial H X of this split
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < ND; ++y)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t i = m0 + wm * 32 + x * 16 + (lane >> 4) + 4 * v;
        const int64_t f = wn * (DMAX / 2) + y * 16 + (lane & 15);
        if (i < g.n && f < g.d) g.part_hx[(int64_t)split * g.n * g.d + i + f * g.n] = hx[x][y][v];
      }
  // row sums: reduce the 16 lanes sharing a row, one partial per column-wave
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      double r = rs[x][v];
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) r += __shfl_xor(r, o);
      const int64_t i = m0 + wm * 32 + x * 16 + (lane >> 4) + 4 * v;
      if ((lane & 15) == 0 && i < g.n) g.part_rs[((int64_t)split * 2 + wn) * g.n + i] = r;
    }
  // sum H o D over the block
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) dg += __shfl_xor(dg, o);
  if (lane == 0) red[wave] = dg;
  __syncthreads();
  if (tid == 0) g.part_dg[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
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
      rs += part_rs[((int64_t)s * 2 + 0) * n + i] + part_rs[((int64_t)s * 2 + 1) * n + i];
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
  const int64_t iblocks = cdiv(n, 64), jblocks = cdiv(n, 64);
  int splits = (int)std::max<int64_t>(1, std::min<int64_t>(jblocks, cdiv(1024, iblocks)));
  const int64_t per = cdiv(jblocks, splits);
  splits = (int)cdiv(jblocks, per);
  DevBuf phx((size_t)splits * n * d * sizeof(double)), prs((size_t)splits * 2 * n * sizeof(double)),
      pdg((size_t)splits * iblocks * sizeof(double));
  fill_f64(st, prs.as<double>(), splits * 2 * n, 0.0);
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
