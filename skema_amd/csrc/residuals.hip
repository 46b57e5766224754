// Residual norms of the final factors (src/Skema_Residuals.hpp:17-217).
//
//   SVD form : rnorms[c] = || [A^T u_c - s_c v_c ; A v_c - s_c u_c] ||_2
//   sym form : rnorms[c] = || A v_c - s_c v_c ||_2   (SketchySPD eigenpairs)
//
// A V and A^T U are formed on the device from the input as it is held: a
// dense column-major A through the FP64 GEMM, a CSR A through the row
// products and a column-compressed copy, and the Gaussian RBF matrix of a
// point set through the fused generator window by window (K is symmetric,
// so A^T U = K U comes out of the same pass as K V, and K is never
// written).  The norms are one block per column with a fixed-order
// reduction, so repeated calls give identical bits.
#include <algorithm>
#include <vector>

#include "gemm.hpp"
#include "linalg.hpp"
#include "rbf.hpp"
#include "rbf_fused.hpp"
#include "residuals.hpp"
#include "sparse.hpp"
#include "util.hpp"

namespace skema {
namespace {

constexpr int kThreads = 256;

// out[c] = sqrt( sum_i (P[i,c] - s_c Q[i,c])^2 + sum_j (R[j,c] - s_c T[j,c])^2 )
// over rows i < m1 of (P, Q) and j < m2 of (R, T); R may be null (m2 = 0).
__global__ __launch_bounds__(kThreads) void k_resid_norms(int64_t m1, const double* P, int64_t ldp, const double* Q,
                                                          int64_t ldq, int64_t m2, const double* R, int64_t ldr,
                                                          const double* T, int64_t ldt, const double* S, double* out) {
  __shared__ double red[kThreads];
  const int64_t c = blockIdx.x;
  const double s = S[c];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < m1; i += kThreads) {
    const double d = P[i + c * ldp] - s * Q[i + c * ldq];
    acc += d * d;
  }
  for (int64_t j = threadIdx.x; j < m2; j += kThreads) {
    const double d = R[j + c * ldr] - s * T[j + c * ldt];
    acc += d * d;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int h = kThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[c] = sqrt(red[0]);
}

void norms_to_host(hipStream_t st, int64_t r, int64_t m1, const double* P, int64_t ldp, const double* Q, int64_t ldq,
                   int64_t m2, const double* R, int64_t ldr, const double* T, int64_t ldt, const double* S,
                   double* rnorms) {
  DevBuf out((size_t)r * sizeof(double));
  hipLaunchKernelGGL(k_resid_norms, dim3((unsigned)r), dim3(kThreads), 0, st, m1, P, ldp, Q, ldq, m2, R, ldr, T, ldt,
                     S, out.as<double>());
  check_launch("k_resid_norms");
  SKEMA_HIP(hipMemcpyAsync(rnorms, out.p, (size_t)r * sizeof(double), hipMemcpyDeviceToHost, st));
  SKEMA_HIP(hipStreamSynchronize(st));
}

// K V (and K U when U is given) over row windows of the RBF matrix of the
// point set: one fused pass with the operator [V^T; U^T].
void rbf_products(Ctx* c, const double* X, int64_t n, int64_t d, int64_t ldp, double gamma, int64_t r,
                  const double* V, int64_t ldv, const double* U, int64_t ldu, int64_t window, double* KV,
                  double* KU) {
  hipStream_t st = c->stream;
  const int64_t h = U ? 2 * r : r;
  DevBuf nrm((size_t)n * sizeof(double)), OP((size_t)h * n * sizeof(double));
  row_norms2(st, X, n, d, ldp, nrm.as<double>());
  transpose(st, n, r, V, ldv, OP.as<double>(), h);
  if (U) transpose(st, n, r, U, ldu, OP.as<double>() + r, h);
  for (int64_t r0 = 0; r0 < n; r0 += window) {
    const int64_t rows = std::min(window, n - r0);
    RbfArgs a;
    a.X = X; a.ldp = ldp; a.d = d; a.nrm = nrm.as<double>(); a.gamma = gamma; a.row0 = r0;
    a.M = rows; a.N = h; a.K = n; a.OP = OP.as<double>(); a.ldop = h;
    a.out.C1 = KV + r0; a.out.ldc1 = n; a.out.alpha1 = 1.0; a.out.beta1 = 0.0; a.out.nsplit = r;
    a.out.C2 = U ? KU + r0 : KV + r0; a.out.ldc2 = n; a.out.alpha2 = 1.0; a.out.beta2 = 0.0;
    rbf_fused_right_f64(st, a, &c->gws.buf);
  }
}

}  // namespace

void residual_fro_stream(hipStream_t st, int64_t m, int64_t n, int64_t r, const double* A, int64_t lda,
                         const double* U, int64_t ldu, const double* S, const double* V, int64_t ldv, int64_t w,
                         double* R, double* US, double* part, double* acc) {
  SKEMA_REQUIRE(w > 0, kErrShape, "residual: window must be positive");
  for (int64_t r0 = 0; r0 < m; r0 += w) {
    const int64_t rows = std::min(w, m - r0);
    copy_matrix(st, rows, n, A + r0, lda, R, rows);
    sum_squares_acc(st, rows, n, R, rows, part, acc + 1);
    if (r > 0) {
      scale_columns(st, rows, r, U + r0, ldu, S, US, rows);
      // no workspace: the product is never split along K (K = r)
      gemm_f64(st, 'N', 'T', rows, n, r, -1.0, US, rows, V, ldv, 1.0, R, rows, nullptr);
    }
    sum_squares_acc(st, rows, n, R, rows, part, acc);
  }
}

void residuals_dense(Ctx* c, int64_t m, int64_t n, int64_t r, const double* A, int64_t lda, const double* U,
                     int64_t ldu, const double* S, const double* V, int64_t ldv, double* rnorms) {
  if (r == 0) return;
  hipStream_t st = c->stream;
  DevBuf AV((size_t)m * r * sizeof(double)), AtU((size_t)n * r * sizeof(double));
  gemm_f64(st, 'N', 'N', m, r, n, 1.0, A, lda, V, ldv, 0.0, AV.as<double>(), m, &c->gws);
  gemm_f64(st, 'T', 'N', n, r, m, 1.0, A, lda, U, ldu, 0.0, AtU.as<double>(), n, &c->gws);
  norms_to_host(st, r, m, AV.as<double>(), m, U, ldu, n, AtU.as<double>(), n, V, ldv, S, rnorms);
}

void residuals_csr(Ctx* c, int64_t m, int64_t n, int64_t r, const int64_t* rowptr, const int32_t* col,
                   const double* val, const double* U, int64_t ldu, const double* S, const double* V, int64_t ldv,
                   double* rnorms) {
  if (r == 0) return;
  hipStream_t st = c->stream;
  DevBuf AV((size_t)m * r * sizeof(double)), AtUt((size_t)r * n * sizeof(double)), Vt((size_t)r * n * sizeof(double)),
      Ut((size_t)r * m * sizeof(double)), AtU((size_t)n * r * sizeof(double)), scratch;
  // A V: row products with the logical r x n operator V^T
  transpose(st, n, r, V, ldv, Vt.as<double>(), r);
  GemmOut o;
  o.C1 = AV.as<double>(); o.ldc1 = m; o.alpha1 = 1.0; o.beta1 = 0.0; o.nsplit = r;
  o.C2 = o.C1; o.ldc2 = m;
  csr_right_dense(st, rowptr, col, val, m, Vt.as<double>(), r, r, o);
  // (A^T U)^T = U^T A: column products over the CSC copy (rows in order)
  transpose(st, m, r, U, ldu, Ut.as<double>(), r);
  const CscWindow H = csr_window_to_csc(st, rowptr, col, val, m, n, scratch, true);
  csc_left_dense(st, H, Ut.as<double>(), r, 0, r, 1.0, 0.0, AtUt.as<double>(), r);
  transpose(st, r, n, AtUt.as<double>(), r, AtU.as<double>(), n);
  norms_to_host(st, r, m, AV.as<double>(), m, U, ldu, n, AtU.as<double>(), n, V, ldv, S, rnorms);
}

void residuals_rbf(Ctx* c, const double* X, int64_t n, int64_t d, int64_t ldp, double gamma, int64_t r,
                   const double* U, int64_t ldu, const double* S, const double* V, int64_t ldv, int64_t window,
                   double* rnorms) {
  if (r == 0) return;
  DevBuf KV((size_t)n * r * sizeof(double)), KU((size_t)n * r * sizeof(double));
  rbf_products(c, X, n, d, ldp, gamma, r, V, ldv, U, ldu, window, KV.as<double>(), KU.as<double>());
  norms_to_host(c->stream, r, n, KV.as<double>(), n, U, ldu, n, KU.as<double>(), n, V, ldv, S, rnorms);
}

void residuals_sym_dense(Ctx* c, int64_t n, int64_t r, const double* A, int64_t lda, const double* V, int64_t ldv,
                         const double* S, double* rnorms) {
  if (r == 0) return;
  hipStream_t st = c->stream;
  DevBuf AV((size_t)n * r * sizeof(double));
  gemm_f64(st, 'N', 'N', n, r, n, 1.0, A, lda, V, ldv, 0.0, AV.as<double>(), n, &c->gws);
  norms_to_host(st, r, n, AV.as<double>(), n, V, ldv, 0, nullptr, 1, nullptr, 1, S, rnorms);
}

void residuals_sym_csr(Ctx* c, int64_t n, int64_t r, const int64_t* rowptr, const int32_t* col, const double* val,
                       const double* V, int64_t ldv, const double* S, double* rnorms) {
  if (r == 0) return;
  hipStream_t st = c->stream;
  DevBuf AV((size_t)n * r * sizeof(double)), Vt((size_t)r * n * sizeof(double));
  transpose(st, n, r, V, ldv, Vt.as<double>(), r);
  GemmOut o;
  o.C1 = AV.as<double>(); o.ldc1 = n; o.alpha1 = 1.0; o.beta1 = 0.0; o.nsplit = r;
  o.C2 = o.C1; o.ldc2 = n;
  csr_right_dense(st, rowptr, col, val, n, Vt.as<double>(), r, r, o);
  norms_to_host(st, r, n, AV.as<double>(), n, V, ldv, 0, nullptr, 1, nullptr, 1, S, rnorms);
}

void residuals_sym_rbf(Ctx* c, const double* X, int64_t n, int64_t d, int64_t ldp, double gamma, int64_t r,
                       const double* V, int64_t ldv, const double* S, int64_t window, double* rnorms) {
  if (r == 0) return;
  DevBuf KV((size_t)n * r * sizeof(double));
  rbf_products(c, X, n, d, ldp, gamma, r, V, ldv, nullptr, 0, window, KV.as<double>(), nullptr);
  norms_to_host(c->stream, r, n, KV.as<double>(), n, V, ldv, 0, nullptr, 1, nullptr, 1, S, rnorms);
}

}  // namespace skema
