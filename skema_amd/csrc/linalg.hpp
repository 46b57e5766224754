// FP64 dense linear algebra used by the fixed-rank finalize
// (reference: linalg::qr / ls / svd in src/Skema_BlasLapack.hpp:71-243).
#pragma once
#include <vector>

#include "common.hpp"
#include "gemm.hpp"

namespace skema {

// Householder QR with LAPACK dgeqr2 conventions (dlarfg: beta =
// -sign(alpha)*||x||, tau = (beta-alpha)/beta, v(0) = 1).  On return A (m x n,
// m >= n, leading dimension lda) holds R on and above the diagonal and the
// reflectors below it; tau has n entries.  With a workspace, blocked like
// dgeqrf: panels of 32 columns are factored column by column (one launch
// for a panel of at most 1024 rows, one launch per column with the rows
// spread over the CUs otherwise) and their compact WY form I - V T V^T is
// applied to the trailing columns with FP64 MFMA GEMMs (same reflectors;
// results equal to rounding).  The T factors stay in the workspace.
void householder_qr(hipStream_t st, int64_t m, int64_t n, double* A, int64_t lda, double* tau,
                    DevBuf* work = nullptr, GemmWorkspace* gws = nullptr);

// Explicit thin Q (m x n) from the reflectors left in A by householder_qr
// (dorg2r semantics; with a workspace blocked like dorgqr, each panel's
// I - V T V^T applied to [I; 0] by GEMMs, last panel first).  Q must not
// alias A.  reuse_t: take the T factors householder_qr left in `work` (A was
// the last matrix factored with this workspace) instead of forming them.
void householder_q(hipStream_t st, int64_t m, int64_t n, const double* A, int64_t lda,
                   const double* tau, double* Q, int64_t ldq, DevBuf* work = nullptr,
                   GemmWorkspace* gws = nullptr, bool reuse_t = false);

// R (n x n, upper, zero below) from the factored A.
void extract_r(hipStream_t st, int64_t n, const double* A, int64_t lda, double* R, int64_t ldr);

// B <- T^{-1} B for upper-triangular T (n x n), B (n x nrhs).  This is what
// dgels('N') reduces to when its input is already upper triangular.
void trsm_upper(hipStream_t st, int64_t n, int64_t nrhs, const double* T, int64_t ldt, double* B,
                int64_t ldb);

// B (cols x rows) <- A^T for A (rows x cols).
void transpose(hipStream_t st, int64_t rows, int64_t cols, const double* A, int64_t lda, double* B,
               int64_t ldb);

// One-sided Jacobi SVD of a square n x n matrix A (overwritten).  Produces
// U (n x n), S (n, descending) and V (n x n) with A = U diag(S) V^T.
// Returns the number of sweeps used; throws kErrNumeric when pairs still
// rotate after 100 sweeps.
int jacobi_svd(hipStream_t st, int64_t n, double* A, int64_t lda, double* U, int64_t ldu, double* S,
               double* V, int64_t ldv, DevBuf* scratch);

}  // namespace skema
