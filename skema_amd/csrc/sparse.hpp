// Sparse products of the windowed sketch path.  All are deterministic: every
// output element is owned by one thread that sums its terms in a fixed order.
#pragma once
#include "common.hpp"
#include "gemm.hpp"
#include "sketch.hpp"

namespace skema {

// C (R x N) = alpha * S[r0:r1, c0:c1] * B + beta * C, B is (c1-c0) x N.
// Rows of S must have sorted column indices (maps are sorted on upload).
void spmm_csr_rows(hipStream_t st, const Csr& S, int64_t r0, int64_t r1, int64_t c0, int64_t c1, double alpha,
                   const double* B, int64_t ldb, int64_t N, double beta, double* C, int64_t ldc);

// out (M x R) = B (M x K) * S^T with S an R x K CSR; written through the
// split epilogue of GemmOut (column p of the product = row p of S).  With a
// workspace, short M and long rows of S split each row's nonzeros over
// blocks (FP64 partials, fixed-order sum).
void spmm_dense_csrT(hipStream_t st, int64_t M, const double* B, int64_t ldb, const Csr& S, const GemmOut& out,
                     DevBuf* ws = nullptr);

// Checks a device CSR window (rowptr non-decreasing, column indices in
// [0, cols)) and returns its nonzero count; flag is one device int of
// scratch; ends (optional): rowptr[0] and rowptr[rows] already read on the
// host.  Throws on a malformed window before any kernel indexes with it.
int64_t validate_csr_window(hipStream_t st, const int64_t* rowptr, const int32_t* col, int64_t rows, int64_t cols,
                            int* flag, const int64_t* ends = nullptr);

// Column-compressed copy of a CSR window (rowptr may start at an offset).
struct CscWindow {
  int64_t rows = 0, cols = 0, nnz = 0;
  int64_t* ptr = nullptr;   // cols+1
  int32_t* row = nullptr;   // nnz, window-relative row, ascending per column
  double* val = nullptr;    // nnz
};
// sorted: rows ascending within each column (a deterministic summation
// order for the kernels that walk a column in order); callers that only
// accumulate with atomics pass false.
CscWindow csr_window_to_csc(hipStream_t st, const int64_t* rowptr, const int32_t* col, const double* val,
                            int64_t rows, int64_t cols, DevBuf& scratch, bool sorted = true);

// X[p,j] = nu*X[p,j] + eta * sum_{(i,a) in column j of H} a * M[p, row0+i]
// with M a dense k x * logical map (ldm).
void csc_left_dense(hipStream_t st, const CscWindow& H, const double* M, int64_t ldm, int64_t row0, int64_t k,
                    double eta, double nu, double* X, int64_t ldx);
// Same with a sparse logical map given as the CSR of its transpose (row
// c of Mt = column c of the map).
void csc_left_sparse(hipStream_t st, const CscWindow& H, const Csr& Mt, int64_t row0, int64_t k, double eta,
                     double nu, double* X, int64_t ldx);

// out[i, p] for the CSR window rows i:  sum_{(j,a) in row i} a * M[p, j]
// with M a dense (kk x n) logical map (ldm), through the split epilogue.
void csr_right_dense(hipStream_t st, const int64_t* rowptr, const int32_t* col, const double* val, int64_t rows,
                     const double* M, int64_t ldm, int64_t kk, const GemmOut& out);
// Same with a sparse logical map given as the CSR of its transpose.
void csr_right_sparse(hipStream_t st, const int64_t* rowptr, const int32_t* col, const double* val, int64_t rows,
                      const Csr& Mt, int64_t kk, const GemmOut& out);

}  // namespace skema
