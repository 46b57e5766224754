// Small device utilities (fills, strided copies, scaling).
#pragma once
#include "common.hpp"

namespace skema {

inline size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// Fills run as kernels (not hipMemsetAsync) so a caller may capture them.
void fill_f64(hipStream_t st, double* p, int64_t n, double v);
void fill_i32(hipStream_t st, int* p, int64_t n, int v);

// B(rows x cols, ldb) = A(rows x cols, lda)
void copy_matrix(hipStream_t st, int64_t rows, int64_t cols, const double* A, int64_t lda, double* B, int64_t ldb);

// A(rows x cols, lda) *= a   (a == 0 writes zeros without reading)
void scale_matrix(hipStream_t st, int64_t rows, int64_t cols, double a, double* A, int64_t lda);

// B(:, j) = A(:, j) * s[j] for a device vector s.
void scale_columns(hipStream_t st, int64_t rows, int64_t cols, const double* A, int64_t lda, const double* s,
                   double* B, int64_t ldb);

// Sum of squares of a column-major matrix, accumulated in FP64 into *out
// (device scalar, overwritten).  Deterministic two-stage reduction.
void sum_squares(hipStream_t st, int64_t rows, int64_t cols, const double* A, int64_t lda, double* out,
                 DevBuf* scratch);

// Same reduction with caller memory: `part` holds kSumSqPartials doubles
// (overwritten), and the total is ADDED to *acc (device scalar), so a
// sequence of calls accumulates in call order on the device.
constexpr int kSumSqPartials = 1024;
void sum_squares_acc(hipStream_t st, int64_t rows, int64_t cols, const double* A, int64_t lda, double* part,
                     double* acc);

}  // namespace skema
