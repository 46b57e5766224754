// Gaussian RBF window generator (src/Skema_Kernel.cpp:5-81).
#pragma once
#include "common.hpp"

namespace skema {

// K (rows x n, ldk) = exp(-gamma * ||x_{row0+i} - x_j||^2) by direct
// differences over the d features of the n x d point set (ldp), with
// K[i, row0+i] = 1 (the reference's diagonal fix, Kernel.cpp:49-74).
void rbf_window(hipStream_t st, const double* X, int64_t n, int64_t d, int64_t ldp, double gamma, int64_t row0,
                int64_t rows, double* K, int64_t ldk);

// Squared row norms of the point set: nrm[i] = sum_f X[i,f]^2.
void row_norms2(hipStream_t st, const double* X, int64_t n, int64_t d, int64_t ldp, double* nrm);

}  // namespace skema
