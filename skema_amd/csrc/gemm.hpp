// Dense FP64 GEMM on CDNA4 matrix cores (v_mfma_f64_16x16x4_f64).
//
// Computes op(A) * op(B) for column-major operands and writes it through a
// two-way column-split epilogue:
//   columns [0, nsplit)  -> C1 = alpha1 * acc + beta1 * C1
//   columns [nsplit, N)  -> C2 = alpha2 * acc + beta2 * C2   (col j-nsplit)
// beta == 0 never reads the output (BLAS semantics).  The split lets one pass
// over a streamed window feed two sketches (Y and the Psi product).
#pragma once
#include "common.hpp"

namespace skema {

struct GemmOut {
  double* C1 = nullptr;
  int64_t ldc1 = 0;
  double alpha1 = 1.0, beta1 = 0.0;
  int64_t nsplit = 0;  // number of columns routed to C1
  double* C2 = nullptr;
  int64_t ldc2 = 0;
  double alpha2 = 1.0, beta2 = 0.0;
};

// Workspace used for split-K partial products (grow-only).
struct GemmWorkspace {
  DevBuf buf;
};

void gemm_f64(hipStream_t stream, char transA, char transB, int64_t M, int64_t N,
              int64_t K, const double* A, int64_t lda, const double* B, int64_t ldb,
              const GemmOut& out, GemmWorkspace* ws);

// C = alpha * op(A) * op(B) + beta * C
void gemm_f64(hipStream_t stream, char transA, char transB, int64_t M, int64_t N,
              int64_t K, double alpha, const double* A, int64_t lda, const double* B,
              int64_t ldb, double beta, double* C, int64_t ldc, GemmWorkspace* ws);

}  // namespace skema
