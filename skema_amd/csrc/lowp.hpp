// Reduced-precision forms of the dense sketch contractions (FP32, BF16,
// split-BF16, FP8 e4m3).
#pragma once
#include "sketch.hpp"

namespace skema {

void prepare_lowp_maps(Sketch& sk);
void update_dense_window_lowp(Sketch& sk, const double* H, int64_t ldh, int64_t row0, int64_t rows);
// Every window of a dense update (A: m x n, ld lda; the state zeroed) in
// one quantise + left-product pass over all rows, then each window's right
// and core products; false (nothing done) where that pass does not apply
// (forms other than fp8, windows not a multiple of 128 rows, ...).
bool update_dense_all_lowp(Sketch& sk, const double* A, int64_t lda);
// Z <- nu Z + eta Phi[:, row0:row0+rows] G for a BF16 G (rows x s,
// column-major, ld ldg) on the BF16 NT GEMM (bf16 form maps).
void lowp_core_bf16(Sketch& sk, const __bf16* G, int64_t ldg, int64_t row0, int64_t rows);
// C (M x N, column-major, FP64) = alpha * A B^T with A (M x K) and B (N x K)
// BF16 row-major (K contiguous), FP32 accumulation.
void gemm_nt_bf16(hipStream_t st, int64_t M, int64_t N, int64_t K, const __bf16* A, int64_t lda, const __bf16* B,
                  int64_t ldb, double alpha, double* C, int64_t ldc, DevBuf* ws);

}  // namespace skema
