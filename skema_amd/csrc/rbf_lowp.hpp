// BF16 form of the fused Gaussian-RBF window + sketch (BF16 Gram and sketch
// MFMAs, FP32 accumulation and exponent; K never written to HBM).
#pragma once
#include "sketch.hpp"

namespace skema {

bool rbf_bf16_supported(const Sketch& sk);
void update_rbf_window_bf16(Sketch& sk, const double* points, int64_t ldp, int64_t row0, int64_t rows);

// Features of the rounded point set: d padded to 32, 64 or 128.
inline int rbf_bf16_padded_features(int64_t d) { return d <= 32 ? 32 : (d <= 64 ? 64 : 128); }
// Points (n x d, column-major FP64) -> BF16 rows padded to dp features, and
// the FP32 squared norms of the rounded points.
void rbf_points_bf16(hipStream_t st, const double* points, int64_t ldp, int64_t n, int64_t d, int dp, __bf16* P,
                     float* nrm);
// [y | G] = K_w [Omega; Psi]^T for window rows row0 .. row0+rows of the
// npts-point RBF matrix: y (columns < k) through the FP64 epilogue of yout
// (C1, alpha1, beta1), G (s columns) as BF16, column-major with ld ldg.
// OP is [Omega; Psi] in BF16, row-major with ld ldop.
void rbf_right_bf16(Ctx* c, const __bf16* P, int dp, const float* nrm, float gamma, int64_t npts, int64_t row0,
                    int64_t rows, const __bf16* OP, int64_t ldop, int64_t k, int64_t s, const GemmOut& yout,
                    __bf16* G, int64_t ldg);

}  // namespace skema
