// BF16 form of the fused Gaussian-RBF window + sketch (BF16 Gram and sketch
// MFMAs, FP32 accumulation and exponent; K never written to HBM).
#pragma once
#include "sketch.hpp"

namespace skema {

bool rbf_bf16_supported(const Sketch& sk);
void update_rbf_window_bf16(Sketch& sk, const double* points, int64_t ldp, int64_t row0, int64_t rows);

}  // namespace skema
