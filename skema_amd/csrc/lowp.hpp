// Reduced-precision forms of the dense sketch contractions (FP32, BF16,
// split-BF16, FP8 e4m3).
#pragma once
#include "sketch.hpp"

namespace skema {

void prepare_lowp_maps(Sketch& sk);
void update_dense_window_lowp(Sketch& sk, const double* H, int64_t ldh, int64_t row0, int64_t rows);

}  // namespace skema
