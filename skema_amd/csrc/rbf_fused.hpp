// Fused Gaussian-RBF window + sketch products (K never written to HBM).
#pragma once
#include "gemm.hpp"
#include "sketch.hpp"

namespace skema {

constexpr int kRbfMaxFeat = 128;  // features held in registers for the Gram

struct RbfArgs {
  int64_t M = 0, N = 0, K = 0;  // product shape
  const double* X = nullptr;    // points, n x d column-major
  int64_t ldp = 0, d = 0;
  const double* nrm = nullptr;  // ||x_j||^2, n entries
  double gamma = 0.0;
  int64_t row0 = 0;             // first window row (point index)
  const double* OP = nullptr;   // right: [Omega; Psi] ((k+s) x n), ld ldop
  int64_t ldop = 0;
  const double* U = nullptr;    // left: Upsilon (k x m), ld ldu
  int64_t ldu = 0;
  GemmOut out;
  int tiles_m = 0;
  int64_t k_per_split = 0;
  double* partial = nullptr;
};

// [y | G] = K_w [Omega; Psi]^T through `out` (split at k), K = n points.
void rbf_fused_right_f64(hipStream_t st, RbfArgs a, DevBuf* ws);
// X = nu X + eta Upsilon[:, idx] K_w through `out`.
void rbf_fused_left_f64(hipStream_t st, RbfArgs a);

bool rbf_fused_supported(const Sketch& sk);
void update_rbf_window_fused(Sketch& sk, const double* points, int64_t ldp, int64_t row0, int64_t rows);

}  // namespace skema
