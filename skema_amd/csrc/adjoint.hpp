// Adjoint of the fused RBF window + Gauss sketch (differentiation form).
#pragma once
#include "sketch.hpp"

namespace skema {

// L = ||K Omega^T||_F^2 + ||Phi K Psi^T||_F^2, K = exp(-gamma * D) over the
// n points; *loss, dX (n x d, ldg; may be null) and *dgamma (host scalars).
void rbf_sketch_loss_grad(Ctx* c, const double* X, int64_t n, int64_t d, int64_t ldp, double gamma, const Map* omega,
                          const Map* phi, const Map* psi, int64_t window, int precision, double* loss, double* dX,
                          int64_t ldg, double* dgamma);

}  // namespace skema
