#include "adjoint.hpp"

namespace skema {

void rbf_sketch_loss_grad(Ctx* c, const double* X, int64_t n, int64_t d, int64_t ldp, double gamma, const Map* omega,
                          const Map* phi, const Map* psi, int64_t window, int precision, double* loss, double* dX,
                          int64_t ldg, double* dgamma) {
  (void)c; (void)X; (void)n; (void)d; (void)ldp; (void)gamma; (void)omega; (void)phi; (void)psi; (void)window;
  (void)precision; (void)loss; (void)dX; (void)ldg; (void)dgamma;
  throw Error(kErrUnsupported, "rbf_sketch_loss_grad is not built yet");
}

}  // namespace skema
