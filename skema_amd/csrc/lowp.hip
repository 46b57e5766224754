#include "lowp.hpp"

namespace skema {

void prepare_lowp_maps(Sketch& sk) {
  (void)sk;
  throw Error(kErrUnsupported, "reduced-precision sketch forms are not built yet");
}
void update_dense_window_lowp(Sketch& sk, const double* H, int64_t ldh, int64_t row0, int64_t rows) {
  (void)sk; (void)H; (void)ldh; (void)row0; (void)rows;
  throw Error(kErrUnsupported, "reduced-precision sketch forms are not built yet");
}
bool rbf_fused_supported(const Sketch& sk) { (void)sk; return false; }
void update_rbf_window_fused(Sketch& sk, const double* points, int64_t ldp, int64_t row0, int64_t rows) {
  (void)sk; (void)points; (void)ldp; (void)row0; (void)rows;
  throw Error(kErrUnsupported, "fused RBF window is not built yet");
}

}  // namespace skema
