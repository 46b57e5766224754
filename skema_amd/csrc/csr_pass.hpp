// Whole-matrix pass of linear_update over CSR input (closed form of the
// window recurrence from a zero state; see csr_pass.hip).
#pragma once
#include "sketch.hpp"

namespace skema {

// ends: rowptr[0] and rowptr[m] as already read by the caller.
void update_csr_all(Sketch& sk, const int64_t* rowptr, const int32_t* col, const double* val, const int64_t* ends);

}  // namespace skema
