// SketchySPD (src/Skema_SketchySVD.cpp:757-1151): the Nystrom sketch of a
// symmetric positive semidefinite matrix and its fixed-rank approximation.
#pragma once
#include "sketch.hpp"

namespace skema {

struct Spd {
  Ctx* ctx = nullptr;
  skema_params p{};
  int64_t k = 0;            // range (4r+1 when range < rank, :763-765)
  Map* omega = nullptr;     // logical n x k
  DevBuf om;                // dense logical Omega (n x k, ld n) when not the map's own storage
  const double* omL = nullptr;
  DevBuf omT;               // Omega^T (k x n, ld k): CSR rows and the RBF generator
  DevBuf Y;                 // m x k
  DevBuf win;               // window scratch
  bool updated = false;
  std::map<std::string, std::map<std::string, double>> timings;
};

void spd_prepare(Spd& sp);
void spd_update_dense(Spd& sp, const double* H, int64_t ldh, int64_t row0, int64_t rows);
void spd_update_csr(Spd& sp, const int64_t* rowptr, const int32_t* col, const double* val, int64_t row0,
                    int64_t rows);
void spd_update_rbf(Spd& sp, const double* points, int64_t ldp, int64_t row0, int64_t rows);
// U (n x r, ldu, device) and S (r, device); shift_removal: 0 removes the
// shift eps ||Y||_2 (Tropp et al. 2017, Alg. 3), 1 subtracts ||Y||_2 as the
// reference does (src/Skema_SketchySVD.cpp:1141-1145).
void spd_low_rank_approx(Spd& sp, int shift_removal, double* U, int64_t ldu, double* S);

}  // namespace skema
