// Library objects behind the C ABI handles.
#pragma once
#include <chrono>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/skema_mi355x.h"
#include "common.hpp"
#include "gemm.hpp"

namespace skema {

// Event-pair timing of named kernel groups on the context stream.
struct KernelStats {
  struct Rec {
    hipEvent_t a, b;
  };
  std::map<std::string, std::vector<Rec>> open;
  std::map<std::string, std::pair<int64_t, double>> done;  // count, total ms
  std::vector<hipEvent_t> pool;
  hipEvent_t get();
  void collect();  // resolves `open` into `done` (synchronises events)
  void reset();
  ~KernelStats();
};

struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool async_calls = false;
  bool profiling = false;
  std::string profile_only;  // when set, only timing scopes of this name record events
  GemmWorkspace gws;
  DevBuf scratch;  // general grow-only scratch
  DevBuf qr;       // blocked QR workspace (panel reflectors, T, products)
  KernelStats stats;
  void sync_if_needed() {
    if (!async_calls) SKEMA_HIP(hipStreamSynchronize(stream));
  }
};

// RAII timing scope around a group of launches.
struct Timed {
  Ctx* c;
  std::string name;
  hipEvent_t a = nullptr;
  hipStream_t st = nullptr;
  Timed(Ctx* ctx, const char* n) : Timed(ctx, n, ctx->stream) {}
  Timed(Ctx* ctx, const char* n, hipStream_t s) : c(ctx), name(n), st(s) {
    if (c->profiling && (c->profile_only.empty() || c->profile_only == name)) {
      a = c->stats.get();
      SKEMA_HIP(hipEventRecord(a, st));
    }
  }
  // Closes the interval now (else at scope exit).
  void stop() {
    if (a) {
      hipEvent_t b = c->stats.get();
      (void)hipEventRecord(b, st);
      c->stats.open[name].push_back({a, b});
      a = nullptr;
    }
  }
  ~Timed() { stop(); }
};

// CSR on the device.
struct Csr {
  int64_t rows = 0, cols = 0, nnz = 0;
  DevBuf rowptr, col, val;  // int64[rows+1], int32[nnz], double[nnz]
  mutable int64_t max_row = -1;  // longest row, filled on first use
  const int64_t* rp() const { return rowptr.as<int64_t>(); }
  const int32_t* ci() const { return col.as<int32_t>(); }
  const double* v() const { return val.as<double>(); }
};

struct Map {
  Ctx* ctx = nullptr;
  int64_t nrow = 0, ncol = 0;  // stored shape
  bool transposed = false;     // init_transposed
  bool sparse = false;
  DevBuf dense;                // stored, column-major, ld = nrow
  Csr csr;                     // stored
  Csr csr_t;                   // transpose of stored
  double init_seconds = 0.0;
  // Logical map (the k x N operator): stored, or its transpose.
  int64_t lrows() const { return transposed ? ncol : nrow; }
  int64_t lcols() const { return transposed ? nrow : ncol; }
  const Csr& logical_csr() const { return transposed ? csr_t : csr; }
  const Csr& logical_csr_t() const { return transposed ? csr : csr_t; }
  // ELL view of the logical transpose when every logical column holds the
  // same number ell_zeta (<= 16) of entries of one magnitude (sparse-sign
  // maps): ell[c * ell_zeta + f] = row | sign << 31, value +-ell_mag.  One
  // 4*zeta-byte load per column instead of a row-pointer pair and two
  // scattered index/value reads.  ell_zeta == 0 when absent.
  DevBuf ell;
  int ell_zeta = 0;
  double ell_mag = 0.0;
  // The same entries in 16 bits (row | sign << 15) when every row is below
  // 2^15: half the bytes, one 16-byte load per column of 8.  Empty otherwise.
  DevBuf ell16;
  // Sparse maps: the largest |entry| and the most entries of one logical
  // column (bounds for exact fixed-point accumulation).
  double sp_vmax = 0.0;
  int64_t sp_zmax = 0;
};

struct Sketch {
  Ctx* ctx = nullptr;
  skema_params p{};
  int64_t k = 0, s = 0;
  Map* maps[4] = {nullptr, nullptr, nullptr, nullptr};  // Upsilon, Omega, Phi, Psi
  bool sparse_maps = false;
  // Logical dense views used by the dense-map kernels (own copies only when
  // a map is stored transposed), and Omega stacked over Psi.
  DevBuf ups;     // k x m (only if Upsilon is stored transposed)
  DevBuf phi;     // s x m (only if Phi is stored transposed)
  DevBuf omps;    // (k+s) x n, Omega over Psi
  const double* upsL = nullptr;
  int64_t ldups = 0;
  const double* phiL = nullptr;
  int64_t ldphi = 0;
  // Sparse maps: [Omega; Psi] as CSR (rows p) and as CSR of its transpose.
  Csr sp_omps, sp_ompsT;
  // Low-precision map copies (precision forms).
  DevBuf lp_ups, lp_omps, lp_phi;
  // fp8 E8M0 block scales of the window copies (L then R), kept apart from the
  // window scratch so their padding (2^0 past the window's K) is written only
  // when the window row count changes
  DevBuf lp_scales;
  int64_t lp_scales_rows = -1;
  int64_t lp_scales_windows = 1;  // windows whose block scales lp_scales holds (one layout each)
  int64_t lp_ldups = 0, lp_ldomps = 0;
  // State.
  DevBuf X, Y, Z;  // k x n, m x k, s x s
  DevBuf G;        // window x s scratch: H * Psi^T
  DevBuf win;      // window scratch (CSC of a CSR window, RBF norms, ...)
  bool updated = false;
  // Finalize results and checkpoints.
  DevBuf P, Q, PhiQ, PsiP, T1, T2, C, Uc, Sv, Vc, work, tau;
  bool finalized = false;
  std::map<std::string, std::map<std::string, double>> timings;
};

double now_seconds();

// Update kernels (per window); implemented in update.hip.
void update_dense_window(Sketch& sk, const double* H, int64_t ldh, int64_t row0, int64_t rows);
void update_csr_window(Sketch& sk, const int64_t* rowptr, const int32_t* col, const double* val, int64_t row0,
                       int64_t rows);
void update_rbf_window(Sketch& sk, const double* points, int64_t ldp, int64_t row0, int64_t rows);
void prepare_sketch_maps(Sketch& sk);
void finalize_sketch(Sketch& sk);

// Map products with dense operands (maps.hip).
void map_lmap(Map& mp, double alpha, const double* B, int64_t brows, int64_t bcols, int64_t ldb, double beta,
              char transA, char transB, int64_t i0, int64_t i1, double* C, int64_t ldc);
void map_rmap(Map& mp, double alpha, const double* A, int64_t arows, int64_t acols, int64_t lda, double beta,
              char transA, char transB, int64_t i0, int64_t i1, double* C, int64_t ldc);
void build_csr_transpose_host(int64_t rows, int64_t cols, int64_t nnz, const int64_t* rp, const int32_t* ci,
                              const double* v, std::vector<int64_t>& trp, std::vector<int32_t>& tci,
                              std::vector<double>& tv);
void upload_csr(Csr& c, int64_t rows, int64_t cols, int64_t nnz, const int64_t* rp, const int32_t* ci,
                const double* v, hipStream_t st);

}  // namespace skema
