// C++17 host layer over the C ABI in skema_mi355x.h, shaped like the
// reference's own classes so Skema's driver code ports line for line:
//   Skema::GaussDimRedux / SparseSignDimRedux  (src/Skema_DimRedux.hpp:95-308)
//   Skema::SketchySVD<MatrixType, DimReduxT>   (src/Skema_SketchySVD.hpp:11-64)
//   Skema::AlgParams (sketch fields)           (src/Skema_AlgParams.hpp:12-90)
// Matrices are device-resident column-major views (the LayoutLeft analogue);
// the library copies nothing it does not own and throws skema_mi355x::Error
// (status + message) where the C ABI returns a status.
#pragma once
#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "skema_mi355x.h"

namespace skema_mi355x {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

inline void check(int code) {
  if (code != SKEMA_OK) throw Error(code, skema_last_error());
}

// ---- device memory ----------------------------------------------------------
// Owning column-major FP64 matrix in device memory (rows x cols, ld = rows).
class DeviceMatrix {
 public:
  DeviceMatrix() = default;
  DeviceMatrix(int64_t rows, int64_t cols) : rows_(rows), cols_(cols), ld_(rows > 0 ? rows : 1) {
    void* p = nullptr;
    if (rows * cols > 0 && hipMalloc(&p, sizeof(double) * rows * cols) != hipSuccess)
      throw Error(SKEMA_ERR_OOM, "DeviceMatrix: hipMalloc failed");
    data_.reset(static_cast<double*>(p), [](double* q) { (void)hipFree(q); });
    if (p) (void)hipMemset(p, 0, sizeof(double) * rows * cols);
  }
  static DeviceMatrix from_host(const double* colmajor, int64_t rows, int64_t cols, int64_t ld) {
    DeviceMatrix m(rows, cols);
    if (rows * cols > 0 &&
        hipMemcpy2D(m.data(), sizeof(double) * rows, colmajor, sizeof(double) * ld, sizeof(double) * rows, cols,
                    hipMemcpyHostToDevice) != hipSuccess)
      throw Error(SKEMA_ERR_HIP, "DeviceMatrix: upload failed");
    return m;
  }
  std::vector<double> to_host() const {
    std::vector<double> h((size_t)(rows_ * cols_));
    if (!h.empty() && hipMemcpy(h.data(), data(), sizeof(double) * h.size(), hipMemcpyDeviceToHost) != hipSuccess)
      throw Error(SKEMA_ERR_HIP, "DeviceMatrix: download failed");
    return h;
  }
  double* data() const { return data_.get(); }
  int64_t extent(int i) const { return i == 0 ? rows_ : cols_; }
  int64_t ld() const { return ld_; }

 private:
  std::shared_ptr<double> data_;
  int64_t rows_ = 0, cols_ = 0, ld_ = 1;
};

// Device CSR (CrsMatrix<double,int,...,size_t> analogue, src/Skema_Utils.hpp:24-25).
class CrsMatrix {
 public:
  CrsMatrix() = default;
  CrsMatrix(int64_t m, int64_t n, const std::vector<int64_t>& rowptr, const std::vector<int32_t>& col,
            const std::vector<double>& val)
      : m_(m), n_(n), nnz_((int64_t)col.size()) {
    rp_ = upload(rowptr);
    ci_ = upload(col);
    v_ = upload(val);
  }
  int64_t numRows() const { return m_; }
  int64_t numCols() const { return n_; }
  int64_t nnz() const { return nnz_; }
  const int64_t* rowptr() const { return static_cast<const int64_t*>(rp_.get()); }
  const int32_t* col() const { return static_cast<const int32_t*>(ci_.get()); }
  const double* val() const { return static_cast<const double*>(v_.get()); }

 private:
  template <class T>
  static std::shared_ptr<void> upload(const std::vector<T>& h) {
    void* p = nullptr;
    const size_t b = std::max<size_t>(h.size(), 1) * sizeof(T);
    if (hipMalloc(&p, b) != hipSuccess) throw Error(SKEMA_ERR_OOM, "CrsMatrix: hipMalloc failed");
    if (!h.empty() && hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
      throw Error(SKEMA_ERR_HIP, "CrsMatrix: upload failed");
    return std::shared_ptr<void>(p, [](void* q) { (void)hipFree(q); });
  }
  int64_t m_ = 0, n_ = 0, nnz_ = 0;
  std::shared_ptr<void> rp_, ci_, v_;
};

// Dense point set for the Gauss RBF window (n points x d features).
struct PointSet {
  DeviceMatrix X;
};

// ---- context ------------------------------------------------------------------
class Context {
 public:
  explicit Context(int device = 0) {
    skema_ctx* c = nullptr;
    check(skema_ctx_create(device, &c));
    h_.reset(c, [](skema_ctx* q) { (void)skema_ctx_destroy(q); });
  }
  skema_ctx* get() const { return h_.get(); }

 private:
  std::shared_ptr<skema_ctx> h_;
};

// ---- parameters ----------------------------------------------------------------
struct AlgParams {
  int64_t matrix_m = 0, matrix_n = 0;
  int64_t rank = 1, sketch_range = 1, sketch_core = 1, window = 1;
  std::array<uint64_t, 4> seeds{{0, 1, 2, 3}};
  double sketch_eta = 1.0, sketch_nu = 1.0;
  int dim_redux = SKEMA_MAP_GAUSS;
  bool issparse = false;
  int kernel_func = SKEMA_KERNEL_NONE;
  double kernel_gamma = 1.0;
  int64_t nfeat = 0;
  int precision = SKEMA_PREC_FP64;
  // Seeded Gauss maps: U(-a, a) like the reference's fill
  // (src/Skema_DimRedux.hpp:112-115); SKEMA_DIST_NORMAL selects N(0, 1).
  int gauss_dist = SKEMA_DIST_UNIFORM;
  // --sketch-compute-svals-iters: rank-r singular values after every window
  // into the history's traces (src/Skema_SketchySVD.cpp:208-216).
  bool sketch_compute_svals_iters = false;
  // --debug-file: text dumps of the maps, the state after every window and
  // the finalize intermediates, named <stem>_<name>.txt in the working
  // directory (src/Skema_SketchySVD.cpp:192-206, :235-245, :368-490).
  std::string debug_filename;
  bool debug = false;    // --debug: print the window and X, Y, Z (:218-232)
  int print_level = 0;   // --print-level: phase banners (:351-544, :587-638)

  skema_params to_c() const {
    skema_params p;
    skema_params_default(&p);
    p.m = matrix_m;
    p.n = matrix_n;
    p.rank = rank;
    p.range = sketch_range;
    p.core = sketch_core;
    p.window = window;
    p.eta = sketch_eta;
    p.nu = sketch_nu;
    p.dim_redux = dim_redux;
    p.issparse = issparse;
    p.kernel = kernel_func;
    p.gamma = kernel_gamma;
    p.nfeat = nfeat;
    p.precision = precision;
    return p;
  }
  // src/Skema_SketchySVD.cpp:20-24
  int64_t range() const { return sketch_range < rank ? 4 * rank + 1 : sketch_range; }
  int64_t core() const { return sketch_core < rank ? 2 * range() + 1 : sketch_core; }
};

using range_type = std::pair<int64_t, int64_t>;

// ---- DimRedux maps ----------------------------------------------------------------
template <class Derived>
class DimRedux {
 public:
  // lmap: alpha * op(map[:, idx]) * op(B) + beta * C  (src/Skema_DimRedux.hpp:60-69)
  DeviceMatrix lmap(double alpha, const DeviceMatrix& B, double beta, char transA = 'N', char transB = 'N',
                    range_type idx = {0, 0}) const {
    const int64_t wc = idx.first != idx.second ? idx.second - idx.first : ncol_;
    char eff = transA;
    if (init_transposed_) eff = (transA == 'N' || transA == 'n') ? 'T' : 'N';
    const int64_t M = (eff == 'N') ? nrow_ : wc;
    const bool tb = !Derived::sparse && (transB == 'T' || transB == 't');
    DeviceMatrix C(M, tb ? B.extent(0) : B.extent(1));
    check(skema_map_lmap(h_.get(), alpha, B.data(), B.extent(0), B.extent(1), B.ld(), beta, transA, transB,
                         idx.first, idx.second, C.data(), C.ld()));
    return C;
  }
  // rmap: alpha * op(A) * op(map) + beta * C  (src/Skema_DimRedux.hpp:71-80)
  DeviceMatrix rmap(double alpha, const DeviceMatrix& A, double beta, char transA = 'N', char transB = 'T',
                    range_type idx = {0, 0}) const {
    const bool ta = !Derived::sparse && (transA == 'T' || transA == 't');
    DeviceMatrix C(ta ? A.extent(1) : A.extent(0), (transB == 'T' || transB == 't') ? nrow_ : ncol_);
    check(skema_map_rmap(h_.get(), alpha, A.data(), A.extent(0), A.extent(1), A.ld(), beta, transA, transB,
                         idx.first, idx.second, C.data(), C.ld()));
    return C;
  }
  bool issparse() const { return Derived::sparse; }
  int64_t nrows() const { return nrow_; }
  int64_t ncols() const { return ncol_; }
  skema_map* handle() const { return h_.get(); }

 protected:
  void adopt(skema_map* m, int64_t nrow, int64_t ncol, bool t) {
    h_.reset(m, [](skema_map* q) { (void)skema_map_destroy(q); });
    nrow_ = nrow;
    ncol_ = ncol;
    init_transposed_ = t;
  }
  std::shared_ptr<skema_map> h_;
  int64_t nrow_ = 0, ncol_ = 0;
  bool init_transposed_ = false;
};

// GaussDimRedux(nrow, ncol, seed, ..., init_transposed) (src/Skema_DimRedux.hpp:98-128)
class GaussDimRedux : public DimRedux<GaussDimRedux> {
 public:
  static constexpr bool sparse = false;
  GaussDimRedux(const Context& ctx, int64_t nrow, int64_t ncol, uint64_t seed, bool init_transposed = false,
                int dist = SKEMA_DIST_UNIFORM) {
    skema_map* m = nullptr;
    check(skema_map_gauss_seeded(ctx.get(), nrow, ncol, seed, dist, init_transposed, &m));
    adopt(m, nrow, ncol, init_transposed);
  }
  GaussDimRedux(const Context& ctx, int64_t nrow, int64_t ncol, const double* host_colmajor,
                bool init_transposed = false) {
    skema_map* m = nullptr;
    check(skema_map_dense(ctx.get(), nrow, ncol, host_colmajor, nrow, init_transposed, &m));
    adopt(m, nrow, ncol, init_transposed);
  }
};

// SparseSignDimRedux (src/Skema_DimRedux.hpp:191-268)
class SparseSignDimRedux : public DimRedux<SparseSignDimRedux> {
 public:
  static constexpr bool sparse = true;
  SparseSignDimRedux(const Context& ctx, int64_t nrow, int64_t ncol, uint64_t seed, bool init_transposed = false,
                     int convention = SKEMA_SSIGN_ROWS, int64_t zeta = 0) {
    skema_map* m = nullptr;
    check(skema_map_sparse_sign_seeded(ctx.get(), nrow, ncol, seed, convention, zeta, init_transposed, &m));
    adopt(m, nrow, ncol, init_transposed);
  }
};

// ---- text dumps (--debug-file / --debug) ---------------------------------------------
namespace detail {
// <stem of the debug file>_<name> in the working directory, as the reference
// names its dumps (debug_filename.filename().stem() + "_" + name).
inline std::string debug_name(const std::string& debug_file, const std::string& name) {
  std::string base = debug_file.substr(debug_file.find_last_of('/') + 1);
  const size_t dot = base.find_last_of('.');
  if (dot != std::string::npos && dot > 0) base = base.substr(0, dot);
  return base + "_" + name;
}
// Impl::write (src/Skema_Common.hpp:165-186): "%.16f " per entry, a row per line.
inline void write_matrix(const std::string& fname, const double* colmajor, int64_t rows, int64_t cols) {
  FILE* f = std::fopen(fname.c_str(), "w");
  if (!f) throw Error(SKEMA_ERR_ARG, "cannot write " + fname);
  for (int64_t i = 0; i < rows; ++i) {
    for (int64_t j = 0; j < cols; ++j) std::fprintf(f, "%.16f ", colmajor[i + j * rows]);
    std::fprintf(f, "\n");
  }
  std::fclose(f);
}
inline void write_vector(const std::string& fname, const std::vector<double>& v) {
  FILE* f = std::fopen(fname.c_str(), "w");
  if (!f) throw Error(SKEMA_ERR_ARG, "cannot write " + fname);
  for (double x : v) std::fprintf(f, "%.16f\n", x);
  std::fclose(f);
}
// MatrixMarket coordinate text of a stored CSR map (1-based indices).
inline void write_csr_mtx(const std::string& fname, int64_t nrow, int64_t ncol, const std::vector<int64_t>& rp,
                          const std::vector<int32_t>& ci, const std::vector<double>& v) {
  FILE* f = std::fopen(fname.c_str(), "w");
  if (!f) throw Error(SKEMA_ERR_ARG, "cannot write " + fname);
  std::fprintf(f, "%%%%MatrixMarket matrix coordinate real general\n%lld %lld %lld\n", (long long)nrow,
               (long long)ncol, (long long)ci.size());
  for (int64_t r = 0; r < nrow; ++r)
    for (int64_t e = rp[r]; e < rp[r + 1]; ++e)
      std::fprintf(f, "%lld %d %.16g\n", (long long)(r + 1), ci[e] + 1, v[e]);
  std::fclose(f);
}
inline void print_matrix(const char* name, const double* colmajor, int64_t rows, int64_t cols) {
  std::cout << name << " = \n";
  for (int64_t i = 0; i < rows; ++i) {
    for (int64_t j = 0; j < cols; ++j) std::cout << " " << colmajor[i + j * rows];
    std::cout << "\n";
  }
  std::cout.flush();
}
inline void dump_map(skema_map* h, const std::string& debug_file, const std::string& label) {
  int64_t nrow = 0, ncol = 0, nnz = 0;
  int sparse = 0, tr = 0;
  check(skema_map_info(h, &nrow, &ncol, &nnz, &sparse, &tr));
  if (sparse) {
    std::vector<int64_t> rp(nrow + 1);
    std::vector<int32_t> ci(std::max<int64_t>(nnz, 1));
    std::vector<double> v(std::max<int64_t>(nnz, 1));
    check(skema_map_download(h, v.data(), rp.data(), ci.data()));
    ci.resize(nnz);
    v.resize(nnz);
    write_csr_mtx(debug_name(debug_file, label + ".mtx"), nrow, ncol, rp, ci, v);
  } else {
    std::vector<double> d((size_t)(nrow * ncol));
    check(skema_map_download(h, d.data(), nullptr, nullptr));
    write_matrix(debug_name(debug_file, label + ".txt"), d.data(), nrow, ncol);
  }
}
}  // namespace detail

// ---- SketchySVD ------------------------------------------------------------------
// MatrixType: DeviceMatrix (dense A), CrsMatrix (CSR A) or PointSet (RBF).
template <class MatrixType, class DimReduxT>
class SketchySVD {
 public:
  // ctor (src/Skema_SketchySVD.cpp:15-117): maps from params.seeds with the
  // reference's storage orientation.
  SketchySVD(const AlgParams& p, const Context& ctx = Context(0)) : p_(p), ctx_(ctx) {
    const int64_t k = p.range(), s = p.core(), m = p.matrix_m, n = p.matrix_n;
    const bool g = p.dim_redux == SKEMA_MAP_GAUSS;
    const bool tU = p.issparse && g, tO = p.issparse;
    maps_.emplace_back(make(tU ? m : k, tU ? k : m, p.seeds[0], tU));
    maps_.emplace_back(make(tO ? n : k, tO ? k : n, p.seeds[1], tO));
    maps_.emplace_back(make(tU ? m : s, tU ? s : m, p.seeds[2], tU));
    maps_.emplace_back(make(tO ? n : s, tO ? s : n, p.seeds[3], tO));
    skema_params cp = p.to_c();
    skema_sketch* h = nullptr;
    check(skema_sketch_create(ctx_.get(), &cp, maps_[0].handle(), maps_[1].handle(), maps_[2].handle(),
                              maps_[3].handle(), &h));
    h_.reset(h, [](skema_sketch* q) { (void)skema_sketch_destroy(q); });
    if (!p_.debug_filename.empty()) {  // the maps as stored (src/Skema_DimRedux.hpp:124-126, :264-266)
      const char* labels[4] = {"Upsilon", "Omega", "Phi", "Psi"};
      for (int i = 0; i < 4; ++i) detail::dump_map(maps_[i].handle(), p_.debug_filename, labels[i]);
    }
  }

  // linear_update (src/Skema_SketchySVD.cpp:119-246).  The per-window hooks
  // (--sketch-compute-svals-iters, --debug-file, --debug) run the window
  // loop here, one library window update at a time; otherwise the library
  // runs the whole loop.  As in the reference, the single-window case
  // (window == m, :137-158) has no hooks.
  void linear_update(const MatrixType& A) {
    const int64_t m = p_.matrix_m, w = p_.window;
    const bool hooks = p_.sketch_compute_svals_iters || !p_.debug_filename.empty() || p_.debug;
    if (!hooks || w >= m) {
      update_all(A);
      return;
    }
    check(skema_sketch_reset(h_.get()));
    int64_t ucnt = 0;
    for (int64_t r0 = 0; r0 < m; r0 += w, ++ucnt) {
      const int64_t rows = std::min(w, m - r0);
      update_window(A, r0, rows);
      if (!p_.debug_filename.empty()) dump_state("." + std::to_string(ucnt));
      if (p_.sketch_compute_svals_iters) {
        // the finalize leaves X, Y, Z intact here (the reference's Q = Y
        // alias would fold this finalize into the later windows, :381)
        auto usv = low_rank_approx(false);
        traces_[std::to_string(ucnt)]["svals"] = std::get<1>(usv);
      }
      if (p_.debug) {
        print_window(A, r0, rows);
        std::vector<double> X, Y, Z;
        state(X, Y, Z);
        detail::print_matrix("X", X.data(), k(), p_.matrix_n);
        detail::print_matrix("Y", Y.data(), p_.matrix_m, k());
        detail::print_matrix("Z", Z.data(), s(), s());
      }
    }
    if (!p_.debug_filename.empty()) dump_state("");
  }

  // low_rank_approx (src/Skema_SketchySVD.cpp:568-664): U (m x r), S (r), V (n x r)
  std::tuple<DeviceMatrix, std::vector<double>, DeviceMatrix> low_rank_approx() { return low_rank_approx(true); }

  // compute_residuals (src/Skema_SketchySVD.cpp:718-727): the per-triplet
  // residual norms rnrms of the factors (U, S, V) returned by low_rank_approx.
  const std::vector<double>& compute_residuals(const MatrixType& A, const DeviceMatrix& U,
                                               const std::vector<double>& S, const DeviceMatrix& V) {
    DeviceMatrix Sd = DeviceMatrix::from_host(S.data(), (int64_t)S.size(), 1, (int64_t)S.size());
    rnrms_.assign(S.size(), 0.0);
    residuals(A, U, Sd, V);
    return rnrms_;
  }
  const std::vector<double>& rnrms() const { return rnrms_; }

  std::string timings_json() const {
    std::vector<char> buf(16384);
    check(skema_sketch_timings_json(h_.get(), buf.data(), buf.size()));
    return std::string(buf.data());
  }
  // The history's traces (src/Skema_SketchySVD.cpp:50, :214-215) as JSON:
  // {"<window>": {"svals": [...]}, ...}; {} without --sketch-compute-svals-iters.
  std::string traces_json() const {
    std::ostringstream os;
    os.precision(17);
    os << "{";
    bool first = true;
    for (auto& w : traces_) {
      os << (first ? "" : ",") << "\"" << w.first << "\":{";
      first = false;
      bool f2 = true;
      for (auto& kv : w.second) {
        os << (f2 ? "" : ",") << "\"" << kv.first << "\":[";
        f2 = false;
        for (size_t i = 0; i < kv.second.size(); ++i) os << (i ? "," : "") << kv.second[i];
        os << "]";
      }
      os << "}";
    }
    os << "}";
    return os.str();
  }
  // X (k x n), Y (m x k), Z (s x s), column-major, on the host.
  void state(std::vector<double>& X, std::vector<double>& Y, std::vector<double>& Z) const {
    X.resize((size_t)(k() * p_.matrix_n));
    Y.resize((size_t)(p_.matrix_m * k()));
    Z.resize((size_t)(s() * s()));
    check(skema_get_state(h_.get(), X.data(), Y.data(), Z.data()));
  }
  int64_t k() const { return p_.range(); }
  int64_t s() const { return p_.core(); }
  skema_sketch* handle() const { return h_.get(); }

 private:
  DimReduxT make(int64_t r, int64_t c, uint64_t seed, bool t) {
    if constexpr (DimReduxT::sparse) return DimReduxT(ctx_, r, c, seed, t);
    else return DimReduxT(ctx_, r, c, seed, t, p_.gauss_dist);
  }
  std::tuple<DeviceMatrix, std::vector<double>, DeviceMatrix> low_rank_approx(bool final_call) {
    const bool talk = final_call && p_.print_level > 0;
    if (talk) {
      std::cout << "\nComputing fixed-rank approximation\n"
                << "SketchSVD::initial_approx: [P,~] = qr(X^T,0)\n"
                << "SketchSVD::initial_approx: [Q,~] = qr(Y,0);\n"
                << "SketchSVD::initial_approx: Phi*Q\n"
                << "SketchSVD::initial_approx: Psi*P\n"
                << "SketchSVD::initial_approx: [U1,T1] = qr(Phi*Q,0)\n"
                << "SketchSVD::initial_approx: [U2,T2] = qr(Psi*P,0)\n"
                << "SketchSVD::initial_approx: W = T1 \\ (U1^T*Z*U2) / T2^T\n"
                << "SketchSVD::fixed_rank_approx: [uu,ss,vv] = svd(Z)\n"
                << "SketchSVD::fixed_rank_approx: U = Q*U\n"
                << "SketchSVD::fixed_rank_approx: V = P*Vt'" << std::endl;
    }
    DeviceMatrix U(p_.matrix_m, p_.rank), V(p_.matrix_n, p_.rank), S(p_.rank, 1);
    check(skema_low_rank_approx_device(h_.get(), U.data(), U.ld(), S.data(), V.data(), V.ld()));
    if (final_call && !p_.debug_filename.empty()) {  // the finalize intermediates (:368-490)
      for (const char* nm : {"P", "Q", "PhiQ", "PsiP", "T1", "T2"}) {
        int64_t r = 0, c = 0;
        check(skema_get_checkpoint(h_.get(), nm, nullptr, &r, &c));
        std::vector<double> buf((size_t)(r * c));
        check(skema_get_checkpoint(h_.get(), nm, buf.data(), &r, &c));
        detail::write_matrix(detail::debug_name(p_.debug_filename, std::string(nm) + ".txt"), buf.data(), r, c);
      }
    }
    return {U, S.to_host(), V};
  }
  void dump_state(const std::string& suffix) const {
    std::vector<double> X, Y, Z;
    state(X, Y, Z);
    detail::write_matrix(detail::debug_name(p_.debug_filename, "X.txt" + suffix), X.data(), k(), p_.matrix_n);
    detail::write_matrix(detail::debug_name(p_.debug_filename, "Y.txt" + suffix), Y.data(), p_.matrix_m, k());
    detail::write_matrix(detail::debug_name(p_.debug_filename, "Z.txt" + suffix), Z.data(), s(), s());
  }
  void update_window(const DeviceMatrix& A, int64_t r0, int64_t rows) {
    check(skema_update_dense(h_.get(), A.data() + r0, A.ld(), r0, rows));
  }
  void update_window(const CrsMatrix& A, int64_t r0, int64_t rows) {
    check(skema_update_csr(h_.get(), A.rowptr() + r0, A.col(), A.val(), r0, rows));
  }
  void update_window(const PointSet& P, int64_t r0, int64_t rows) {
    check(skema_update_rbf(h_.get(), P.X.data(), P.X.ld(), r0, rows));
  }
  // the window H (src/Skema_SketchySVD.cpp:220-221)
  void print_window(const DeviceMatrix& A, int64_t r0, int64_t rows) const {
    DeviceMatrix H(rows, p_.matrix_n);
    if (hipMemcpy2D(H.data(), sizeof(double) * rows, A.data() + r0, sizeof(double) * A.ld(), sizeof(double) * rows,
                    p_.matrix_n, hipMemcpyDeviceToDevice) != hipSuccess)
      throw Error(SKEMA_ERR_HIP, "print_window: copy failed");
    auto h = H.to_host();
    detail::print_matrix("H", h.data(), rows, p_.matrix_n);
  }
  void print_window(const CrsMatrix& A, int64_t r0, int64_t rows) const {
    std::vector<int64_t> rp(rows + 1);
    if (hipMemcpy(rp.data(), A.rowptr() + r0, sizeof(int64_t) * (rows + 1), hipMemcpyDeviceToHost) != hipSuccess)
      throw Error(SKEMA_ERR_HIP, "print_window: copy failed");
    const int64_t nnz = rp[rows] - rp[0];
    std::vector<int32_t> ci(std::max<int64_t>(nnz, 1));
    std::vector<double> v(std::max<int64_t>(nnz, 1));
    if (nnz > 0 && (hipMemcpy(ci.data(), A.col() + rp[0], sizeof(int32_t) * nnz, hipMemcpyDeviceToHost) !=
                        hipSuccess ||
                    hipMemcpy(v.data(), A.val() + rp[0], sizeof(double) * nnz, hipMemcpyDeviceToHost) != hipSuccess))
      throw Error(SKEMA_ERR_HIP, "print_window: copy failed");
    std::cout << "H = \n";
    for (int64_t r = 0; r < rows; ++r)
      for (int64_t e = rp[r] - rp[0]; e < rp[r + 1] - rp[0]; ++e)
        std::cout << " (" << r << ", " << ci[e] << ") " << v[e] << "\n";
    std::cout.flush();
  }
  void print_window(const PointSet& P, int64_t r0, int64_t rows) const {
    const int64_t n = P.X.extent(0);
    DeviceMatrix K(rows, n);
    check(skema_rbf_window(ctx_.get(), P.X.data(), n, P.X.extent(1), P.X.ld(), p_.kernel_gamma, r0, rows, K.data(),
                           K.ld()));
    auto h = K.to_host();
    detail::print_matrix("H", h.data(), rows, n);
  }
  void update_all(const DeviceMatrix& A) { check(skema_linear_update_dense(h_.get(), A.data(), A.ld())); }
  void update_all(const CrsMatrix& A) {
    check(skema_linear_update_csr(h_.get(), A.rowptr(), A.col(), A.val()));
  }
  void update_all(const PointSet& P) { check(skema_linear_update_rbf(h_.get(), P.X.data(), P.X.ld())); }
  void residuals(const DeviceMatrix& A, const DeviceMatrix& U, const DeviceMatrix& S, const DeviceMatrix& V) {
    check(skema_residuals_dense(ctx_.get(), p_.matrix_m, p_.matrix_n, S.extent(0), A.data(), A.ld(), U.data(),
                                U.ld(), S.data(), V.data(), V.ld(), rnrms_.data()));
  }
  void residuals(const CrsMatrix& A, const DeviceMatrix& U, const DeviceMatrix& S, const DeviceMatrix& V) {
    check(skema_residuals_csr(ctx_.get(), p_.matrix_m, p_.matrix_n, S.extent(0), A.rowptr(), A.col(), A.val(),
                              U.data(), U.ld(), S.data(), V.data(), V.ld(), rnrms_.data()));
  }
  void residuals(const PointSet& P, const DeviceMatrix& U, const DeviceMatrix& S, const DeviceMatrix& V) {
    check(skema_residuals_rbf(ctx_.get(), P.X.data(), P.X.extent(0), P.X.extent(1), P.X.ld(), p_.kernel_gamma,
                              S.extent(0), U.data(), U.ld(), S.data(), V.data(), V.ld(), p_.window,
                              rnrms_.data()));
  }

  AlgParams p_;
  Context ctx_;
  std::vector<DimReduxT> maps_;
  std::shared_ptr<skema_sketch> h_;
  std::vector<double> rnrms_;
  std::map<std::string, std::map<std::string, std::vector<double>>> traces_;
};

// ---- SketchySPD ------------------------------------------------------------------
// MatrixType: DeviceMatrix (dense A), CrsMatrix (CSR A) or PointSet (RBF);
// A symmetric positive semidefinite, m == n.
template <class MatrixType, class DimReduxT>
class SketchySPD {
 public:
  // ctor (src/Skema_SketchySVD.cpp:757-789): Omega = DimReduxT(ncol, range, seeds[0])
  SketchySPD(const AlgParams& p, const Context& ctx = Context(0))
      : p_(p), ctx_(ctx), omega_(make_omega(p, ctx)) {
    skema_params cp = p.to_c();
    skema_spd* h = nullptr;
    check(skema_spd_create(ctx_.get(), &cp, omega_.handle(), &h));
    h_.reset(h, [](skema_spd* q) { (void)skema_spd_destroy(q); });
    if (!p_.debug_filename.empty()) detail::dump_map(omega_.handle(), p_.debug_filename, "Omega");
  }

  // linear_update (src/Skema_SketchySVD.cpp:790-880), every window sketched;
  // --debug-file writes Y at the end (:881-885)
  void linear_update(const MatrixType& A) {
    update_all(A);
    if (!p_.debug_filename.empty()) {
      std::vector<double> Y((size_t)(p_.matrix_m * p_.range()));
      check(skema_spd_get_state(h_.get(), Y.data()));
      detail::write_matrix(detail::debug_name(p_.debug_filename, "Y.txt"), Y.data(), p_.matrix_m, p_.range());
    }
  }

  std::string timings_json() const {
    std::vector<char> buf(8192);
    check(skema_spd_timings_json(h_.get(), buf.data(), buf.size()));
    return std::string(buf.data());
  }

  // low_rank_approx (src/Skema_SketchySVD.cpp:934-1151): U (n x r), eigenvalues (r)
  std::tuple<DeviceMatrix, std::vector<double>> low_rank_approx(int shift_removal = SKEMA_SPD_SHIFT_EPS) {
    DeviceMatrix U(p_.matrix_n, p_.rank), S(p_.rank, 1);
    check(skema_spd_low_rank_approx_device(h_.get(), shift_removal, U.data(), U.ld(), S.data()));
    return {U, S.to_host()};
  }

  // compute_residuals (src/Skema_SketchySVD.cpp:1205-1215 via
  // src/Skema_Residuals.hpp:123-158): ||A u_i - lambda_i u_i||_2
  const std::vector<double>& compute_residuals(const MatrixType& A, const DeviceMatrix& U,
                                               const std::vector<double>& S) {
    DeviceMatrix Sd = DeviceMatrix::from_host(S.data(), (int64_t)S.size(), 1, (int64_t)S.size());
    rnrms_.assign(S.size(), 0.0);
    residuals(A, U, Sd);
    return rnrms_;
  }
  skema_spd* handle() const { return h_.get(); }

 private:
  static DimReduxT make_omega(const AlgParams& p, const Context& ctx) {
    if constexpr (DimReduxT::sparse) return DimReduxT(ctx, p.matrix_n, p.range(), p.seeds[0], false);
    else return DimReduxT(ctx, p.matrix_n, p.range(), p.seeds[0], false, p.gauss_dist);
  }
  void update_all(const DeviceMatrix& A) { check(skema_spd_linear_update_dense(h_.get(), A.data(), A.ld())); }
  void update_all(const CrsMatrix& A) {
    check(skema_spd_linear_update_csr(h_.get(), A.rowptr(), A.col(), A.val()));
  }
  void update_all(const PointSet& P) { check(skema_spd_linear_update_rbf(h_.get(), P.X.data(), P.X.ld())); }
  void residuals(const DeviceMatrix& A, const DeviceMatrix& U, const DeviceMatrix& S) {
    check(skema_residuals_sym_dense(ctx_.get(), p_.matrix_n, S.extent(0), A.data(), A.ld(), U.data(), U.ld(),
                                    S.data(), rnrms_.data()));
  }
  void residuals(const CrsMatrix& A, const DeviceMatrix& U, const DeviceMatrix& S) {
    check(skema_residuals_sym_csr(ctx_.get(), p_.matrix_n, S.extent(0), A.rowptr(), A.col(), A.val(), U.data(),
                                  U.ld(), S.data(), rnrms_.data()));
  }
  void residuals(const PointSet& P, const DeviceMatrix& U, const DeviceMatrix& S) {
    check(skema_residuals_sym_rbf(ctx_.get(), P.X.data(), P.X.extent(0), P.X.extent(1), P.X.ld(), p_.kernel_gamma,
                                  S.extent(0), U.data(), U.ld(), S.data(), p_.window, rnrms_.data()));
  }

  AlgParams p_;
  Context ctx_;
  DimReduxT omega_;
  std::shared_ptr<skema_spd> h_;
  std::vector<double> rnrms_;
};

}  // namespace skema_mi355x
