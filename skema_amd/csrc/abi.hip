// extern "C" boundary of libskema_mi355x (declared in include/skema_mi355x.h).
// Every entry point validates its arguments on the host (shapes, leading
// dimensions, device-pointer extents) before any kernel sees them, converts
// library errors into status codes, and never throws across the ABI.
#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>
#include <sstream>

#include "adjoint.hpp"
#include "linalg.hpp"
#include "lowp.hpp"
#include "rbf.hpp"
#include "rbf_fused.hpp"
#include "csr_pass.hpp"
#include "residuals.hpp"
#include "spd.hpp"
#include "sketch.hpp"
#include "sparse.hpp"
#include "util.hpp"

struct skema_ctx_s {
  skema::Ctx c;
};
struct skema_map_s {
  skema::Map m;
};
struct skema_sketch_s {
  skema::Sketch s;
};
struct skema_spd_s {
  skema::Spd s;
};

namespace skema {
Map* make_sparse_map(Ctx* ctx, int64_t nrow, int64_t ncol, int64_t nnz, const int64_t* rp, const int32_t* ci,
                     const double* v, bool transposed);
Map* make_dense_map(Ctx* ctx, int64_t nrow, int64_t ncol, const double* src, int64_t ld, bool src_device,
                    bool transposed);
void gauss_fill(Ctx* ctx, Map* m, uint64_t seed, int dist);
void sparse_sign_host(int64_t nrow, int64_t ncol, uint64_t seed, int convention, int64_t zeta, bool transposed,
                      std::vector<int64_t>& rp, std::vector<int32_t>& ci, std::vector<double>& v);

hipEvent_t KernelStats::get() {
  if (!pool.empty()) {
    hipEvent_t e = pool.back();
    pool.pop_back();
    return e;
  }
  hipEvent_t e;
  SKEMA_HIP(hipEventCreate(&e));
  return e;
}
void KernelStats::collect() {
  for (auto& kv : open) {
    for (auto& r : kv.second) {
      SKEMA_HIP(hipEventSynchronize(r.b));
      float ms = 0.f;
      SKEMA_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      auto& d = done[kv.first];
      d.first += 1;
      d.second += ms;
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
  }
  open.clear();
}
void KernelStats::reset() {
  collect();
  done.clear();
}
KernelStats::~KernelStats() {
  for (auto& kv : open)
    for (auto& r : kv.second) {
      (void)hipEventDestroy(r.a);
      (void)hipEventDestroy(r.b);
    }
  for (auto e : pool) (void)hipEventDestroy(e);
}
}  // namespace skema

namespace {
using namespace skema;

thread_local std::string g_err;

template <class F>
int guarded(F&& f) {
  try {
    f();
    g_err.clear();
    return SKEMA_OK;
  } catch (const Error& e) {
    g_err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    g_err = "host allocation failed";
    return SKEMA_ERR_OOM;
  } catch (const std::exception& e) {
    g_err = e.what();
    return SKEMA_ERR_ARG;
  } catch (...) {
    g_err = "unknown error";
    return SKEMA_ERR_ARG;
  }
}

void set_device(Ctx* c) { SKEMA_HIP(hipSetDevice(c->device)); }

// The [p, p+bytes) range must lie inside one device allocation.
void check_dev(const void* p, size_t bytes, const char* what) {
  SKEMA_REQUIRE(p != nullptr, kErrArg, std::string(what) + ": null device pointer");
  if (bytes == 0) return;
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  hipError_t e = hipMemGetAddressRange(&base, &size, const_cast<void*>(p));
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(kErrArg, std::string(what) + ": not a device allocation");
  }
  const char* b = static_cast<const char*>(base);
  const char* q = static_cast<const char*>(p);
  SKEMA_REQUIRE(q >= b && q + bytes <= b + size, kErrShape,
                std::string(what) + ": extent exceeds its device allocation");
}
size_t mat_bytes(int64_t rows, int64_t cols, int64_t ld) {
  if (rows <= 0 || cols <= 0) return 0;
  return (size_t)((cols - 1) * ld + rows) * sizeof(double);
}
void check_mat(const double* p, int64_t rows, int64_t cols, int64_t ld, const char* what) {
  SKEMA_REQUIRE(rows >= 0 && cols >= 0, kErrShape, std::string(what) + ": negative dimension");
  SKEMA_REQUIRE(ld >= (rows > 0 ? rows : 1), kErrShape, std::string(what) + ": leading dimension < rows");
  check_dev(p, mat_bytes(rows, cols, ld), what);
}
char norm_trans(char t) {
  if (t == 'n' || t == 'N') return 'N';
  if (t == 't' || t == 'T' || t == 'c' || t == 'C') return 'T';
  throw Error(kErrArg, std::string("transpose flag must be 'N' or 'T', got '") + t + "'");
}

void zero_state(Sketch& sk) {
  hipStream_t st = sk.ctx->stream;
  fill_f64(st, sk.X.as<double>(), sk.k * sk.p.n, 0.0);
  fill_f64(st, sk.Y.as<double>(), sk.p.m * sk.k, 0.0);
  fill_f64(st, sk.Z.as<double>(), sk.s * sk.s, 0.0);
  sk.finalized = false;
}

void check_window(const Sketch& sk, int64_t row0, int64_t rows) {
  SKEMA_REQUIRE(rows > 0 && row0 >= 0 && row0 + rows <= sk.p.m, kErrShape, "update: window outside [0, m)");
}

// Reads rowptr[0] and rowptr[rows] of a device CSR window and checks that
// col/val cover them.
// ends_out (optional): rowptr[0] and rowptr[rows], read here anyway, so the
// caller's own validation needs no second round trip.
void check_csr_window(const int64_t* rowptr, const int32_t* col, const double* val, int64_t rows, hipStream_t st,
                      int64_t* ends_out = nullptr) {
  check_dev(rowptr, (size_t)(rows + 1) * sizeof(int64_t), "rowptr");
  int64_t ends[2];
  SKEMA_HIP(hipMemcpyAsync(&ends[0], rowptr, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  SKEMA_HIP(hipMemcpyAsync(&ends[1], rowptr + rows, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  SKEMA_HIP(hipStreamSynchronize(st));
  SKEMA_REQUIRE(ends[0] >= 0 && ends[1] >= ends[0], kErrArg, "CSR: bad rowptr");
  if (ends[1] > 0) {
    check_dev(col, (size_t)ends[1] * sizeof(int32_t), "col");
    check_dev(val, (size_t)ends[1] * sizeof(double), "val");
  }
  if (ends_out) {
    ends_out[0] = ends[0];
    ends_out[1] = ends[1];
  }
}

template <class Fn>
void for_each_window(const Sketch& sk, Fn&& fn) {
  const int64_t m = sk.p.m;
  int64_t w = sk.p.window;
  if (w == m) {  // single-shot branch (src/Skema_SketchySVD.cpp:137-158)
    fn((int64_t)0, m);
    return;
  }
  for (int64_t irow = 0; irow < m; irow += w) {  // :167-173
    int64_t hi = irow + w;
    if (!(irow + w < m)) {
      hi = m;
      w = hi - irow;
    }
    fn(irow, hi - irow);
  }
}

void low_rank_to_device(Sketch& sk, double* U, int64_t ldu, double* S, double* V, int64_t ldv) {
  Ctx* c = sk.ctx;
  finalize_sketch(sk);
  const int64_t k = sk.k, r = sk.p.rank, m = sk.p.m, n = sk.p.n;
  const double t0 = now_seconds();
  // U = Q*Uc(:,1:r), V = P*Vc(:,1:r)  (:627, :645, :657-660)
  gemm_f64(c->stream, 'N', 'N', m, r, k, 1.0, sk.Q.as<double>(), m, sk.Uc.as<double>(), k, 0.0, U, ldu, &c->gws);
  gemm_f64(c->stream, 'N', 'N', n, r, k, 1.0, sk.P.as<double>(), n, sk.Vc.as<double>(), k, 0.0, V, ldv, &c->gws);
  SKEMA_HIP(hipMemcpyAsync(S, sk.Sv.as<double>(), r * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  SKEMA_HIP(hipStreamSynchronize(c->stream));
  sk.timings["approx"]["dgemm"] += now_seconds() - t0;
}

}  // namespace

extern "C" {

void skema_params_default(skema_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->rank = 1;
  p->range = 1;
  p->core = 1;
  p->window = 1;
  p->eta = 1.0;
  p->nu = 1.0;
  p->dim_redux = SKEMA_MAP_GAUSS;
  p->kernel = SKEMA_KERNEL_NONE;
  p->precision = SKEMA_PREC_FP64;
  p->gamma = 1.0;
}

const char* skema_last_error(void) { return g_err.c_str(); }
int skema_version(void) { return 10000; }

namespace {
void noalloc_lookup(int device);  // static scratch set-up (no-allocation residual pass, below)
}  // namespace

int skema_ctx_create(int device, skema_ctx** out) {
  return guarded([&] {
    SKEMA_REQUIRE(out != nullptr, kErrArg, "ctx_create: null out");
    int count = 0;
    SKEMA_HIP(hipGetDeviceCount(&count));
    SKEMA_REQUIRE(device >= 0 && device < count, kErrArg, "ctx_create: no such device");
    SKEMA_HIP(hipSetDevice(device));
    noalloc_lookup(device);
    std::unique_ptr<skema_ctx_s> c(new skema_ctx_s);
    c->c.device = device;
    *out = c.release();
  });
}

int skema_ctx_destroy(skema_ctx* ctx) {
  return guarded([&] {
    if (!ctx) return;
    set_device(&ctx->c);
    (void)hipStreamSynchronize(ctx->c.stream);
    delete ctx;
  });
}

int skema_ctx_set_stream(skema_ctx* ctx, void* stream) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx, kErrArg, "null ctx");
    ctx->c.stream = static_cast<hipStream_t>(stream);
  });
}

int skema_ctx_set_async(skema_ctx* ctx, int async_calls) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx, kErrArg, "null ctx");
    ctx->c.async_calls = async_calls != 0;
  });
}

int skema_ctx_synchronize(skema_ctx* ctx) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx, kErrArg, "null ctx");
    set_device(&ctx->c);
    SKEMA_HIP(hipStreamSynchronize(ctx->c.stream));
  });
}

int skema_ctx_set_profiling(skema_ctx* ctx, int enable) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx, kErrArg, "null ctx");
    ctx->c.profiling = enable != 0;
  });
}

int skema_ctx_set_profiling_only(skema_ctx* ctx, const char* name) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx, kErrArg, "null ctx");
    ctx->c.profile_only = name ? name : "";
  });
}

int skema_ctx_kernel_stats(skema_ctx* ctx, const char* name, int64_t* count, double* total_ms) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx && name, kErrArg, "kernel_stats: null argument");
    set_device(&ctx->c);
    ctx->c.stats.collect();
    auto it = ctx->c.stats.done.find(name);
    if (count) *count = (it == ctx->c.stats.done.end()) ? 0 : it->second.first;
    if (total_ms) *total_ms = (it == ctx->c.stats.done.end()) ? 0.0 : it->second.second;
  });
}

int skema_ctx_reset_stats(skema_ctx* ctx) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx, kErrArg, "null ctx");
    set_device(&ctx->c);
    ctx->c.stats.reset();
  });
}

// ---- maps -------------------------------------------------------------------
int skema_map_dense(skema_ctx* ctx, int64_t nrow, int64_t ncol, const double* host, int64_t ld, int init_transposed,
                    skema_map** out) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx && host && out, kErrArg, "map_dense: null argument");
    set_device(&ctx->c);
    const double t0 = now_seconds();
    Map* m = make_dense_map(&ctx->c, nrow, ncol, host, ld, false, init_transposed != 0);
    m->init_seconds = now_seconds() - t0;
    *out = reinterpret_cast<skema_map*>(m);
  });
}

int skema_map_dense_device(skema_ctx* ctx, int64_t nrow, int64_t ncol, const double* dev, int64_t ld,
                           int init_transposed, skema_map** out) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx && out, kErrArg, "map_dense_device: null argument");
    set_device(&ctx->c);
    check_mat(dev, nrow, ncol, ld, "map_dense_device");
    const double t0 = now_seconds();
    Map* m = make_dense_map(&ctx->c, nrow, ncol, dev, ld, true, init_transposed != 0);
    m->init_seconds = now_seconds() - t0;
    *out = reinterpret_cast<skema_map*>(m);
  });
}

int skema_map_gauss_seeded(skema_ctx* ctx, int64_t nrow, int64_t ncol, uint64_t seed, int dist, int init_transposed,
                           skema_map** out) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx && out, kErrArg, "map_gauss_seeded: null argument");
    SKEMA_REQUIRE(dist == SKEMA_DIST_NORMAL || dist == SKEMA_DIST_UNIFORM, kErrArg, "map_gauss_seeded: bad dist");
    set_device(&ctx->c);
    const double t0 = now_seconds();
    Map* m = make_dense_map(&ctx->c, nrow, ncol, nullptr, nrow, false, init_transposed != 0);
    gauss_fill(&ctx->c, m, seed, dist);
    m->init_seconds = now_seconds() - t0;
    *out = reinterpret_cast<skema_map*>(m);
  });
}

int skema_map_csr(skema_ctx* ctx, int64_t nrow, int64_t ncol, int64_t nnz, const int64_t* rowptr, const int32_t* col,
                  const double* val, int init_transposed, skema_map** out) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx && rowptr && out, kErrArg, "map_csr: null argument");
    SKEMA_REQUIRE(nnz == 0 || (col && val), kErrArg, "map_csr: null col/val");
    set_device(&ctx->c);
    const double t0 = now_seconds();
    Map* m = make_sparse_map(&ctx->c, nrow, ncol, nnz, rowptr, col, val, init_transposed != 0);
    m->init_seconds = now_seconds() - t0;
    *out = reinterpret_cast<skema_map*>(m);
  });
}

int skema_map_sparse_sign_seeded(skema_ctx* ctx, int64_t nrow, int64_t ncol, uint64_t seed, int convention,
                                 int64_t zeta, int init_transposed, skema_map** out) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx && out, kErrArg, "map_sparse_sign_seeded: null argument");
    SKEMA_REQUIRE(convention == SKEMA_SSIGN_ROWS || convention == SKEMA_SSIGN_COLS, kErrArg,
                  "map_sparse_sign_seeded: bad convention");
    SKEMA_REQUIRE(nrow > 0 && ncol > 0, kErrShape, "map_sparse_sign_seeded: empty shape");
    set_device(&ctx->c);
    const double t0 = now_seconds();
    std::vector<int64_t> rp;
    std::vector<int32_t> ci;
    std::vector<double> v;
    sparse_sign_host(nrow, ncol, seed, convention, zeta, init_transposed != 0, rp, ci, v);
    Map* m = make_sparse_map(&ctx->c, nrow, ncol, (int64_t)ci.size(), rp.data(), ci.data(), v.data(),
                             init_transposed != 0);
    m->init_seconds = now_seconds() - t0;
    *out = reinterpret_cast<skema_map*>(m);
  });
}

int skema_map_destroy(skema_map* map) {
  return guarded([&] {
    if (!map) return;
    Map* m = reinterpret_cast<Map*>(map);
    set_device(m->ctx);
    delete m;
  });
}

int skema_map_info(const skema_map* map, int64_t* nrow, int64_t* ncol, int64_t* nnz, int* issparse,
                   int* init_transposed) {
  return guarded([&] {
    SKEMA_REQUIRE(map, kErrArg, "map_info: null map");
    const Map* m = reinterpret_cast<const Map*>(map);
    if (nrow) *nrow = m->nrow;
    if (ncol) *ncol = m->ncol;
    if (nnz) *nnz = m->sparse ? m->csr.nnz : m->nrow * m->ncol;
    if (issparse) *issparse = m->sparse;
    if (init_transposed) *init_transposed = m->transposed;
  });
}

int skema_map_download(const skema_map* map, double* dense_or_val, int64_t* rowptr, int32_t* col) {
  return guarded([&] {
    SKEMA_REQUIRE(map, kErrArg, "map_download: null map");
    const Map* m = reinterpret_cast<const Map*>(map);
    set_device(m->ctx);
    hipStream_t st = m->ctx->stream;
    if (!m->sparse) {
      if (dense_or_val)
        SKEMA_HIP(hipMemcpyAsync(dense_or_val, m->dense.p, m->nrow * m->ncol * sizeof(double), hipMemcpyDeviceToHost,
                                 st));
    } else {
      if (rowptr)
        SKEMA_HIP(hipMemcpyAsync(rowptr, m->csr.rowptr.p, (m->nrow + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, st));
      if (col && m->csr.nnz)
        SKEMA_HIP(hipMemcpyAsync(col, m->csr.col.p, m->csr.nnz * sizeof(int32_t), hipMemcpyDeviceToHost, st));
      if (dense_or_val && m->csr.nnz)
        SKEMA_HIP(hipMemcpyAsync(dense_or_val, m->csr.val.p, m->csr.nnz * sizeof(double), hipMemcpyDeviceToHost, st));
    }
    SKEMA_HIP(hipStreamSynchronize(st));
  });
}

int skema_map_lmap(skema_map* map, double alpha, const double* B, int64_t brows, int64_t bcols, int64_t ldb,
                   double beta, char transA, char transB, int64_t idx_first, int64_t idx_second, double* C,
                   int64_t ldc) {
  return guarded([&] {
    SKEMA_REQUIRE(map, kErrArg, "lmap: null map");
    Map* m = reinterpret_cast<Map*>(map);
    set_device(m->ctx);
    const char ta = norm_trans(transA), tb = norm_trans(transB);
    check_mat(B, brows, bcols, ldb, "lmap B");
    const bool win = idx_first != idx_second;
    const int64_t wc = win ? idx_second - idx_first : m->ncol;
    const char eff = m->transposed ? (ta == 'N' ? 'T' : 'N') : ta;
    const int64_t M = (eff == 'N') ? m->nrow : wc;
    const int64_t N = (!m->sparse && tb == 'T') ? brows : bcols;
    check_mat(C, M, N, ldc, "lmap C");
    map_lmap(*m, alpha, B, brows, bcols, ldb, beta, ta, tb, idx_first, idx_second, C, ldc);
    m->ctx->sync_if_needed();
  });
}

int skema_map_rmap(skema_map* map, double alpha, const double* A, int64_t arows, int64_t acols, int64_t lda,
                   double beta, char transA, char transB, int64_t idx_first, int64_t idx_second, double* C,
                   int64_t ldc) {
  return guarded([&] {
    SKEMA_REQUIRE(map, kErrArg, "rmap: null map");
    Map* m = reinterpret_cast<Map*>(map);
    set_device(m->ctx);
    const char ta = norm_trans(transA), tb = norm_trans(transB);
    check_mat(A, arows, acols, lda, "rmap A");
    const int64_t M = (!m->sparse && ta == 'T') ? acols : arows;
    const int64_t N = (tb == 'T') ? m->nrow : m->ncol;
    check_mat(C, M, N, ldc, "rmap C");
    map_rmap(*m, alpha, A, arows, acols, lda, beta, ta, tb, idx_first, idx_second, C, ldc);
    m->ctx->sync_if_needed();
  });
}

// ---- windows -----------------------------------------------------------------
int skema_rbf_window(skema_ctx* ctx, const double* points, int64_t n, int64_t d, int64_t ldp, double gamma,
                     int64_t row0, int64_t rows, double* K, int64_t ldk) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx, kErrArg, "rbf_window: null ctx");
    set_device(&ctx->c);
    SKEMA_REQUIRE(n > 0 && d > 0, kErrShape, "rbf_window: empty point set");
    SKEMA_REQUIRE(rows > 0 && row0 >= 0 && row0 + rows <= n, kErrShape, "rbf_window: rows outside [0, n)");
    check_mat(points, n, d, ldp, "rbf_window points");
    check_mat(K, rows, n, ldk, "rbf_window K");
    rbf_window(ctx->c.stream, points, n, d, ldp, gamma, row0, rows, K, ldk);
    ctx->c.sync_if_needed();
  });
}

// ---- sketch ------------------------------------------------------------------
int skema_sketch_create(skema_ctx* ctx, const skema_params* p, skema_map* upsilon, skema_map* omega, skema_map* phi,
                        skema_map* psi, skema_sketch** out) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx && p && out, kErrArg, "sketch_create: null argument");
    set_device(&ctx->c);
    SKEMA_REQUIRE(p->m > 0 && p->n > 0, kErrShape, "sketch_create: m and n must be positive");
    SKEMA_REQUIRE(p->rank >= 1, kErrArg, "sketch_create: rank must be >= 1");
    SKEMA_REQUIRE(p->window >= 1, kErrArg, "sketch_create: window must be >= 1");
    SKEMA_REQUIRE(p->eta >= 0.0 && p->nu >= 0.0, kErrArg, "sketch_create: eta and nu must be >= 0");
    SKEMA_REQUIRE(p->dim_redux == SKEMA_MAP_GAUSS || p->dim_redux == SKEMA_MAP_SPARSE_SIGN, kErrArg,
                  "sketch_create: bad dim_redux");
    SKEMA_REQUIRE(p->kernel == SKEMA_KERNEL_NONE || p->kernel == SKEMA_KERNEL_GAUSSRBF, kErrArg,
                  "sketch_create: bad kernel");
    SKEMA_REQUIRE(p->precision >= SKEMA_PREC_FP64 && p->precision <= SKEMA_PREC_BF16, kErrArg,
                  "sketch_create: bad precision");
    if (p->kernel == SKEMA_KERNEL_GAUSSRBF) {
      SKEMA_REQUIRE(!p->issparse, kErrUnsupported, "Gauss RBF window on a sparse matrix is not available");
      SKEMA_REQUIRE(p->m == p->n, kErrShape, "RBF window: n must equal m (one row and column per point)");
      SKEMA_REQUIRE(p->nfeat >= 1, kErrShape, "RBF window: nfeat must be >= 1");
      SKEMA_REQUIRE(p->gamma >= 0.0, kErrArg, "RBF window: gamma must be >= 0");
    }
    std::unique_ptr<skema_sketch_s> h(new skema_sketch_s);
    Sketch& sk = h->s;
    sk.ctx = &ctx->c;
    sk.p = *p;
    // src/Skema_SketchySVD.cpp:20-24
    sk.k = (p->range < p->rank) ? 4 * p->rank + 1 : p->range;
    sk.s = (p->core < p->rank) ? 2 * sk.k + 1 : p->core;
    sk.maps[0] = reinterpret_cast<Map*>(upsilon);
    sk.maps[1] = reinterpret_cast<Map*>(omega);
    sk.maps[2] = reinterpret_cast<Map*>(phi);
    sk.maps[3] = reinterpret_cast<Map*>(psi);
    SKEMA_REQUIRE(p->rank <= sk.k, kErrArg, "sketch_create: rank must not exceed range");
    prepare_sketch_maps(sk);
    sk.X.alloc((size_t)sk.k * p->n * sizeof(double));
    sk.Y.alloc((size_t)p->m * sk.k * sizeof(double));
    sk.Z.alloc((size_t)sk.s * sk.s * sizeof(double));
    zero_state(sk);
    auto& ti = sk.timings["init"];
    const char* names[4] = {"upsilon", "omega", "phi", "psi"};
    for (int i = 0; i < 4; ++i) ti[names[i]] = sk.maps[i]->init_seconds;
    SKEMA_HIP(hipStreamSynchronize(ctx->c.stream));
    *out = h.release();
  });
}

int skema_sketch_destroy(skema_sketch* sk) {
  return guarded([&] {
    if (!sk) return;
    set_device(sk->s.ctx);
    (void)hipStreamSynchronize(sk->s.ctx->stream);
    delete sk;
  });
}

int skema_sketch_sizes(const skema_sketch* sk, int64_t* range, int64_t* core) {
  return guarded([&] {
    SKEMA_REQUIRE(sk, kErrArg, "null sketch");
    if (range) *range = sk->s.k;
    if (core) *core = sk->s.s;
  });
}

int skema_sketch_reset(skema_sketch* sk) {
  return guarded([&] {
    SKEMA_REQUIRE(sk, kErrArg, "null sketch");
    set_device(sk->s.ctx);
    zero_state(sk->s);
    sk->s.ctx->sync_if_needed();
  });
}

int skema_update_dense(skema_sketch* h, const double* H, int64_t ldh, int64_t row0, int64_t rows) {
  return guarded([&] {
    SKEMA_REQUIRE(h, kErrArg, "null sketch");
    Sketch& sk = h->s;
    set_device(sk.ctx);
    SKEMA_REQUIRE(sk.p.kernel == SKEMA_KERNEL_NONE, kErrArg, "update_dense on an RBF sketch: use update_rbf");
    check_window(sk, row0, rows);
    check_mat(H, rows, sk.p.n, ldh, "update_dense H");
    update_dense_window(sk, H, ldh, row0, rows);
    sk.finalized = false;
    sk.ctx->sync_if_needed();
  });
}

int skema_update_csr(skema_sketch* h, const int64_t* rowptr, const int32_t* col, const double* val, int64_t row0,
                     int64_t rows) {
  return guarded([&] {
    SKEMA_REQUIRE(h, kErrArg, "null sketch");
    Sketch& sk = h->s;
    set_device(sk.ctx);
    SKEMA_REQUIRE(sk.p.kernel == SKEMA_KERNEL_NONE, kErrArg, "update_csr on an RBF sketch");
    check_window(sk, row0, rows);
    check_csr_window(rowptr, col, val, rows, sk.ctx->stream);
    update_csr_window(sk, rowptr, col, val, row0, rows);
    sk.finalized = false;
    sk.ctx->sync_if_needed();
  });
}

int skema_update_rbf(skema_sketch* h, const double* points, int64_t ldp, int64_t row0, int64_t rows) {
  return guarded([&] {
    SKEMA_REQUIRE(h, kErrArg, "null sketch");
    Sketch& sk = h->s;
    set_device(sk.ctx);
    SKEMA_REQUIRE(sk.p.kernel == SKEMA_KERNEL_GAUSSRBF, kErrArg, "update_rbf on a sketch without an RBF window");
    check_window(sk, row0, rows);
    check_mat(points, sk.p.n, sk.p.nfeat, ldp, "update_rbf points");
    update_rbf_window(sk, points, ldp, row0, rows);
    sk.finalized = false;
    sk.ctx->sync_if_needed();
  });
}

int skema_linear_update_dense(skema_sketch* h, const double* A, int64_t lda) {
  return guarded([&] {
    SKEMA_REQUIRE(h, kErrArg, "null sketch");
    Sketch& sk = h->s;
    set_device(sk.ctx);
    SKEMA_REQUIRE(sk.p.kernel == SKEMA_KERNEL_NONE, kErrArg, "linear_update_dense on an RBF sketch");
    check_mat(A, sk.p.m, sk.p.n, lda, "linear_update_dense A");
    // nu = 1 in the fp8, bf16 and split-BF16 forms: the whole matrix as one
    // window (fp8: one quantise + left-product pass, then one right and one
    // core product; bf16 forms: one quantise pass and the three products);
    // they write X, Y and Z whole without reading them, so the state is not
    // cleared first
    if (update_dense_all_lowp(sk, A, lda)) {
      sk.finalized = false;
    } else {
      zero_state(sk);
      for_each_window(sk, [&](int64_t r0, int64_t w) { update_dense_window(sk, A + r0, lda, r0, w); });
    }
    sk.ctx->sync_if_needed();
  });
}

int skema_linear_update_csr(skema_sketch* h, const int64_t* rowptr, const int32_t* col, const double* val) {
  return guarded([&] {
    SKEMA_REQUIRE(h, kErrArg, "null sketch");
    Sketch& sk = h->s;
    set_device(sk.ctx);
    SKEMA_REQUIRE(sk.p.kernel == SKEMA_KERNEL_NONE, kErrArg, "linear_update_csr on an RBF sketch");
    int64_t ends[2];
    check_csr_window(rowptr, col, val, sk.p.m, sk.ctx->stream, ends);
    // the window loop from a zero state in one pass over the whole CSR; the
    // pass writes every entry of X, Y and Z, so the state is not cleared first
    sk.finalized = false;
    update_csr_all(sk, rowptr, col, val, ends);
    sk.ctx->sync_if_needed();
  });
}

int skema_linear_update_rbf(skema_sketch* h, const double* points, int64_t ldp) {
  return guarded([&] {
    SKEMA_REQUIRE(h, kErrArg, "null sketch");
    Sketch& sk = h->s;
    set_device(sk.ctx);
    SKEMA_REQUIRE(sk.p.kernel == SKEMA_KERNEL_GAUSSRBF, kErrArg, "linear_update_rbf without an RBF window");
    check_mat(points, sk.p.n, sk.p.nfeat, ldp, "linear_update_rbf points");
    zero_state(sk);
    for_each_window(sk, [&](int64_t r0, int64_t w) { update_rbf_window(sk, points, ldp, r0, w); });
    sk.ctx->sync_if_needed();
  });
}

int skema_get_state(const skema_sketch* h, double* X, double* Y, double* Z) {
  return guarded([&] {
    SKEMA_REQUIRE(h, kErrArg, "null sketch");
    const Sketch& sk = h->s;
    set_device(sk.ctx);
    hipStream_t st = sk.ctx->stream;
    if (X) SKEMA_HIP(hipMemcpyAsync(X, sk.X.p, sk.k * sk.p.n * sizeof(double), hipMemcpyDeviceToHost, st));
    if (Y) SKEMA_HIP(hipMemcpyAsync(Y, sk.Y.p, sk.p.m * sk.k * sizeof(double), hipMemcpyDeviceToHost, st));
    if (Z) SKEMA_HIP(hipMemcpyAsync(Z, sk.Z.p, sk.s * sk.s * sizeof(double), hipMemcpyDeviceToHost, st));
    SKEMA_HIP(hipStreamSynchronize(st));
  });
}

int skema_set_state(skema_sketch* h, const double* X, const double* Y, const double* Z) {
  return guarded([&] {
    SKEMA_REQUIRE(h, kErrArg, "null sketch");
    Sketch& sk = h->s;
    set_device(sk.ctx);
    hipStream_t st = sk.ctx->stream;
    if (X) SKEMA_HIP(hipMemcpyAsync(sk.X.p, X, sk.k * sk.p.n * sizeof(double), hipMemcpyHostToDevice, st));
    if (Y) SKEMA_HIP(hipMemcpyAsync(sk.Y.p, Y, sk.p.m * sk.k * sizeof(double), hipMemcpyHostToDevice, st));
    if (Z) SKEMA_HIP(hipMemcpyAsync(sk.Z.p, Z, sk.s * sk.s * sizeof(double), hipMemcpyHostToDevice, st));
    SKEMA_HIP(hipStreamSynchronize(st));
    sk.finalized = false;
  });
}

int skema_state_device(skema_sketch* h, double** X, double** Y, double** Z) {
  return guarded([&] {
    SKEMA_REQUIRE(h, kErrArg, "null sketch");
    if (X) *X = h->s.X.as<double>();
    if (Y) *Y = h->s.Y.as<double>();
    if (Z) *Z = h->s.Z.as<double>();
  });
}

int skema_low_rank_approx_device(skema_sketch* h, double* U, int64_t ldu, double* S, double* V, int64_t ldv) {
  return guarded([&] {
    SKEMA_REQUIRE(h, kErrArg, "null sketch");
    Sketch& sk = h->s;
    set_device(sk.ctx);
    check_mat(U, sk.p.m, sk.p.rank, ldu, "low_rank_approx U");
    check_mat(V, sk.p.n, sk.p.rank, ldv, "low_rank_approx V");
    check_dev(S, sk.p.rank * sizeof(double), "low_rank_approx S");
    low_rank_to_device(sk, U, ldu, S, V, ldv);
  });
}

int skema_low_rank_approx(skema_sketch* h, double* U, double* S, double* V) {
  return guarded([&] {
    SKEMA_REQUIRE(h && U && S && V, kErrArg, "low_rank_approx: null argument");
    Sketch& sk = h->s;
    set_device(sk.ctx);
    const int64_t m = sk.p.m, n = sk.p.n, r = sk.p.rank;
    DevBuf dU((size_t)m * r * sizeof(double)), dV((size_t)n * r * sizeof(double)), dS((size_t)r * sizeof(double));
    low_rank_to_device(sk, dU.as<double>(), m, dS.as<double>(), dV.as<double>(), n);
    hipStream_t st = sk.ctx->stream;
    SKEMA_HIP(hipMemcpyAsync(U, dU.p, m * r * sizeof(double), hipMemcpyDeviceToHost, st));
    SKEMA_HIP(hipMemcpyAsync(V, dV.p, n * r * sizeof(double), hipMemcpyDeviceToHost, st));
    SKEMA_HIP(hipMemcpyAsync(S, dS.p, r * sizeof(double), hipMemcpyDeviceToHost, st));
    SKEMA_HIP(hipStreamSynchronize(st));
  });
}

int skema_get_checkpoint(const skema_sketch* h, const char* name, double* host, int64_t* rows, int64_t* cols) {
  return guarded([&] {
    SKEMA_REQUIRE(h && name, kErrArg, "get_checkpoint: null argument");
    const Sketch& sk = h->s;
    SKEMA_REQUIRE(sk.finalized, kErrState, "get_checkpoint: call low_rank_approx first");
    set_device(sk.ctx);
    const int64_t k = sk.k, s = sk.s, m = sk.p.m, n = sk.p.n;
    const std::string nm(name);
    const DevBuf* b = nullptr;
    int64_t r = 0, c = 0;
    if (nm == "P") { b = &sk.P; r = n; c = k; }
    else if (nm == "Q") { b = &sk.Q; r = m; c = k; }
    else if (nm == "PhiQ") { b = &sk.PhiQ; r = s; c = k; }
    else if (nm == "PsiP") { b = &sk.PsiP; r = s; c = k; }
    else if (nm == "T1") { b = &sk.T1; r = k; c = k; }
    else if (nm == "T2") { b = &sk.T2; r = k; c = k; }
    else if (nm == "C") { b = &sk.C; r = k; c = k; }
    else if (nm == "S") { b = &sk.Sv; r = k; c = 1; }
    else throw Error(kErrArg, "get_checkpoint: unknown name " + nm);
    if (rows) *rows = r;
    if (cols) *cols = c;
    if (host) {
      SKEMA_HIP(hipMemcpyAsync(host, b->p, r * c * sizeof(double), hipMemcpyDeviceToHost, sk.ctx->stream));
      SKEMA_HIP(hipStreamSynchronize(sk.ctx->stream));
    }
  });
}

int skema_sketch_timings_json(const skema_sketch* h, char* buf, size_t len) {
  return guarded([&] {
    SKEMA_REQUIRE(h && buf && len > 0, kErrArg, "timings_json: null argument");
    const Sketch& sk = h->s;
    Ctx* c = sk.ctx;
    set_device(c);
    c->stats.collect();
    auto stat_s = [&](const char* k) {
      auto it = c->stats.done.find(k);
      return it == c->stats.done.end() ? 0.0 : it->second.second * 1e-3;
    };
    std::map<std::string, std::map<std::string, double>> t = sk.timings;
    auto& u = t["update"];
    // the fused RBF generator times its left / right products under rbf_*
    u["upsilon"] = stat_s("sketch_left") + stat_s("rbf_left");
    u["omega"] = stat_s("sketch_right") + stat_s("rbf_right");
    u["phi"] = 0.0;  // fused into sketch_right / sketch_core
    u["psi"] = stat_s("sketch_core");
    u["window"] = stat_s("window");
    u["daxpy"] = 0.0;  // folded into the epilogues
    std::ostringstream os;
    os.precision(17);
    os << "{\"timings\":{";
    bool f1 = true;
    for (auto& a : t) {
      os << (f1 ? "" : ",") << "\"" << a.first << "\":{";
      f1 = false;
      bool f2 = true;
      for (auto& b : a.second) {
        os << (f2 ? "" : ",") << "\"" << b.first << "\":" << b.second;
        f2 = false;
      }
      os << "}";
    }
    os << "}}";
    const std::string js = os.str();
    SKEMA_REQUIRE(js.size() + 1 <= len, kErrArg, "timings_json: buffer too small");
    std::memcpy(buf, js.c_str(), js.size() + 1);
  });
}

namespace {
void check_residual_fro_args(int64_t m, int64_t n, int64_t r, const double* A, int64_t lda, const double* U,
                             int64_t ldu, const double* S, const double* V, int64_t ldv, int64_t window,
                             const double* err2, const double* norm2) {
  SKEMA_REQUIRE(err2 && norm2, kErrArg, "residual: null argument");
  SKEMA_REQUIRE(m > 0 && n > 0 && r >= 0 && window > 0, kErrShape, "residual: bad sizes");
  check_mat(A, m, n, lda, "residual A");
  check_mat(U, m, r, ldu, "residual U");
  check_mat(V, n, r, ldv, "residual V");
  if (r) check_dev(S, r * sizeof(double), "residual S");
}
}  // namespace

int skema_residual_fro_dense(skema_ctx* ctx, int64_t m, int64_t n, int64_t r, const double* A, int64_t lda,
                             const double* U, int64_t ldu, const double* S, const double* V, int64_t ldv,
                             int64_t window, double* err2, double* norm2) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx, kErrArg, "residual: null argument");
    Ctx* c = &ctx->c;
    set_device(c);
    check_residual_fro_args(m, n, r, A, lda, U, ldu, S, V, ldv, window, err2, norm2);
    hipStream_t st = c->stream;
    const int64_t w = std::min(window, m);
    DevBuf R((size_t)w * n * sizeof(double)), US((size_t)w * std::max<int64_t>(r, 1) * sizeof(double)),
        acc(2 * sizeof(double)), part(kSumSqPartials * sizeof(double));
    fill_f64(st, acc.as<double>(), 2, 0.0);
    residual_fro_stream(st, m, n, r, A, lda, U, ldu, S, V, ldv, w, R.as<double>(), US.as<double>(),
                        part.as<double>(), acc.as<double>());
    double h[2];
    SKEMA_HIP(hipMemcpyAsync(h, acc.p, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    SKEMA_HIP(hipStreamSynchronize(st));
    *err2 = h[0];
    *norm2 = h[1];
  });
}

// ---- no-allocation residual pass -------------------------------------------
// One static device array per GPU holds the pass's scratch: the two sums
// (acc), the partial sums of the reduction, U diag(S) rows and the residual
// window.  In host code `scratch` names a host-side stand-in, never the
// device copy: its device address is looked up once per device, with that
// device current, when a context is created on it (skema_ctx_create), and
// only that address is handed to kernels and copies.
namespace {
constexpr size_t kNoAllocBytes = size_t(32) << 20;
static __device__ __align__(256) double scratch[kNoAllocBytes / sizeof(double)];

constexpr size_t kNoAllocAccOff = 0;                                   // 2 doubles
constexpr size_t kNoAllocPartOff = 256;                                // kSumSqPartials doubles
constexpr size_t kNoAllocRowsOff = kNoAllocPartOff + kSumSqPartials * sizeof(double);  // R, then US

std::mutex g_noalloc_mu;
std::vector<void*> g_noalloc_ptr;  // per device; nullptr until looked up

// Called with `device` current.
void noalloc_lookup(int device) {
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if ((int)g_noalloc_ptr.size() <= device) g_noalloc_ptr.resize(device + 1, nullptr);
  if (g_noalloc_ptr[device]) return;
  void* ptr = nullptr;
  SKEMA_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  SKEMA_REQUIRE(ptr != nullptr, kErrHip, "static scratch: no device address");
  g_noalloc_ptr[device] = ptr;
}

char* noalloc_ptr(int device) {
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  SKEMA_REQUIRE(device < (int)g_noalloc_ptr.size() && g_noalloc_ptr[device], kErrState,
                "static scratch: not set up on this device (create the context first)");
  return static_cast<char*>(g_noalloc_ptr[device]);
}
}  // namespace

int skema_residual_fro_dense_noalloc(skema_ctx* ctx, int64_t m, int64_t n, int64_t r, const double* A, int64_t lda,
                                     const double* U, int64_t ldu, const double* S, const double* V, int64_t ldv,
                                     int64_t window, double* err2, double* norm2) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx, kErrArg, "residual: null argument");
    Ctx* c = &ctx->c;
    set_device(c);
    check_residual_fro_args(m, n, r, A, lda, U, ldu, S, V, ldv, window, err2, norm2);
    const int64_t rcols = std::max<int64_t>(r, 1);
    const int64_t per_row = (int64_t)sizeof(double) * (n + rcols);
    const int64_t fit = (int64_t)(kNoAllocBytes - kNoAllocRowsOff) / per_row;
    SKEMA_REQUIRE(fit >= 1, kErrShape,
                  "residual (no-alloc): one row needs " + std::to_string(per_row) + " bytes of the " +
                      std::to_string(kNoAllocBytes - kNoAllocRowsOff) + "-byte static scratch (n + max(r,1) <= " +
                      std::to_string((kNoAllocBytes - kNoAllocRowsOff) / sizeof(double)) + ")");
    const int64_t w = std::min({window, m, fit});
    char* base = noalloc_ptr(c->device);
    double* acc = reinterpret_cast<double*>(base + kNoAllocAccOff);
    double* part = reinterpret_cast<double*>(base + kNoAllocPartOff);
    double* R = reinterpret_cast<double*>(base + kNoAllocRowsOff);
    double* US = R + w * n;
    hipStream_t st = c->stream;
    fill_f64(st, acc, 2, 0.0);  // the array keeps the previous call's sums
    residual_fro_stream(st, m, n, r, A, lda, U, ldu, S, V, ldv, w, R, US, part, acc);
    double h[2];
    SKEMA_HIP(hipMemcpyAsync(h, acc, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    SKEMA_HIP(hipStreamSynchronize(st));
    *err2 = h[0];
    *norm2 = h[1];
  });
}

namespace {
void check_factors(int64_t m, int64_t n, int64_t r, const double* U, int64_t ldu, const double* S, const double* V,
                   int64_t ldv, const double* rnorms) {
  SKEMA_REQUIRE(m > 0 && n > 0 && r >= 0 && r <= std::min(m, n), kErrShape, "residuals: bad sizes");
  SKEMA_REQUIRE(rnorms || r == 0, kErrArg, "residuals: null rnorms");
  if (U) check_mat(U, m, r, ldu, "residuals U");
  check_mat(V, n, r, ldv, "residuals V");
  if (r) check_dev(S, r * sizeof(double), "residuals S");
}
}  // namespace

int skema_residuals_dense(skema_ctx* ctx, int64_t m, int64_t n, int64_t r, const double* A, int64_t lda,
                          const double* U, int64_t ldu, const double* S, const double* V, int64_t ldv,
                          double* rnorms) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx && U, kErrArg, "residuals: null argument");
    Ctx* c = &ctx->c;
    set_device(c);
    check_factors(m, n, r, U, ldu, S, V, ldv, rnorms);
    check_mat(A, m, n, lda, "residuals A");
    residuals_dense(c, m, n, r, A, lda, U, ldu, S, V, ldv, rnorms);
  });
}

int skema_residuals_csr(skema_ctx* ctx, int64_t m, int64_t n, int64_t r, const int64_t* rowptr, const int32_t* col,
                        const double* val, const double* U, int64_t ldu, const double* S, const double* V,
                        int64_t ldv, double* rnorms) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx && U, kErrArg, "residuals: null argument");
    Ctx* c = &ctx->c;
    set_device(c);
    check_factors(m, n, r, U, ldu, S, V, ldv, rnorms);
    check_csr_window(rowptr, col, val, m, c->stream);
    residuals_csr(c, m, n, r, rowptr, col, val, U, ldu, S, V, ldv, rnorms);
  });
}

int skema_residuals_rbf(skema_ctx* ctx, const double* points, int64_t n, int64_t d, int64_t ldp, double gamma,
                        int64_t r, const double* U, int64_t ldu, const double* S, const double* V, int64_t ldv,
                        int64_t window, double* rnorms) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx && U, kErrArg, "residuals: null argument");
    SKEMA_REQUIRE(window > 0, kErrShape, "residuals: window must be positive");
    SKEMA_REQUIRE(d >= 1 && d <= kRbfMaxFeat, kErrUnsupported, "residuals: RBF features outside [1, 128]");
    Ctx* c = &ctx->c;
    set_device(c);
    check_factors(n, n, r, U, ldu, S, V, ldv, rnorms);
    check_mat(points, n, d, ldp, "residuals points");
    residuals_rbf(c, points, n, d, ldp, gamma, r, U, ldu, S, V, ldv, window, rnorms);
  });
}

int skema_residuals_sym_dense(skema_ctx* ctx, int64_t n, int64_t r, const double* A, int64_t lda, const double* V,
                              int64_t ldv, const double* S, double* rnorms) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx, kErrArg, "residuals: null context");
    Ctx* c = &ctx->c;
    set_device(c);
    check_factors(n, n, r, nullptr, 0, S, V, ldv, rnorms);
    check_mat(A, n, n, lda, "residuals A");
    residuals_sym_dense(c, n, r, A, lda, V, ldv, S, rnorms);
  });
}

int skema_residuals_sym_csr(skema_ctx* ctx, int64_t n, int64_t r, const int64_t* rowptr, const int32_t* col,
                            const double* val, const double* V, int64_t ldv, const double* S, double* rnorms) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx, kErrArg, "residuals: null context");
    Ctx* c = &ctx->c;
    set_device(c);
    check_factors(n, n, r, nullptr, 0, S, V, ldv, rnorms);
    check_csr_window(rowptr, col, val, n, c->stream);
    residuals_sym_csr(c, n, r, rowptr, col, val, V, ldv, S, rnorms);
  });
}

int skema_residuals_sym_rbf(skema_ctx* ctx, const double* points, int64_t n, int64_t d, int64_t ldp, double gamma,
                            int64_t r, const double* V, int64_t ldv, const double* S, int64_t window,
                            double* rnorms) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx, kErrArg, "residuals: null context");
    SKEMA_REQUIRE(window > 0, kErrShape, "residuals: window must be positive");
    SKEMA_REQUIRE(d >= 1 && d <= kRbfMaxFeat, kErrUnsupported, "residuals: RBF features outside [1, 128]");
    Ctx* c = &ctx->c;
    set_device(c);
    check_factors(n, n, r, nullptr, 0, S, V, ldv, rnorms);
    check_mat(points, n, d, ldp, "residuals points");
    residuals_sym_rbf(c, points, n, d, ldp, gamma, r, V, ldv, S, window, rnorms);
  });
}

int skema_spd_create(skema_ctx* ctx, const skema_params* p, skema_map* omega, skema_spd** out) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx && p && omega && out, kErrArg, "spd_create: null argument");
    set_device(&ctx->c);
    SKEMA_REQUIRE(p->m > 0 && p->m == p->n, kErrShape, "spd_create: A must be square (m == n > 0)");
    SKEMA_REQUIRE(p->rank >= 1 && p->window >= 1, kErrArg, "spd_create: rank and window must be >= 1");
    SKEMA_REQUIRE(p->eta >= 0.0 && p->nu >= 0.0, kErrArg, "spd_create: eta and nu must be >= 0");
    SKEMA_REQUIRE(p->precision == SKEMA_PREC_FP64, kErrUnsupported, "spd: only the FP64 form is built");
    if (p->kernel == SKEMA_KERNEL_GAUSSRBF) {
      SKEMA_REQUIRE(!p->issparse, kErrUnsupported, "Gauss RBF window on a sparse matrix is not available");
      SKEMA_REQUIRE(p->nfeat >= 1 && p->nfeat <= kRbfMaxFeat, kErrUnsupported, "spd: RBF features outside [1, 128]");
      SKEMA_REQUIRE(p->gamma >= 0.0, kErrArg, "RBF window: gamma must be >= 0");
    }
    std::unique_ptr<skema_spd_s> h(new skema_spd_s);
    Spd& sp = h->s;
    sp.ctx = &ctx->c;
    sp.p = *p;
    sp.k = (p->range < p->rank) ? 4 * p->rank + 1 : p->range;
    SKEMA_REQUIRE(p->rank <= sp.k && sp.k <= p->n, kErrShape, "spd_create: need rank <= range <= n");
    sp.omega = reinterpret_cast<Map*>(omega);
    SKEMA_REQUIRE(sp.omega->lrows() == p->n && sp.omega->lcols() == sp.k, kErrShape,
                  "spd_create: Omega must be the logical n x range map");
    spd_prepare(sp);
    sp.Y.alloc((size_t)p->m * sp.k * sizeof(double));
    fill_f64(ctx->c.stream, sp.Y.as<double>(), p->m * sp.k, 0.0);
    sp.timings["init"]["omega"] = sp.omega->init_seconds;
    SKEMA_HIP(hipStreamSynchronize(ctx->c.stream));
    *out = h.release();
  });
}

int skema_spd_destroy(skema_spd* sp) {
  return guarded([&] {
    if (!sp) return;
    set_device(sp->s.ctx);
    (void)hipStreamSynchronize(sp->s.ctx->stream);
    delete sp;
  });
}

extern "C++" {
namespace {
void spd_check_window(const Spd& sp, int64_t row0, int64_t rows) {
  SKEMA_REQUIRE(rows > 0 && row0 >= 0 && row0 + rows <= sp.p.m, kErrShape, "spd update: window outside [0, m)");
}
// The window loop of linear_update; its wall time (synchronous calls) goes
// to update/omega, the key of the reference's fused-here rmap (the window
// generator and the axpy are folded into the same kernels).
template <class Fn>
void spd_windows(Spd& sp, Fn&& fn) {
  const int64_t m = sp.p.m, w = std::min(sp.p.window, m);
  auto& tu = sp.timings["update"];
  for (const char* key : {"omega", "window", "daxpy"}) tu.emplace(key, 0.0);
  const double t0 = now_seconds();
  for (int64_t r0 = 0; r0 < m; r0 += w) fn(r0, std::min(w, m - r0));
  if (!sp.ctx->async_calls) {
    SKEMA_HIP(hipStreamSynchronize(sp.ctx->stream));
    tu["omega"] += now_seconds() - t0;
  }
}
void spd_zero(Spd& sp) { fill_f64(sp.ctx->stream, sp.Y.as<double>(), sp.p.m * sp.k, 0.0); }
}  // namespace
}  // extern "C++"

int skema_spd_update_dense(skema_spd* h, const double* H, int64_t ldh, int64_t row0, int64_t rows) {
  return guarded([&] {
    SKEMA_REQUIRE(h, kErrArg, "null spd");
    Spd& sp = h->s;
    set_device(sp.ctx);
    SKEMA_REQUIRE(sp.p.kernel == SKEMA_KERNEL_NONE, kErrArg, "spd_update_dense on an RBF sketch");
    spd_check_window(sp, row0, rows);
    check_mat(H, rows, sp.p.n, ldh, "spd window");
    spd_update_dense(sp, H, ldh, row0, rows);
    sp.updated = true;
    sp.ctx->sync_if_needed();
  });
}

int skema_spd_update_csr(skema_spd* h, const int64_t* rowptr, const int32_t* col, const double* val, int64_t row0,
                         int64_t rows) {
  return guarded([&] {
    SKEMA_REQUIRE(h, kErrArg, "null spd");
    Spd& sp = h->s;
    set_device(sp.ctx);
    SKEMA_REQUIRE(sp.p.kernel == SKEMA_KERNEL_NONE, kErrArg, "spd_update_csr on an RBF sketch");
    spd_check_window(sp, row0, rows);
    check_csr_window(rowptr, col, val, rows, sp.ctx->stream);
    spd_update_csr(sp, rowptr, col, val, row0, rows);
    sp.updated = true;
    sp.ctx->sync_if_needed();
  });
}

int skema_spd_update_rbf(skema_spd* h, const double* points, int64_t ldp, int64_t row0, int64_t rows) {
  return guarded([&] {
    SKEMA_REQUIRE(h, kErrArg, "null spd");
    Spd& sp = h->s;
    set_device(sp.ctx);
    SKEMA_REQUIRE(sp.p.kernel == SKEMA_KERNEL_GAUSSRBF, kErrArg, "spd_update_rbf without an RBF window");
    spd_check_window(sp, row0, rows);
    check_mat(points, sp.p.n, sp.p.nfeat, ldp, "spd points");
    spd_update_rbf(sp, points, ldp, row0, rows);
    sp.updated = true;
    sp.ctx->sync_if_needed();
  });
}

int skema_spd_linear_update_dense(skema_spd* h, const double* A, int64_t lda) {
  return guarded([&] {
    SKEMA_REQUIRE(h, kErrArg, "null spd");
    Spd& sp = h->s;
    set_device(sp.ctx);
    SKEMA_REQUIRE(sp.p.kernel == SKEMA_KERNEL_NONE, kErrArg, "spd linear_update_dense on an RBF sketch");
    check_mat(A, sp.p.m, sp.p.n, lda, "spd A");
    spd_zero(sp);
    spd_windows(sp, [&](int64_t r0, int64_t w) { spd_update_dense(sp, A + r0, lda, r0, w); });
    sp.updated = true;
    sp.ctx->sync_if_needed();
  });
}

int skema_spd_linear_update_csr(skema_spd* h, const int64_t* rowptr, const int32_t* col, const double* val) {
  return guarded([&] {
    SKEMA_REQUIRE(h, kErrArg, "null spd");
    Spd& sp = h->s;
    set_device(sp.ctx);
    SKEMA_REQUIRE(sp.p.kernel == SKEMA_KERNEL_NONE, kErrArg, "spd linear_update_csr on an RBF sketch");
    check_csr_window(rowptr, col, val, sp.p.m, sp.ctx->stream);
    spd_zero(sp);
    spd_windows(sp, [&](int64_t r0, int64_t w) { spd_update_csr(sp, rowptr + r0, col, val, r0, w); });
    sp.updated = true;
    sp.ctx->sync_if_needed();
  });
}

int skema_spd_linear_update_rbf(skema_spd* h, const double* points, int64_t ldp) {
  return guarded([&] {
    SKEMA_REQUIRE(h, kErrArg, "null spd");
    Spd& sp = h->s;
    set_device(sp.ctx);
    SKEMA_REQUIRE(sp.p.kernel == SKEMA_KERNEL_GAUSSRBF, kErrArg, "spd linear_update_rbf without an RBF window");
    check_mat(points, sp.p.n, sp.p.nfeat, ldp, "spd points");
    spd_zero(sp);
    spd_windows(sp, [&](int64_t r0, int64_t w) { spd_update_rbf(sp, points, ldp, r0, w); });
    sp.updated = true;
    sp.ctx->sync_if_needed();
  });
}

int skema_spd_get_state(const skema_spd* h, double* Y) {
  return guarded([&] {
    SKEMA_REQUIRE(h && Y, kErrArg, "spd_get_state: null argument");
    const Spd& sp = h->s;
    set_device(sp.ctx);
    SKEMA_HIP(hipMemcpyAsync(Y, sp.Y.p, (size_t)sp.p.m * sp.k * sizeof(double), hipMemcpyDeviceToHost,
                             sp.ctx->stream));
    SKEMA_HIP(hipStreamSynchronize(sp.ctx->stream));
  });
}

int skema_spd_low_rank_approx_device(skema_spd* h, int shift_removal, double* U, int64_t ldu, double* S) {
  return guarded([&] {
    SKEMA_REQUIRE(h && U && S, kErrArg, "spd_low_rank_approx: null argument");
    Spd& sp = h->s;
    set_device(sp.ctx);
    SKEMA_REQUIRE(shift_removal == SKEMA_SPD_SHIFT_EPS || shift_removal == SKEMA_SPD_SHIFT_REFERENCE, kErrArg,
                  "spd_low_rank_approx: bad shift_removal");
    SKEMA_REQUIRE(sp.updated, kErrState, "spd_low_rank_approx before any update");
    check_mat(U, sp.p.n, sp.p.rank, ldu, "spd U");
    check_dev(S, (size_t)sp.p.rank * sizeof(double), "spd S");
    spd_low_rank_approx(sp, shift_removal, U, ldu, S);
    SKEMA_HIP(hipStreamSynchronize(sp.ctx->stream));
  });
}

int skema_spd_timings_json(const skema_spd* h, char* buf, size_t len) {
  return guarded([&] {
    SKEMA_REQUIRE(h && buf && len > 0, kErrArg, "spd_timings_json: null argument");
    std::ostringstream os;
    os.precision(17);
    os << "{\"timings\":{";
    bool f1 = true;
    for (auto& a : h->s.timings) {
      os << (f1 ? "" : ",") << "\"" << a.first << "\":{";
      f1 = false;
      bool f2 = true;
      for (auto& b : a.second) {
        os << (f2 ? "" : ",") << "\"" << b.first << "\":" << b.second;
        f2 = false;
      }
      os << "}";
    }
    os << "}}";
    const std::string js = os.str();
    SKEMA_REQUIRE(js.size() + 1 <= len, kErrArg, "spd_timings_json: buffer too small");
    std::memcpy(buf, js.c_str(), js.size() + 1);
  });
}

int skema_spd_low_rank_approx(skema_spd* h, int shift_removal, double* U, double* S) {
  return guarded([&] {
    SKEMA_REQUIRE(h && U && S, kErrArg, "spd_low_rank_approx: null argument");
    Spd& sp = h->s;
    set_device(sp.ctx);
    SKEMA_REQUIRE(shift_removal == SKEMA_SPD_SHIFT_EPS || shift_removal == SKEMA_SPD_SHIFT_REFERENCE, kErrArg,
                  "spd_low_rank_approx: bad shift_removal");
    SKEMA_REQUIRE(sp.updated, kErrState, "spd_low_rank_approx before any update");
    const int64_t n = sp.p.n, r = sp.p.rank;
    DevBuf Ud((size_t)n * r * sizeof(double)), Sd((size_t)r * sizeof(double));
    spd_low_rank_approx(sp, shift_removal, Ud.as<double>(), n, Sd.as<double>());
    SKEMA_HIP(hipMemcpyAsync(U, Ud.p, (size_t)n * r * sizeof(double), hipMemcpyDeviceToHost, sp.ctx->stream));
    SKEMA_HIP(hipMemcpyAsync(S, Sd.p, (size_t)r * sizeof(double), hipMemcpyDeviceToHost, sp.ctx->stream));
    SKEMA_HIP(hipStreamSynchronize(sp.ctx->stream));
  });
}

int skema_rbf_sketch_loss_grad(skema_ctx* ctx, const double* points, int64_t n, int64_t d, int64_t ldp, double gamma,
                               const skema_map* omega, const skema_map* phi, const skema_map* psi, int64_t window,
                               int precision, double* loss, double* dX, int64_t ldg, double* dgamma) {
  return guarded([&] {
    SKEMA_REQUIRE(ctx && omega && phi && psi && loss && dgamma, kErrArg, "rbf_loss_grad: null argument");
    Ctx* c = &ctx->c;
    set_device(c);
    SKEMA_REQUIRE(n > 0 && d > 0 && window > 0, kErrShape, "rbf_loss_grad: bad sizes");
    check_mat(points, n, d, ldp, "rbf_loss_grad points");
    if (dX) check_mat(dX, n, d, ldg, "rbf_loss_grad dX");
    rbf_sketch_loss_grad(c, points, n, d, ldp, gamma, reinterpret_cast<const Map*>(omega),
                         reinterpret_cast<const Map*>(phi), reinterpret_cast<const Map*>(psi), window, precision, loss,
                         dX, ldg, dgamma);
  });
}

}  // extern "C"
