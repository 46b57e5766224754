/*
 * libskema_mi355x -- C ABI of the MI355X (gfx950) streaming SketchySVD path.
 *
 * Replaces, on one GPU, the path of jeremy-myers/skema that streams row
 * windows of A through three sketches and finalizes a fixed-rank U, S, V:
 *   SketchySVD<M,D>::SketchySVD / linear_update / low_rank_approx
 *     (src/Skema_SketchySVD.hpp:11-64, src/Skema_SketchySVD.cpp:15-664)
 *   DimRedux<Derived>::lmap / rmap  (src/Skema_DimRedux.hpp:60-80)
 *   WindowBase<M>::get, GaussRBF<M>::compute
 *     (src/Skema_Window.hpp:21-23, src/Skema_Kernel.hpp:18-25)
 *
 * Conventions (same as the reference's Kokkos views, src/Skema_Utils.hpp:9-25):
 *   - scalars are double; dense matrices are column-major with an explicit
 *     leading dimension (ld >= rows);
 *   - CSR uses int64 row pointers, int32 column indices, double values;
 *   - index ranges are half-open [first, second) like range_type.
 * Pointers marked "device" must be device memory of the context's GPU
 * (hipMalloc or any allocator on that device); "host" pointers are ordinary
 * host memory and are copied.  The library never aliases caller memory after
 * a call returns, except maps built with *_device which copy too.
 *
 * Errors: every function returns SKEMA_OK (0) or a status code; a
 * human-readable message for the calling thread is available from
 * skema_last_error().  Nothing exits, aborts or throws across this ABI.
 *
 * Threading: a context and everything created from it must be used by one
 * host thread at a time; distinct contexts may be used concurrently.  Calls
 * are synchronous unless skema_ctx_set_async(ctx, 1) was called, in which case
 * work is only enqueued on the context's stream (skema_ctx_synchronize waits).
 */
#ifndef SKEMA_MI355X_H
#define SKEMA_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define SKEMA_OK 0
#define SKEMA_ERR_ARG 1         /* invalid argument value                   */
#define SKEMA_ERR_SHAPE 2       /* inconsistent dimensions / ld / ranges     */
#define SKEMA_ERR_HIP 3         /* HIP runtime or launch failure             */
#define SKEMA_ERR_OOM 4         /* device allocation failed                  */
#define SKEMA_ERR_NUMERIC 5     /* non-finite or rank-deficient result       */
#define SKEMA_ERR_STATE 6       /* call out of order (e.g. approx before update) */
#define SKEMA_ERR_UNSUPPORTED 7 /* valid but not supported combination      */

/* DimRedux map families (src/Skema_Utils.hpp:78-84). */
#define SKEMA_MAP_GAUSS 0
#define SKEMA_MAP_SPARSE_SIGN 1

/* Window generators (src/Skema_Utils.hpp:64-70). */
#define SKEMA_KERNEL_NONE 0
#define SKEMA_KERNEL_GAUSSRBF 1

/* Arithmetic of the sketch contractions (the finalize is always FP64).
 * Dense input takes every form; the fused RBF generator takes FP64 and BF16
 * (BF16 Gram and sketch, FP32 exponent and accumulation); CSR input and
 * sparse-sign maps are FP64.  The state X, Y, Z is FP64 in every form. */
#define SKEMA_PREC_FP64 0   /* plain: FP64 MFMA, the reference's arithmetic   */
#define SKEMA_PREC_FP32 1   /* FP32 inputs and accumulation                   */
#define SKEMA_PREC_BF16X3 2 /* split-BF16 (hi*hi + hi*lo + lo*hi), FP32 acc.  */
#define SKEMA_PREC_FP8 3    /* FP8 e4m3: maps with one power-of-two scale per
                               row, the window with an E8M0 scale per
                               32-element block of each row and column (one
                               scale per row and column when a window starts
                               off a 16-row boundary), FP32 accumulation;
                               core product in BF16                         */
#define SKEMA_PREC_BF16 4   /* single-term BF16, FP32 acc.                    */

/* Seeded Gauss map distributions. */
#define SKEMA_DIST_NORMAL 0  /* N(0,1)                                         */
#define SKEMA_DIST_UNIFORM 1 /* U(-a,a), a = sqrt(2 ln(nrow*ncol)): the
                                reference's GaussDimRedux fill,
                                src/Skema_DimRedux.hpp:112,115               */

/* Sparse-sign conventions. */
#define SKEMA_SSIGN_ROWS 0 /* zeta = max(2,min(ncol,8)) per stored row, with
                              replacement, +-1 (src/Skema_DimRedux.hpp:205-251) */
#define SKEMA_SSIGN_COLS 1 /* zeta distinct rows per column, +-1/sqrt(zeta)    */

typedef struct skema_ctx_s skema_ctx;
typedef struct skema_map_s skema_map;
typedef struct skema_sketch_s skema_sketch;
typedef struct skema_spd_s skema_spd;

/* Sketch parameters: the sketch fields of AlgParams
 * (src/Skema_AlgParams.hpp:12-90; defaults src/Skema_AlgParams.cpp:10-57). */
typedef struct skema_params {
  int64_t m;         /* rows of A (matrix_m)                                    */
  int64_t n;         /* cols of A (matrix_n); for an RBF window, n = m points   */
  int64_t rank;      /* r                                                       */
  int64_t range;     /* k; range < rank selects 4r+1 (SketchySVD.cpp:20-22)     */
  int64_t core;      /* s; core < rank selects 2k+1 (SketchySVD.cpp:23-24)      */
  int64_t window;    /* rows per window (--window, default 1)                   */
  double eta;        /* S <- nu*S + eta*s (SketchySVD.cpp:187-189, :666-716)    */
  double nu;
  int32_t dim_redux; /* SKEMA_MAP_*                                             */
  int32_t issparse;  /* A is CSR: selects the transposed map storage           */
  int32_t kernel;    /* SKEMA_KERNEL_*                                          */
  int32_t precision; /* SKEMA_PREC_*                                            */
  double gamma;      /* RBF bandwidth (--gamma, default 1.0)                    */
  int64_t nfeat;     /* RBF: features per point (d)                             */
} skema_params;

/* Fills p with the reference defaults (rank 1, range 1, core 1, window 1,
 * eta = nu = 1, gauss, dense, no kernel, FP64, gamma 1). */
void skema_params_default(skema_params* p);

/* ---- errors, version ---------------------------------------------------- */
const char* skema_last_error(void);
int skema_version(void);

/* ---- context (replaces Kokkos::initialize/finalize, driver/main.cpp:150,221) */
int skema_ctx_create(int device, skema_ctx** out);
int skema_ctx_destroy(skema_ctx* ctx);
/* Use a caller stream (a hipStream_t); NULL selects the device's null stream. */
int skema_ctx_set_stream(skema_ctx* ctx, void* hip_stream);
int skema_ctx_set_async(skema_ctx* ctx, int async_calls);
int skema_ctx_synchronize(skema_ctx* ctx);
/* Kernel timing with HIP events recorded around every library kernel launch
 * group; see skema_ctx_kernel_stats. */
int skema_ctx_set_profiling(skema_ctx* ctx, int enable);
/* Restrict profiling to the launch groups named `name` (NULL: all).  Each
 * recorded event pair costs the stream a few microseconds, so timing one
 * group over a run perturbs it far less than timing all of them.  No
 * counterpart in the reference (its timers are host-side,
 * src/Skema_SketchySVD.cpp:126-246). */
int skema_ctx_set_profiling_only(skema_ctx* ctx, const char* name);
/* Average/total milliseconds and launch count recorded under `name` since
 * profiling was enabled (names: "window", "sketch_left", "sketch_right",
 * "sketch_core", "rbf_left", "rbf_right", "rbf_adjoint"). */
int skema_ctx_kernel_stats(skema_ctx* ctx, const char* name, int64_t* count, double* total_ms);
int skema_ctx_reset_stats(skema_ctx* ctx);

/* ---- DimRedux maps (src/Skema_DimRedux.hpp:95-308) ----------------------- */
/* A stored (nrow x ncol) map; init_transposed as in the reference ctor
 * (src/Skema_DimRedux.hpp:38-53): the map it represents is the transpose. */
int skema_map_dense(skema_ctx* ctx, int64_t nrow, int64_t ncol, const double* host, int64_t ld,
                    int init_transposed, skema_map** out);
int skema_map_dense_device(skema_ctx* ctx, int64_t nrow, int64_t ncol, const double* dev, int64_t ld,
                           int init_transposed, skema_map** out);
/* GaussDimRedux ctor (src/Skema_DimRedux.hpp:98-128) on this library's
 * counter-based stream: entry (i,j) uses counter i + j*nrow. */
int skema_map_gauss_seeded(skema_ctx* ctx, int64_t nrow, int64_t ncol, uint64_t seed, int dist,
                           int init_transposed, skema_map** out);
int skema_map_csr(skema_ctx* ctx, int64_t nrow, int64_t ncol, int64_t nnz, const int64_t* rowptr_host,
                  const int32_t* col_host, const double* val_host, int init_transposed, skema_map** out);
/* SparseSignDimRedux ctor (src/Skema_DimRedux.hpp:191-268) on this
 * library's counter-based stream; convention SKEMA_SSIGN_*; zeta <= 0 picks
 * the reference's max(2, min(ncol, 8)) (ROWS) or 8 (COLS). */
int skema_map_sparse_sign_seeded(skema_ctx* ctx, int64_t nrow, int64_t ncol, uint64_t seed, int convention,
                                 int64_t zeta, int init_transposed, skema_map** out);
int skema_map_destroy(skema_map* map);
int skema_map_info(const skema_map* map, int64_t* nrow, int64_t* ncol, int64_t* nnz, int* issparse,
                   int* init_transposed);
/* Copy the stored map to host: dense -> nrow*ncol column-major (ld = nrow);
 * sparse -> rowptr[nrow+1], col[nnz], val[nnz]. Unused pointers may be NULL. */
int skema_map_download(const skema_map* map, double* dense_or_val, int64_t* rowptr, int32_t* col);

/* DimRedux::lmap (src/Skema_DimRedux.hpp:60-69; Gauss: Gauss.cpp:10-33,
 * SparseSign: SparseSign.cpp:9-31) with a dense device operand:
 *   C = alpha * op_A(map[:, idx]) * op_B(B) + beta * C
 * init_transposed flips transA as in the reference; idx_first == idx_second
 * means the whole map.  B is brows x bcols (ldb), C must be sized for the
 * product (ldc >= rows of the product).  The SparseSign form ignores transB
 * like the reference's spmv dispatch (src/Skema_Common.hpp:42-50). */
int skema_map_lmap(skema_map* map, double alpha, const double* B, int64_t brows, int64_t bcols, int64_t ldb,
                   double beta, char transA, char transB, int64_t idx_first, int64_t idx_second, double* C,
                   int64_t ldc);
/* DimRedux::rmap (src/Skema_DimRedux.hpp:71-80; Gauss.cpp:35-52,
 * SparseSign.cpp:33-51) with a dense device operand:
 *   Gauss:      C = alpha * op_A(A) * op_B(map) + beta * C   (idx ignored)
 *   SparseSign: C = alpha * op_A(A) * op_B(map) + beta * C   (the reference
 *               computes it as (op(map) A^T)^T; same value) */
int skema_map_rmap(skema_map* map, double alpha, const double* A, int64_t arows, int64_t acols, int64_t lda,
                   double beta, char transA, char transB, int64_t idx_first, int64_t idx_second, double* C,
                   int64_t ldc);

/* ---- windows (src/Skema_Window.hpp, src/Skema_Kernel.cpp) ---------------- */
/* GaussRBF<matrix_type>::compute for rows [row0,row0+rows) of the point set
 * (src/Skema_Kernel.cpp:5-81): K[i,j] = exp(-gamma*||x_{row0+i}-x_j||^2) for
 * all n points, K[i,row0+i] = 1.  points: device, n x d (ldp); K: device
 * rows x n (ldk).  This is the materialising form; the sketch update uses the
 * fused generator and never writes K. */
int skema_rbf_window(skema_ctx* ctx, const double* points, int64_t n, int64_t d, int64_t ldp, double gamma,
                     int64_t row0, int64_t rows, double* K, int64_t ldk);

/* ---- SketchySVD (src/Skema_SketchySVD.hpp:11-64) ------------------------- */
/* ctor (src/Skema_SketchySVD.cpp:15-117).  The four maps must have the
 * stored shapes the reference builds for (dim_redux, issparse):
 *   Upsilon k x m, Omega k x n, Phi s x m, Psi s x n, stored transposed
 *   (init_transposed = 1) exactly where the reference does (:30-95).
 * The sketch keeps references to the maps: destroy the sketch first. */
int skema_sketch_create(skema_ctx* ctx, const skema_params* p, skema_map* upsilon, skema_map* omega,
                        skema_map* phi, skema_map* psi, skema_sketch** out);
int skema_sketch_destroy(skema_sketch* sk);
/* k (range) and s (core) actually used. */
int skema_sketch_sizes(const skema_sketch* sk, int64_t* range, int64_t* core);
/* X, Y, Z := 0 (linear_update's allocation, src/Skema_SketchySVD.cpp:131-133). */
int skema_sketch_reset(skema_sketch* sk);

/* One window: update(H, idx) + the three axpy's (SketchySVD.cpp:176-189).
 * Dense: H = rows x n device block (ldh), idx = [row0, row0+rows). */
int skema_update_dense(skema_sketch* sk, const double* H, int64_t ldh, int64_t row0, int64_t rows);
/* CSR window with device arrays; rowptr has rows+1 entries and may carry an
 * offset (rowptr[0] != 0 addresses col/val from rowptr[0], like a row slice of
 * the whole matrix: Window<crs>::get, src/Skema_Window.hpp:50-78). */
int skema_update_csr(skema_sketch* sk, const int64_t* rowptr, const int32_t* col, const double* val,
                     int64_t row0, int64_t rows);
/* RBF window rows [row0,row0+rows) generated from the device point set
 * (n x nfeat, ldp) and sketched without materialising K. */
int skema_update_rbf(skema_sketch* sk, const double* points, int64_t ldp, int64_t row0, int64_t rows);

/* Whole linear_update loop (SketchySVD.cpp:119-246): resets X, Y, Z and
 * streams all m rows in params.window sized windows. */
int skema_linear_update_dense(skema_sketch* sk, const double* A, int64_t lda);
int skema_linear_update_csr(skema_sketch* sk, const int64_t* rowptr, const int32_t* col, const double* val);
int skema_linear_update_rbf(skema_sketch* sk, const double* points, int64_t ldp);

/* Sketch state (host, column-major: X k x n, Y m x k, Z s x s); NULL skips. */
int skema_get_state(const skema_sketch* sk, double* X, double* Y, double* Z);
int skema_set_state(skema_sketch* sk, const double* X, const double* Y, const double* Z);
/* Device pointers to the live state (valid until destroy). */
int skema_state_device(skema_sketch* sk, double** X, double** Y, double** Z);

/* low_rank_approx (SketchySVD.cpp:568-664 with initial_approx :337-566), all
 * FP64 on the GPU.  Host outputs: U m x r, S r, V n x r (column-major). Y is
 * left intact (the reference overwrites it, :381). */
int skema_low_rank_approx(skema_sketch* sk, double* U, double* S, double* V);
/* Same, writing device outputs (U ldu >= m, V ldv >= n). */
int skema_low_rank_approx_device(skema_sketch* sk, double* U, int64_t ldu, double* S, double* V, int64_t ldv);
/* Intermediate of the last finalize, copied to host: "P" (n x k), "Q" (m x k),
 * "PhiQ" and "PsiP" (s x k, before their QR), "T1", "T2", "C" (k x k). */
int skema_get_checkpoint(const skema_sketch* sk, const char* name, double* host, int64_t* rows, int64_t* cols);

/* Timings with the reference's keys ({"init":{...},"update":{...},
 * "approx":{...}}, src/Skema_SketchySVD.cpp:101-116) as JSON text. */
int skema_sketch_timings_json(const skema_sketch* sk, char* buf, size_t len);

/* ---- SketchySPD (Nystrom sketch of a symmetric PSD matrix) ---------------- */
/* SketchySPD<M,D> (src/Skema_SketchySVD.cpp:757-1151): Y = A Omega (m x k)
 * over row windows, then the fixed-rank Nystrom approximation.  p->m == p->n;
 * k = range (4r+1 when range < rank, :763-765); omega is the logical n x k
 * map (the reference's DimReduxT(ncol, range, seeds[0]), :768-773).  The
 * reference's window loop never resets its pass counter (:795-858) and so
 * sketches window 0 only; every window is sketched here. */
int skema_spd_create(skema_ctx* ctx, const skema_params* p, skema_map* omega, skema_spd** out);
int skema_spd_destroy(skema_spd* sp);
/* One window of rows [row0, row0+rows): Y[idx] = nu Y[idx] + eta A_w Omega
 * (update + axpy, :896-931, :1153-1200).  Dense A_w (ldh), a CSR window
 * (rowptr at row row0 of the whole matrix), or the RBF rows of the point set. */
int skema_spd_update_dense(skema_spd* sp, const double* H, int64_t ldh, int64_t row0, int64_t rows);
int skema_spd_update_csr(skema_spd* sp, const int64_t* rowptr, const int32_t* col, const double* val,
                         int64_t row0, int64_t rows);
int skema_spd_update_rbf(skema_spd* sp, const double* points, int64_t ldp, int64_t row0, int64_t rows);
/* linear_update (:790-880): zero Y, then every window. */
int skema_spd_linear_update_dense(skema_spd* sp, const double* A, int64_t lda);
int skema_spd_linear_update_csr(skema_spd* sp, const int64_t* rowptr, const int32_t* col, const double* val);
int skema_spd_linear_update_rbf(skema_spd* sp, const double* points, int64_t ldp);
/* Y (m x k, column-major) to the host. */
int skema_spd_get_state(const skema_spd* sp, double* Y);
/* low_rank_approx (:934-1151): U (n x r) and the eigenvalues lambda (r).
 * shift_removal SKEMA_SPD_SHIFT_EPS subtracts the shift eps ||Y||_2 (Tropp et
 * al. 2017, Alg. 3); SKEMA_SPD_SHIFT_REFERENCE subtracts ||Y||_2 as the
 * reference does (:1141-1145), which clips every eigenvalue below ||Y||_2. */
/* Timings with the reference's SketchySPD keys ({"init":{"omega"},
 * "update":{"omega","window","daxpy"}, "approx":{"norm2","daxpy","omega",
 * "update","dpotrf","dgels","dgesvd"}}, src/Skema_SketchySVD.cpp:776-786) as
 * JSON text; the fused window update's wall time is under update/omega. */
int skema_spd_timings_json(const skema_spd* sp, char* buf, size_t len);
#define SKEMA_SPD_SHIFT_EPS 0
#define SKEMA_SPD_SHIFT_REFERENCE 1
int skema_spd_low_rank_approx(skema_spd* sp, int shift_removal, double* U, double* S);
int skema_spd_low_rank_approx_device(skema_spd* sp, int shift_removal, double* U, int64_t ldu, double* S);

/* ---- streamed quality pass ----------------------------------------------- */
/* ||A - U diag(S) V^T||_F^2 and ||A||_F^2 for dense device A (m x n, lda),
 * U (m x r), V (n x r), streamed in row windows of `window` rows. */
int skema_residual_fro_dense(skema_ctx* ctx, int64_t m, int64_t n, int64_t r, const double* A, int64_t lda,
                             const double* U, int64_t ldu, const double* S, const double* V, int64_t ldv,
                             int64_t window, double* err2, double* norm2);

/* The same pass without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * scratch (the two sums, the partial sums, U diag(S) rows and the residual
 * window) is one static 32 MiB device array per GPU, held until the process
 * ends.  Windows are cut to the rows that fit, at most
 * floor((32 MiB - 8448) / (8 (n + max(r, 1)))) rows; when one row does not
 * fit (n + max(r, 1) > 4193248) the call fails with SKEMA_ERR_SHAPE before
 * any kernel runs.  Every call on one GPU shares that array, so calls that
 * overlap on one GPU (two streams or two host threads) race.  Results equal
 * skema_residual_fro_dense's, bit for bit when `window` rows fit.  A context
 * must have been created on the GPU first (skema_ctx_create sets the array
 * up). */
int skema_residual_fro_dense_noalloc(skema_ctx* ctx, int64_t m, int64_t n, int64_t r, const double* A, int64_t lda,
                                     const double* U, int64_t ldu, const double* S, const double* V, int64_t ldv,
                                     int64_t window, double* err2, double* norm2);

/* ---- residual norms of the final factors ---------------------------------- */
/* rnorms[c] = ||[A^T u_c - s_c v_c ; A v_c - s_c u_c]||_2 for c < r: the
 * reference's residuals(A, U, S, V, rank, algParams, window)
 * (src/Skema_Residuals.hpp:160-217, windowed A V / A^T U at :60-121), called
 * by SketchySVD::compute_residuals (src/Skema_SketchySVD.cpp:718-727).
 * U (m x r, ldu), S (r), V (n x r, ldv) on the device; rnorms (r) on the host. */
int skema_residuals_dense(skema_ctx* ctx, int64_t m, int64_t n, int64_t r, const double* A, int64_t lda,
                          const double* U, int64_t ldu, const double* S, const double* V, int64_t ldv,
                          double* rnorms);
/* Same for a device CSR A (m x n; rowptr int64[m+1], col int32, val). */
int skema_residuals_csr(skema_ctx* ctx, int64_t m, int64_t n, int64_t r, const int64_t* rowptr, const int32_t* col,
                        const double* val, const double* U, int64_t ldu, const double* S, const double* V,
                        int64_t ldv, double* rnorms);
/* Same for A = the n x n Gaussian RBF matrix of the device point set
 * (n x d, ldp), generated in row windows by the fused generator (the
 * GaussRBFWindow path of map_A_by_window, src/Skema_Residuals.hpp:60-121). */
int skema_residuals_rbf(skema_ctx* ctx, const double* points, int64_t n, int64_t d, int64_t ldp, double gamma,
                        int64_t r, const double* U, int64_t ldu, const double* S, const double* V, int64_t ldv,
                        int64_t window, double* rnorms);
/* Eigenpair residuals rnorms[c] = ||A v_c - s_c v_c||_2 of a symmetric A
 * (residuals(A, V, S, rank, ...), src/Skema_Residuals.hpp:123-158, used by
 * SketchySPD). */
int skema_residuals_sym_dense(skema_ctx* ctx, int64_t n, int64_t r, const double* A, int64_t lda, const double* V,
                              int64_t ldv, const double* S, double* rnorms);
int skema_residuals_sym_csr(skema_ctx* ctx, int64_t n, int64_t r, const int64_t* rowptr, const int32_t* col,
                            const double* val, const double* V, int64_t ldv, const double* S, double* rnorms);
int skema_residuals_sym_rbf(skema_ctx* ctx, const double* points, int64_t n, int64_t d, int64_t ldp, double gamma,
                            int64_t r, const double* V, int64_t ldv, const double* S, int64_t window,
                            double* rnorms);

/* ---- adjoint of the fused RBF window + Gauss sketch ----------------------- */
/* Loss L = ||K Omega^T||_F^2 + ||Phi K Psi^T||_F^2 with K = exp(-gamma D) over
 * the n points (n x d, ldp, device).  Omega (k x n), Phi (s x n), Psi (s x n)
 * are dense maps.  Outputs: *loss, dX (n x d, ldg, device), *dgamma. Kernel
 * tiles are recomputed in the backward pass, never stored. */
int skema_rbf_sketch_loss_grad(skema_ctx* ctx, const double* points, int64_t n, int64_t d, int64_t ldp,
                               double gamma, const skema_map* omega, const skema_map* phi, const skema_map* psi,
                               int64_t window, int precision, double* loss, double* dX, int64_t ldg,
                               double* dgamma);

#ifdef __cplusplus
}
#endif
#endif /* SKEMA_MI355X_H */
