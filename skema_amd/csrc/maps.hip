// DimRedux maps: construction (uploaded or seeded) and the lmap / rmap
// products with dense operands (src/Skema_DimRedux.hpp:60-80,
// src/Skema_DimRedux_Gauss.cpp:10-52, src/Skema_DimRedux_SparseSign.cpp:9-51).
#include <algorithm>
#include <cmath>
#include <numeric>

#include "sketch.hpp"
#include "sparse.hpp"
#include "util.hpp"

namespace skema {
namespace {

// Counter-based generator: splitmix64 finaliser over key(seed) + (c+1)*golden.
__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__host__ __device__ inline double counter_uniform(uint64_t seed, uint64_t c) {
  const uint64_t key = mix64(seed * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull);
  const uint64_t h = mix64(key + (c + 1) * 0x9E3779B97F4A7C15ull);
  return ((double)(h >> 11) + 0.5) * 0x1.0p-53;
}

__global__ void k_gauss_fill(double* out, int64_t nrow, int64_t ncol, uint64_t seed, int dist, double a) {
  const int64_t total = nrow * ncol;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
    double v;
    if (dist == SKEMA_DIST_NORMAL) {
      const double u1 = counter_uniform(seed, 2 * (uint64_t)e);
      const double u2 = counter_uniform(seed, 2 * (uint64_t)e + 1);
      v = sqrt(-2.0 * log(u1)) * cos(2.0 * M_PI * u2);
    } else {
      v = a * (2.0 * counter_uniform(seed, (uint64_t)e) - 1.0);
    }
    out[e] = v;  // e = i + j*nrow: stored column-major with ld = nrow
  }
}

// Sort each row's (col, val) by column (stable), so window sub-ranges of a
// row are contiguous.
void sort_rows(int64_t rows, const int64_t* rp, std::vector<int32_t>& ci, std::vector<double>& v) {
  std::vector<int64_t> idx;
  std::vector<int32_t> tc;
  std::vector<double> tv;
  for (int64_t r = 0; r < rows; ++r) {
    const int64_t lo = rp[r] - rp[0], hi = rp[r + 1] - rp[0];
    bool sorted = true;
    for (int64_t e = lo + 1; e < hi && sorted; ++e) sorted = ci[e - 1] <= ci[e];
    if (sorted) continue;
    idx.resize(hi - lo);
    std::iota(idx.begin(), idx.end(), lo);
    std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return ci[a] < ci[b]; });
    tc.resize(hi - lo);
    tv.resize(hi - lo);
    for (int64_t t = 0; t < hi - lo; ++t) { tc[t] = ci[idx[t]]; tv[t] = v[idx[t]]; }
    std::copy(tc.begin(), tc.end(), ci.begin() + lo);
    std::copy(tv.begin(), tv.end(), v.begin() + lo);
  }
}

}  // namespace

double now_seconds() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void build_csr_transpose_host(int64_t rows, int64_t cols, int64_t nnz, const int64_t* rp, const int32_t* ci,
                              const double* v, std::vector<int64_t>& trp, std::vector<int32_t>& tci,
                              std::vector<double>& tv) {
  trp.assign(cols + 1, 0);
  for (int64_t e = 0; e < nnz; ++e) trp[ci[e] + 1]++;
  for (int64_t c = 0; c < cols; ++c) trp[c + 1] += trp[c];
  std::vector<int64_t> cur(trp.begin(), trp.end() - 1);
  tci.resize(nnz);
  tv.resize(nnz);
  for (int64_t r = 0; r < rows; ++r)
    for (int64_t e = rp[r] - rp[0]; e < rp[r + 1] - rp[0]; ++e) {
      const int64_t pos = cur[ci[e]]++;
      tci[pos] = (int32_t)r;
      tv[pos] = v[e];
    }
}

void upload_csr(Csr& c, int64_t rows, int64_t cols, int64_t nnz, const int64_t* rp, const int32_t* ci,
                const double* v, hipStream_t st) {
  c.rows = rows;
  c.cols = cols;
  c.nnz = nnz;
  std::vector<int64_t> r0(rows + 1);
  for (int64_t r = 0; r <= rows; ++r) r0[r] = rp[r] - rp[0];
  c.rowptr.alloc((rows + 1) * sizeof(int64_t));
  c.col.alloc(std::max<int64_t>(nnz, 1) * sizeof(int32_t));
  c.val.alloc(std::max<int64_t>(nnz, 1) * sizeof(double));
  SKEMA_HIP(hipMemcpyAsync(c.rowptr.p, r0.data(), (rows + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
  if (nnz > 0) {
    SKEMA_HIP(hipMemcpyAsync(c.col.p, ci, nnz * sizeof(int32_t), hipMemcpyHostToDevice, st));
    SKEMA_HIP(hipMemcpyAsync(c.val.p, v, nnz * sizeof(double), hipMemcpyHostToDevice, st));
  }
  SKEMA_HIP(hipStreamSynchronize(st));
}

// Builds a sparse map from host CSR (validated, rows sorted), plus the CSR of
// its transpose.
Map* make_sparse_map(Ctx* ctx, int64_t nrow, int64_t ncol, int64_t nnz, const int64_t* rp, const int32_t* ci_in,
                     const double* v_in, bool transposed) {
  SKEMA_REQUIRE(nrow > 0 && ncol > 0 && nnz >= 0, kErrShape, "map: empty or negative shape");
  SKEMA_REQUIRE(ncol < ((int64_t)1 << 31) && nrow < ((int64_t)1 << 31), kErrShape, "map: dimension above 2^31");
  SKEMA_REQUIRE(rp[nrow] - rp[0] == nnz, kErrArg, "map: rowptr[nrow]-rowptr[0] != nnz");
  for (int64_t r = 0; r < nrow; ++r) SKEMA_REQUIRE(rp[r + 1] >= rp[r], kErrArg, "map: rowptr not non-decreasing");
  std::vector<int32_t> ci(ci_in, ci_in + nnz);
  std::vector<double> v(v_in, v_in + nnz);
  for (int64_t e = 0; e < nnz; ++e) SKEMA_REQUIRE(ci[e] >= 0 && ci[e] < ncol, kErrArg, "map: column index out of range");
  sort_rows(nrow, rp, ci, v);
  std::unique_ptr<Map> m(new Map);
  m->ctx = ctx;
  m->nrow = nrow;
  m->ncol = ncol;
  m->transposed = transposed;
  m->sparse = true;
  upload_csr(m->csr, nrow, ncol, nnz, rp, ci.data(), v.data(), ctx->stream);
  std::vector<int64_t> trp;
  std::vector<int32_t> tci;
  std::vector<double> tv;
  build_csr_transpose_host(nrow, ncol, nnz, rp, ci.data(), v.data(), trp, tci, tv);
  upload_csr(m->csr_t, ncol, nrow, nnz, trp.data(), tci.data(), tv.data(), ctx->stream);
  // ELL view of the logical transpose: rows of the stored CSR when the map
  // is stored transposed, rows of its transpose otherwise
  {
    const bool st = transposed;
    const int64_t lc = st ? nrow : ncol;  // logical columns = rows of the logical transpose
    const int64_t* p = st ? rp : trp.data();
    const int32_t* c = st ? ci.data() : tci.data();
    const double* val = st ? v.data() : tv.data();
    for (int64_t r = 0; r < lc; ++r) m->sp_zmax = std::max<int64_t>(m->sp_zmax, p[r + 1] - p[r]);
    for (int64_t e = 0; e < nnz; ++e) m->sp_vmax = std::max(m->sp_vmax, std::fabs(val[e]));
    const int64_t z = lc > 0 ? p[1] - p[0] : 0;
    bool ok = z > 0 && z <= 16 && nnz == z * lc && val[0] != 0.0;
    const double mag = ok ? std::fabs(val[0]) : 0.0;
    for (int64_t r = 0; ok && r < lc; ++r) ok = (p[r + 1] - p[r] == z);
    for (int64_t e = 0; ok && e < nnz; ++e) ok = (std::fabs(val[e]) == mag);
    if (ok) {
      std::vector<uint32_t> h((size_t)nnz);
      for (int64_t e = 0; e < nnz; ++e)  // entry e is (row e / z, slot e % z): rows hold z entries each
        h[(size_t)e] = (uint32_t)c[e] | (val[e] < 0.0 ? 0x80000000u : 0u);
      m->ell.alloc((size_t)nnz * sizeof(uint32_t));
      SKEMA_HIP(hipMemcpyAsync(m->ell.p, h.data(), (size_t)nnz * sizeof(uint32_t), hipMemcpyHostToDevice,
                               ctx->stream));
      SKEMA_HIP(hipStreamSynchronize(ctx->stream));
      m->ell_zeta = (int)z;
      m->ell_mag = mag;
      if ((st ? ncol : nrow) <= 0x7fff) {  // every row index fits 15 bits
        std::vector<uint16_t> h16((size_t)nnz);
        for (int64_t e = 0; e < nnz; ++e) h16[(size_t)e] = (uint16_t)(c[e] | (val[e] < 0.0 ? 0x8000 : 0));
        m->ell16.alloc((size_t)nnz * sizeof(uint16_t));
        SKEMA_HIP(hipMemcpyAsync(m->ell16.p, h16.data(), (size_t)nnz * sizeof(uint16_t), hipMemcpyHostToDevice,
                                 ctx->stream));
        SKEMA_HIP(hipStreamSynchronize(ctx->stream));
      }
    }
  }
  return m.release();
}

Map* make_dense_map(Ctx* ctx, int64_t nrow, int64_t ncol, const double* src, int64_t ld, bool src_device,
                    bool transposed) {
  SKEMA_REQUIRE(nrow > 0 && ncol > 0, kErrShape, "map: empty or negative shape");
  SKEMA_REQUIRE(ld >= nrow, kErrShape, "map: ld < nrow");
  std::unique_ptr<Map> m(new Map);
  m->ctx = ctx;
  m->nrow = nrow;
  m->ncol = ncol;
  m->transposed = transposed;
  m->dense.alloc((size_t)nrow * ncol * sizeof(double));
  if (src) {
    SKEMA_HIP(hipMemcpy2DAsync(m->dense.p, nrow * sizeof(double), src, ld * sizeof(double), nrow * sizeof(double),
                               ncol, src_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, ctx->stream));
    SKEMA_HIP(hipStreamSynchronize(ctx->stream));
  }
  return m.release();
}

void gauss_fill(Ctx* ctx, Map* m, uint64_t seed, int dist) {
  const double a = std::sqrt(2.0 * std::log((double)m->nrow * (double)m->ncol));
  const int64_t total = m->nrow * m->ncol;
  int64_t nb = cdiv(total, 256);
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(k_gauss_fill, dim3((unsigned)nb), dim3(256), 0, ctx->stream, m->dense.as<double>(), m->nrow,
                     m->ncol, seed, dist, a);
  check_launch("k_gauss_fill");
  SKEMA_HIP(hipStreamSynchronize(ctx->stream));
}

// Sparse-sign materialisation on the host (same counter stream as above).
void sparse_sign_host(int64_t nrow, int64_t ncol, uint64_t seed, int convention, int64_t zeta, bool transposed,
                      std::vector<int64_t>& rp, std::vector<int32_t>& ci, std::vector<double>& v) {
  if (convention == SKEMA_SSIGN_ROWS) {
    // zeta entries per stored row, uniform columns with replacement, sorted,
    // exact +-1 (src/Skema_DimRedux.hpp:205-251).
    const int64_t z = zeta > 0 ? zeta : std::max<int64_t>(2, std::min<int64_t>(ncol, 8));
    rp.resize(nrow + 1);
    ci.resize(nrow * z);
    v.resize(nrow * z);
    for (int64_t r = 0; r <= nrow; ++r) rp[r] = r * z;
    for (int64_t r = 0; r < nrow; ++r) {
      for (int64_t t = 0; t < z; ++t) {
        int64_t c = (int64_t)(counter_uniform(seed, (uint64_t)(r * z + t)) * (double)ncol);
        ci[r * z + t] = (int32_t)std::min<int64_t>(c, ncol - 1);
      }
      std::sort(ci.begin() + r * z, ci.begin() + (r + 1) * z);
      for (int64_t t = 0; t < z; ++t)
        v[r * z + t] = counter_uniform(seed, (uint64_t)(r * z + t + nrow * z)) < 0.5 ? -1.0 : 1.0;
    }
    return;
  }
  // zeta distinct rows per column of the logical map (Floyd sampling),
  // values +-1/sqrt(zeta).  Logical shape: (lr x lc).
  const int64_t lr = transposed ? ncol : nrow, lc = transposed ? nrow : ncol;
  const int64_t z = std::min<int64_t>(zeta > 0 ? zeta : 8, lr);
  const double mag = 1.0 / std::sqrt((double)z);
  std::vector<int64_t> sel(z);
  // Entries of the logical map grouped by logical column.
  std::vector<int32_t> lrow(lc * z);
  std::vector<double> lval(lc * z);
  for (int64_t c = 0; c < lc; ++c) {
    for (int64_t t = 0; t < z; ++t) {
      const int64_t j = lr - z + t;
      const double u = counter_uniform(seed, (uint64_t)(c * z + t));
      int64_t cand = std::min<int64_t>((int64_t)(u * (double)(j + 1)), j);
      bool dup = false;
      for (int64_t q = 0; q < t; ++q) dup |= (sel[q] == cand);
      sel[t] = dup ? j : cand;
    }
    std::sort(sel.begin(), sel.end());
    for (int64_t t = 0; t < z; ++t) {
      lrow[c * z + t] = (int32_t)sel[t];
      lval[c * z + t] = counter_uniform(seed ^ 0x5bd1e995ull, (uint64_t)(c * z + t)) < 0.5 ? -mag : mag;
    }
  }
  if (transposed) {
    // stored = logical^T: stored row c holds logical column c.
    rp.resize(lc + 1);
    for (int64_t c = 0; c <= lc; ++c) rp[c] = c * z;
    ci = lrow;
    v = lval;
  } else {
    // stored = logical: CSR rows = logical rows; build from column groups.
    std::vector<int64_t> colptr(lc + 1);
    for (int64_t c = 0; c <= lc; ++c) colptr[c] = c * z;
    build_csr_transpose_host(lc, lr, lc * z, colptr.data(), lrow.data(), lval.data(), rp, ci, v);
  }
}

// ---- lmap / rmap -----------------------------------------------------------
void map_lmap(Map& mp, double alpha, const double* B, int64_t brows, int64_t bcols, int64_t ldb, double beta,
              char transA, char transB, int64_t i0, int64_t i1, double* C, int64_t ldc) {
  Ctx* ctx = mp.ctx;
  const bool win = (i0 != i1);
  SKEMA_REQUIRE(!win || (i0 >= 0 && i0 < i1 && i1 <= mp.ncol), kErrShape, "lmap: idx outside the map's columns");
  const int64_t wc = win ? i1 - i0 : mp.ncol;  // columns of map[:, idx]
  char tA = (transA == 't') ? 'T' : (transA == 'n' ? 'N' : transA);
  if (mp.transposed) tA = (tA == 'N') ? 'T' : 'N';  // Gauss.cpp:18-20, SparseSign.cpp:18-20
  const int64_t M = (tA == 'N') ? mp.nrow : wc;
  const int64_t Kd = (tA == 'N') ? wc : mp.nrow;
  if (!mp.sparse) {
    const bool tb = (transB == 'T' || transB == 't');
    const int64_t K = tb ? bcols : brows, N = tb ? brows : bcols;
    SKEMA_REQUIRE(K == Kd, kErrShape, "lmap: inner dimensions differ");
    SKEMA_REQUIRE(ldc >= M, kErrShape, "lmap: ldc too small");
    const double* data = mp.dense.as<double>() + (win ? i0 * mp.nrow : 0);
    gemm_f64(ctx->stream, tA, tb ? 'T' : 'N', M, N, K, alpha, data, mp.nrow, B, ldb, beta, C, ldc, &ctx->gws);
    return;
  }
  // SparseSign: spmv(transA) with transB ignored (src/Skema_Common.hpp:42-50).
  SKEMA_REQUIRE(brows == Kd, kErrShape, "lmap: inner dimensions differ");
  SKEMA_REQUIRE(ldc >= M, kErrShape, "lmap: ldc too small");
  const int64_t c0 = win ? i0 : 0, c1 = win ? i1 : mp.ncol;
  if (tA == 'N')
    spmm_csr_rows(ctx->stream, mp.csr, 0, mp.nrow, c0, c1, alpha, B, ldb, bcols, beta, C, ldc);
  else
    spmm_csr_rows(ctx->stream, mp.csr_t, c0, c1, 0, mp.nrow, alpha, B, ldb, bcols, beta, C, ldc);
}

void map_rmap(Map& mp, double alpha, const double* A, int64_t arows, int64_t acols, int64_t lda, double beta,
              char transA, char transB, int64_t i0, int64_t i1, double* C, int64_t ldc) {
  Ctx* ctx = mp.ctx;
  (void)i0;
  (void)i1;  // idx is ignored by the dense-operand rmap (Gauss.cpp:35-52)
  const bool tb = (transB == 'T' || transB == 't');
  if (!mp.sparse) {
    const bool ta = (transA == 'T' || transA == 't');
    const int64_t M = ta ? acols : arows, K = ta ? arows : acols;
    const int64_t Kd = tb ? mp.ncol : mp.nrow, N = tb ? mp.nrow : mp.ncol;
    SKEMA_REQUIRE(K == Kd, kErrShape, "rmap: inner dimensions differ");
    SKEMA_REQUIRE(ldc >= M, kErrShape, "rmap: ldc too small");
    gemm_f64(ctx->stream, ta ? 'T' : 'N', tb ? 'T' : 'N', M, N, K, alpha, A, lda, mp.dense.as<double>(), mp.nrow,
             beta, C, ldc, &ctx->gws);
    return;
  }
  // SparseSign: (op_swap(transB)(map) * A^T)^T = A * op_transB(map); transA is
  // ignored by the reference's spmv dispatch (SparseSign.cpp:33-51).
  const Csr& S = tb ? mp.csr : mp.csr_t;  // rows of S = columns of the product
  SKEMA_REQUIRE(acols == S.cols, kErrShape, "rmap: inner dimensions differ");
  SKEMA_REQUIRE(ldc >= arows, kErrShape, "rmap: ldc too small");
  GemmOut o;
  o.C1 = C;
  o.ldc1 = ldc;
  o.alpha1 = alpha;
  o.beta1 = beta;
  o.nsplit = S.rows;
  o.C2 = C;
  o.ldc2 = ldc;
  spmm_dense_csrT(ctx->stream, arows, A, lda, S, o);
}

}  // namespace skema
