// SketchySVD window update and fixed-rank finalize on one GPU.
//
// Per window H (rows [row0, row0+w)) with logical maps Upsilon (k x m),
// Omega (k x n), Phi (s x m), Psi (s x n):
//   X        <- nu*X        + eta * Upsilon[:, idx] * H          (k x n)
//   Y[idx,:] <- nu*Y[idx,:] + eta * H * Omega^T                  (w x k)
//   Z        <- nu*Z        + eta * Phi[:, idx] * (H * Psi^T)    (s x s)
// (src/Skema_SketchySVD.cpp:167-189 with update() :255-335).  The reference
// forms W = Phi[:,idx]*H and then W*Psi^T; the same Z is formed here as
// Phi[:,idx]*(H*Psi^T) so H*Omega^T and H*Psi^T share one pass over H
// (one GEMM against Omega stacked over Psi) and the s x n intermediate
// disappears.  The nu/eta axpy's are folded into the GEMM epilogues.
#include <cmath>

#include "linalg.hpp"
#include "lowp.hpp"
#include "rbf.hpp"
#include "rbf_fused.hpp"
#include "rbf_lowp.hpp"
#include "sketch.hpp"
#include "sparse.hpp"
#include "util.hpp"

namespace skema {
namespace {

std::vector<int64_t> host_rowptr(const Csr& c, hipStream_t st) {
  std::vector<int64_t> rp(c.rows + 1);
  SKEMA_HIP(hipMemcpyAsync(rp.data(), c.rp(), (c.rows + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, st));
  SKEMA_HIP(hipStreamSynchronize(st));
  return rp;
}

void download_csr(const Csr& c, hipStream_t st, std::vector<int64_t>& rp, std::vector<int32_t>& ci,
                  std::vector<double>& v) {
  rp = host_rowptr(c, st);
  ci.resize(c.nnz);
  v.resize(c.nnz);
  if (c.nnz) {
    SKEMA_HIP(hipMemcpyAsync(ci.data(), c.ci(), c.nnz * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    SKEMA_HIP(hipMemcpyAsync(v.data(), c.v(), c.nnz * sizeof(double), hipMemcpyDeviceToHost, st));
    SKEMA_HIP(hipStreamSynchronize(st));
  }
}

// Row-concatenation [A; B] of two CSR matrices with the same column count.
void stack_csr(const Csr& A, const Csr& B, hipStream_t st, Csr& out) {
  std::vector<int64_t> ra, rb;
  std::vector<int32_t> ca, cb;
  std::vector<double> va, vb;
  download_csr(A, st, ra, ca, va);
  download_csr(B, st, rb, cb, vb);
  std::vector<int64_t> rp(A.rows + B.rows + 1);
  for (int64_t r = 0; r <= A.rows; ++r) rp[r] = ra[r];
  for (int64_t r = 1; r <= B.rows; ++r) rp[A.rows + r] = ra[A.rows] + rb[r];
  ca.insert(ca.end(), cb.begin(), cb.end());
  va.insert(va.end(), vb.begin(), vb.end());
  upload_csr(out, A.rows + B.rows, A.cols, A.nnz + B.nnz, rp.data(), ca.data(), va.data(), st);
}

}  // namespace

// Logical dense views (pointer, ld) of the maps, copied once when a map is
// stored transposed, and Omega stacked over Psi.
void prepare_sketch_maps(Sketch& sk) {
  Ctx* c = sk.ctx;
  hipStream_t st = c->stream;
  const int64_t k = sk.k, s = sk.s, m = sk.p.m, n = sk.p.n;
  Map** mp = sk.maps;
  const int64_t want_r[4] = {k, k, s, s}, want_c[4] = {m, n, m, n};
  const char* names[4] = {"Upsilon", "Omega", "Phi", "Psi"};
  for (int i = 0; i < 4; ++i) {
    SKEMA_REQUIRE(mp[i] != nullptr, kErrArg, std::string("sketch: missing map ") + names[i]);
    SKEMA_REQUIRE(mp[i]->ctx == c, kErrArg, std::string("sketch: map from another context: ") + names[i]);
    SKEMA_REQUIRE(mp[i]->lrows() == want_r[i] && mp[i]->lcols() == want_c[i], kErrShape,
                  std::string("sketch: map ") + names[i] + " has shape " + std::to_string(mp[i]->nrow) + "x" +
                      std::to_string(mp[i]->ncol) + (mp[i]->transposed ? " (transposed)" : "") + ", expected " +
                      std::to_string(want_r[i]) + "x" + std::to_string(want_c[i]));
    SKEMA_REQUIRE(mp[i]->sparse == mp[0]->sparse, kErrArg, "sketch: maps must all be dense or all sparse");
  }
  sk.sparse_maps = mp[0]->sparse;
  SKEMA_REQUIRE(sk.sparse_maps == (sk.p.dim_redux == SKEMA_MAP_SPARSE_SIGN), kErrArg,
                "sketch: map storage does not match dim_redux");
  if (sk.sparse_maps) {
    stack_csr(mp[1]->logical_csr(), mp[3]->logical_csr(), st, sk.sp_omps);
    // Transposed stack: row j lists column j of Omega (p) and of Psi (k+p).
    std::vector<int64_t> ro, rs;
    std::vector<int32_t> co, cs;
    std::vector<double> vo, vs;
    download_csr(mp[1]->logical_csr_t(), st, ro, co, vo);
    download_csr(mp[3]->logical_csr_t(), st, rs, cs, vs);
    std::vector<int64_t> rp(n + 1, 0);
    std::vector<int32_t> ci;
    std::vector<double> v;
    ci.reserve(co.size() + cs.size());
    v.reserve(co.size() + cs.size());
    for (int64_t j = 0; j < n; ++j) {
      for (int64_t e = ro[j]; e < ro[j + 1]; ++e) { ci.push_back(co[e]); v.push_back(vo[e]); }
      for (int64_t e = rs[j]; e < rs[j + 1]; ++e) { ci.push_back((int32_t)(k + cs[e])); v.push_back(vs[e]); }
      rp[j + 1] = (int64_t)ci.size();
    }
    upload_csr(sk.sp_ompsT, n, k + s, (int64_t)ci.size(), rp.data(), ci.data(), v.data(), st);
    return;
  }
  auto logical = [&](Map* m, DevBuf& own, const double*& ptr, int64_t& ld) {
    if (!m->transposed) {
      ptr = m->dense.as<double>();
      ld = m->nrow;
    } else {
      own.alloc((size_t)m->ncol * m->nrow * sizeof(double));
      transpose(st, m->nrow, m->ncol, m->dense.as<double>(), m->nrow, own.as<double>(), m->ncol);
      ptr = own.as<double>();
      ld = m->ncol;
    }
  };
  logical(mp[0], sk.ups, sk.upsL, sk.ldups);
  logical(mp[2], sk.phi, sk.phiL, sk.ldphi);
  sk.omps.alloc((size_t)(k + s) * n * sizeof(double));
  double* om = sk.omps.as<double>();
  for (int i : {1, 3}) {
    Map* m = mp[i];
    double* dst = om + (i == 3 ? k : 0);
    if (!m->transposed) copy_matrix(st, m->nrow, m->ncol, m->dense.as<double>(), m->nrow, dst, k + s);
    else transpose(st, m->nrow, m->ncol, m->dense.as<double>(), m->nrow, dst, k + s);
  }
  SKEMA_HIP(hipStreamSynchronize(st));
  if (sk.p.precision != SKEMA_PREC_FP64) prepare_lowp_maps(sk);
}

void update_dense_window(Sketch& sk, const double* H, int64_t ldh, int64_t row0, int64_t rows) {
  Ctx* c = sk.ctx;
  hipStream_t st = c->stream;
  const int64_t k = sk.k, s = sk.s, m = sk.p.m, n = sk.p.n;
  const double eta = sk.p.eta, nu = sk.p.nu;
  sk.G.reserve((size_t)rows * s * sizeof(double));
  double* G = sk.G.as<double>();
  double* X = sk.X.as<double>();
  double* Y = sk.Y.as<double>();
  double* Z = sk.Z.as<double>();
  GemmOut o;
  o.C1 = Y + row0;
  o.ldc1 = m;
  o.alpha1 = eta;
  o.beta1 = nu;
  o.nsplit = k;
  o.C2 = G;
  o.ldc2 = rows;
  o.alpha2 = 1.0;
  o.beta2 = 0.0;
  if (sk.sparse_maps) {
    const Csr& U = sk.maps[0]->logical_csr();
    const Csr& F = sk.maps[2]->logical_csr();
    {
      Timed t(c, "sketch_left");
      spmm_csr_rows(st, U, 0, k, row0, row0 + rows, eta, H, ldh, n, nu, X, k);
    }
    {
      Timed t(c, "sketch_right");
      spmm_dense_csrT(st, rows, H, ldh, sk.sp_omps, o);
    }
    {
      Timed t(c, "sketch_core");
      spmm_csr_rows(st, F, 0, s, row0, row0 + rows, eta, G, rows, s, nu, Z, s);
    }
    return;
  }
  if (sk.p.precision != SKEMA_PREC_FP64) {
    update_dense_window_lowp(sk, H, ldh, row0, rows);
    return;
  }
  {
    Timed t(c, "sketch_left");
    gemm_f64(st, 'N', 'N', k, n, rows, eta, sk.upsL + row0 * sk.ldups, sk.ldups, H, ldh, nu, X, k, &c->gws);
  }
  {
    Timed t(c, "sketch_right");
    gemm_f64(st, 'N', 'T', rows, k + s, n, H, ldh, sk.omps.as<double>(), k + s, o, &c->gws);
  }
  {
    Timed t(c, "sketch_core");
    gemm_f64(st, 'N', 'N', s, s, rows, eta, sk.phiL + row0 * sk.ldphi, sk.ldphi, G, rows, nu, Z, s, &c->gws);
  }
}

void update_csr_window(Sketch& sk, const int64_t* rowptr, const int32_t* col, const double* val, int64_t row0,
                       int64_t rows) {
  Ctx* c = sk.ctx;
  hipStream_t st = c->stream;
  const int64_t k = sk.k, s = sk.s, m = sk.p.m, n = sk.p.n;
  const double eta = sk.p.eta, nu = sk.p.nu;
  sk.G.reserve((size_t)rows * s * sizeof(double));
  double* G = sk.G.as<double>();
  double* X = sk.X.as<double>();
  double* Y = sk.Y.as<double>();
  double* Z = sk.Z.as<double>();
  GemmOut o;
  o.C1 = Y + row0;
  o.ldc1 = m;
  o.alpha1 = eta;
  o.beta1 = nu;
  o.nsplit = k;
  o.C2 = G;
  o.ldc2 = rows;
  o.alpha2 = 1.0;
  o.beta2 = 0.0;
  CscWindow Hc;
  {
    Timed t(c, "window");
    Hc = csr_window_to_csc(st, rowptr, col, val, rows, n, sk.win);
  }
  if (sk.sparse_maps) {
    {
      Timed t(c, "sketch_left");
      csc_left_sparse(st, Hc, sk.maps[0]->logical_csr_t(), row0, k, eta, nu, X, k);
    }
    {
      Timed t(c, "sketch_right");
      csr_right_sparse(st, rowptr, col, val, rows, sk.sp_ompsT, k + s, o);
    }
    {
      Timed t(c, "sketch_core");
      spmm_csr_rows(st, sk.maps[2]->logical_csr(), 0, s, row0, row0 + rows, eta, G, rows, s, nu, Z, s);
    }
    return;
  }
  {
    Timed t(c, "sketch_left");
    csc_left_dense(st, Hc, sk.upsL, sk.ldups, row0, k, eta, nu, X, k);
  }
  {
    Timed t(c, "sketch_right");
    csr_right_dense(st, rowptr, col, val, rows, sk.omps.as<double>(), k + s, k + s, o);
  }
  {
    Timed t(c, "sketch_core");
    gemm_f64(st, 'N', 'N', s, s, rows, eta, sk.phiL + row0 * sk.ldphi, sk.ldphi, G, rows, nu, Z, s, &c->gws);
  }
}

void update_rbf_window(Sketch& sk, const double* points, int64_t ldp, int64_t row0, int64_t rows) {
  Ctx* c = sk.ctx;
  const int64_t n = sk.p.n;
  if (!sk.sparse_maps && rbf_fused_supported(sk)) {
    update_rbf_window_fused(sk, points, ldp, row0, rows);
    return;
  }
  if (rbf_bf16_supported(sk)) {
    update_rbf_window_bf16(sk, points, ldp, row0, rows);
    return;
  }
  sk.win.reserve((size_t)rows * n * sizeof(double));
  {
    Timed t(c, "window");
    rbf_window(c->stream, points, n, sk.p.nfeat, ldp, sk.p.gamma, row0, rows, sk.win.as<double>(), rows);
  }
  update_dense_window(sk, sk.win.as<double>(), rows, row0, rows);
}

// ---- finalize ---------------------------------------------------------------
namespace {
struct Stopwatch {
  hipStream_t st;
  double t0;
  explicit Stopwatch(hipStream_t s) : st(s) { t0 = now_seconds(); }
  double lap() {
    SKEMA_HIP(hipStreamSynchronize(st));
    const double t = now_seconds(), d = t - t0;
    t0 = t;
    return d;
  }
};
}  // namespace

// initial_approx + low_rank_approx (src/Skema_SketchySVD.cpp:337-664); every
// product and factorisation in FP64 on the GPU.
void finalize_sketch(Sketch& sk) {
  Ctx* c = sk.ctx;
  hipStream_t st = c->stream;
  const int64_t k = sk.k, s = sk.s, m = sk.p.m, n = sk.p.n;
  SKEMA_REQUIRE(k <= m && k <= n, kErrShape, "low_rank_approx: range must not exceed min(m, n) (QR of Y and X^T)");
  SKEMA_REQUIRE(s >= k, kErrShape, "low_rank_approx: core must be >= range (QR of Phi*Q and Psi*P)");
  auto D = [](DevBuf& b, int64_t count) {
    b.reserve((size_t)std::max<int64_t>(count, 1) * sizeof(double));
    return b.as<double>();
  };
  double* P = D(sk.P, n * k);
  double* Q = D(sk.Q, m * k);
  double* PhiQ = D(sk.PhiQ, s * k);
  double* PsiP = D(sk.PsiP, s * k);
  double* T1 = D(sk.T1, k * k);
  double* T2 = D(sk.T2, k * k);
  double* C = D(sk.C, k * k);
  double* Uc = D(sk.Uc, k * k);
  double* Sv = D(sk.Sv, k);
  double* Vc = D(sk.Vc, k * k);
  double* tau = D(sk.tau, k);
  const int64_t wlen = std::max<int64_t>(std::max(m, n) * k, 4 * s * k + 3 * k * k);
  double* W = D(sk.work, wlen);
  auto& tm = sk.timings["approx"];
  for (const char* key : {"phi", "psi", "dgeqrf", "dgemm", "dgels", "dgesvd"}) tm.emplace(key, 0.0);
  Stopwatch sw(st);
  sw.lap();

  // [P,~] = qr(X^T,0)  (:355-357)
  transpose(st, k, n, sk.X.as<double>(), k, W, n);
  householder_qr(st, n, k, W, n, tau, &c->qr, &c->gws);
  householder_q(st, n, k, W, n, tau, P, n, &c->qr, &c->gws, true);
  // [Q,~] = qr(Y,0)  (:381-383), on a copy: Y stays intact.
  copy_matrix(st, m, k, sk.Y.as<double>(), m, W, m);
  householder_qr(st, m, k, W, m, tau, &c->qr, &c->gws);
  householder_q(st, m, k, W, m, tau, Q, m, &c->qr, &c->gws, true);
  tm["dgeqrf"] += sw.lap();

  // U1 = Phi*Q, U2 = Psi*P  (:409, :430)
  if (sk.sparse_maps) {
    spmm_csr_rows(st, sk.maps[2]->logical_csr(), 0, s, 0, m, 1.0, Q, m, k, 0.0, PhiQ, s);
    tm["phi"] += sw.lap();
    spmm_csr_rows(st, sk.maps[3]->logical_csr(), 0, s, 0, n, 1.0, P, n, k, 0.0, PsiP, s);
    tm["psi"] += sw.lap();
  } else {
    gemm_f64(st, 'N', 'N', s, k, m, 1.0, sk.phiL, sk.ldphi, Q, m, 0.0, PhiQ, s, &c->gws);
    tm["phi"] += sw.lap();
    gemm_f64(st, 'N', 'N', s, k, n, 1.0, sk.omps.as<double>() + k, k + s, P, n, 0.0, PsiP, s, &c->gws);
    tm["psi"] += sw.lap();
  }

  // [U1,T1] = qr(Phi*Q,0); [U2,T2] = qr(Psi*P,0)  (:453, :476)
  double* F = W;            // s x k factor scratch
  double* U1 = W + s * k;   // s x k
  double* U2 = U1 + s * k;  // s x k
  double* Z1 = U2 + s * k;  // k x s
  double* Z2 = Z1 + k * s;  // k x k
  double* Z2t = Z2 + k * k; // k x k
  double* Cw = Z2t + k * k; // k x k (Jacobi works in place)
  copy_matrix(st, s, k, PhiQ, s, F, s);
  householder_qr(st, s, k, F, s, tau, &c->qr, &c->gws);
  extract_r(st, k, F, s, T1, k);
  householder_q(st, s, k, F, s, tau, U1, s, &c->qr, &c->gws, true);
  copy_matrix(st, s, k, PsiP, s, F, s);
  householder_qr(st, s, k, F, s, tau, &c->qr, &c->gws);
  extract_r(st, k, F, s, T2, k);
  householder_q(st, s, k, F, s, tau, U2, s, &c->qr, &c->gws, true);
  tm["dgeqrf"] += sw.lap();

  // Z2 = (U1^T * Z) * U2  (:500, :515)
  gemm_f64(st, 'T', 'N', k, s, s, 1.0, U1, s, sk.Z.as<double>(), s, 0.0, Z1, k, &c->gws);
  gemm_f64(st, 'N', 'N', k, k, s, 1.0, Z1, k, U2, s, 0.0, Z2, k, &c->gws);
  tm["dgemm"] += sw.lap();

  // Z2 = T1 \ Z2;  C = (T2 \ Z2^T)^T  (:532, :549-562)
  trsm_upper(st, k, k, T1, k, Z2, k);
  transpose(st, k, k, Z2, k, Z2t, k);
  trsm_upper(st, k, k, T2, k, Z2t, k);
  transpose(st, k, k, Z2t, k, C, k);
  tm["dgels"] += sw.lap();

  // [Uc, S, Vc] = svd(C)  (:609).  C = Qc R by Householder QR, then the
  // one-sided Jacobi SVD of R^T = Ur S Vr^T, so C = (Qc Vr) S Ur^T: the QR
  // leaves R graded (row norms decaying), which cuts the Jacobi sweeps about
  // threefold on geometrically decaying spectra.  F, U1, U2 and Z1 (each at
  // least k x k) are free by now.
  {
    double* Qc = F;
    double* Rt = U1;
    double* Ur = U2;
    double* Vr = Z1;
    copy_matrix(st, k, k, C, k, Cw, k);
    householder_qr(st, k, k, Cw, k, tau, &c->qr, &c->gws);
    extract_r(st, k, Cw, k, Z2, k);
    transpose(st, k, k, Z2, k, Rt, k);
    householder_q(st, k, k, Cw, k, tau, Qc, k, &c->qr, &c->gws, true);
    jacobi_svd(st, k, Rt, k, Ur, k, Sv, Vr, k, &c->scratch);
    gemm_f64(st, 'N', 'N', k, k, k, 1.0, Qc, k, Vr, k, 0.0, Uc, k, &c->gws);
    copy_matrix(st, k, k, Ur, k, Vc, k);
  }
  tm["dgesvd"] += sw.lap();
  sk.finalized = true;
}

}  // namespace skema
