"""CPU oracle for Skema's streaming SketchySVD path -- TEST INFRASTRUCTURE ONLY.

This module is the checker, never the product.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only to compute expected values or to time the CPU baseline.
The ``skema_amd`` package never imports it and has no CPU fallback.

What it restates (reference = jeremy-myers/skema, paths relative to its root):

* ``SketchySVD<M,D>`` three-sketch streaming SVD
  (src/Skema_SketchySVD.cpp:15-117 ctor, :119-246 linear_update, :255-335 the
  four ``update`` specialisations, :337-566 initial_approx, :568-664
  low_rank_approx, :666-716 axpy).
* ``GaussDimRedux`` / ``SparseSignDimRedux`` lmap / rmap with the reference's
  storage orientation and mode flips (src/Skema_DimRedux_Gauss.cpp:10-99,
  src/Skema_DimRedux_SparseSign.cpp:9-155, src/Skema_DimRedux.hpp:95-308).
* ``Window::get`` and ``GaussRBFWindow::get`` / ``GaussRBF::compute``
  (src/Skema_Window.hpp:36-107, src/Skema_Kernel.cpp:5-81).
* ``residuals`` / ``map_A_by_window`` (src/Skema_Residuals.hpp:17-217): the
  per-triplet residual norms of the final factors.
* ``SketchySPD`` (src/Skema_SketchySVD.cpp:757-1151): the Nystrom sketch and
  its fixed-rank approximation (two reference defects documented there).
* ``linalg::qr`` / ``ls`` / ``svd`` (src/Skema_BlasLapack.hpp:71-97,168-243):
  the SAME LAPACK routines (dgeqrf+dorgqr with lwork=n, dgels, dgesvd 'S','S')
  called through ``scipy.linalg.lapack`` with the reference's arguments.

Dense products use numpy (an optimised dgemm, the analogue of
KokkosBlas::gemm); sparse products use scipy.sparse (the analogue of
KokkosSparse spmv / spgemm).  All arithmetic is float64, matrices are 2-D
arrays indexed (row, col) exactly like the reference's LayoutLeft views.

Maps are *materialised inputs*: the reference draws them from Kokkos'
XorShift64 pool, whose stream cannot be reproduced offline, so both this
oracle and the GPU library are fed identical materialised maps.  The
generators below (counter-based splitmix64 + Box-Muller, sparse-sign in the
reference's per-row convention and in the per-column convention) are the
oracle's restatement of the seeded generators the GPU library implements.

Pinning (tests/test_oracle_*.py): the reference's own fixtures
(test/data/*_svals.txt, exact-recovery regime), closed forms (X = Upsilon*A,
Y = A*Omega^T, Z = Phi*A*Psi^T, window-size invariance, the nu/eta
recurrence), brute-force loops on tiny inputs, LAPACK invariants
(Q^T Q = I, QR = A, the dlarfg sign rule), and planted-mistake checks.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple

import numpy as np
import scipy.sparse as sp
from scipy.linalg import lapack

# --------------------------------------------------------------------------
# Counter-based generator (restated in skema_amd/csrc/maps.hip)
# --------------------------------------------------------------------------
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_GOLD = np.uint64(0x9E3779B97F4A7C15)
_SALT = np.uint64(0x632BE59BD9B4E019)


def _mix64(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def counter_uniform(seed: int, counters) -> np.ndarray:
    """Doubles in (0,1): splitmix64(key(seed) + (c+1)*golden) -> 53 bits."""
    with np.errstate(over="ignore"):
        c = np.asarray(counters, dtype=np.uint64)
        key = _mix64(np.uint64(seed) * _GOLD + _SALT)
        h = _mix64(key + (c + np.uint64(1)) * _GOLD)
    return ((h >> np.uint64(11)).astype(np.float64) + 0.5) * (2.0 ** -53)


def gauss_map(nrow: int, ncol: int, seed: int, dist: str = "normal") -> np.ndarray:
    """Dense map, entry (i,j) drawn from counter e = i + j*nrow.

    dist="normal": N(0,1) by Box-Muller on counters (2e, 2e+1).
    dist="uniform": U(-a,a), a = sqrt(2 ln(nrow*ncol)) -- the reference's
    GaussDimRedux fill (src/Skema_DimRedux.hpp:112,115) on our counter stream.
    """
    e = np.arange(nrow * ncol, dtype=np.uint64)
    if dist == "normal":
        u1 = counter_uniform(seed, np.uint64(2) * e)
        u2 = counter_uniform(seed, np.uint64(2) * e + np.uint64(1))
        v = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)
    elif dist == "uniform":
        a = math.sqrt(2.0 * math.log(nrow * ncol))
        v = a * (2.0 * counter_uniform(seed, e) - 1.0)
    else:
        raise ValueError(dist)
    return v.reshape(ncol, nrow).T.copy()


def sparse_sign_rows(nrow: int, ncol: int, seed: int) -> sp.csr_matrix:
    """Reference convention (src/Skema_DimRedux.hpp:205-251): zeta =
    max(2, min(ncol, 8)) entries per STORED row, column indices uniform in
    [0, ncol) drawn WITH replacement and sorted, values exactly +-1."""
    zeta = max(2, min(ncol, 8))
    e = np.arange(nrow * zeta, dtype=np.uint64)
    col = np.minimum((counter_uniform(seed, e) * ncol).astype(np.int64), ncol - 1)
    col = np.sort(col.reshape(nrow, zeta), axis=1).reshape(-1)
    sgn = np.where(counter_uniform(seed, e + np.uint64(nrow * zeta)) < 0.5, -1.0, 1.0)
    rowptr = np.arange(nrow + 1, dtype=np.int64) * zeta
    return sp.csr_matrix((sgn, col.astype(np.int32), rowptr), shape=(nrow, ncol))


def sparse_sign_cols(nrow: int, ncol: int, seed: int, zeta: int = 8) -> sp.csr_matrix:
    """Per-column convention: each column gets min(zeta,nrow) DISTINCT rows
    chosen uniformly (Floyd's algorithm on the counter stream) with values
    +-1/sqrt(zeta_eff).  Returned as CSR of the (nrow x ncol) map."""
    z = min(zeta, nrow)
    sel = np.empty((ncol, z), dtype=np.int64)
    cols = np.arange(ncol, dtype=np.uint64)
    for t, j in enumerate(range(nrow - z, nrow)):
        u = counter_uniform(seed, cols * np.uint64(z) + np.uint64(t))
        cand = np.minimum((u * (j + 1)).astype(np.int64), j)
        dup = (sel[:, :t] == cand[:, None]).any(axis=1) if t else np.zeros(ncol, bool)
        sel[:, t] = np.where(dup, j, cand)
    sel.sort(axis=1)
    e = cols[:, None] * np.uint64(z) + np.arange(z, dtype=np.uint64)[None, :]
    u = counter_uniform(seed ^ 0x5bd1e995, e.reshape(-1))
    val = np.where(u < 0.5, -1.0, 1.0) / math.sqrt(z)
    rows = sel.reshape(-1)
    colidx = np.repeat(np.arange(ncol), z)
    return sp.csr_matrix((val, (rows, colidx)), shape=(nrow, ncol))


# --------------------------------------------------------------------------
# linalg shim: src/Skema_BlasLapack.hpp
# --------------------------------------------------------------------------
def qr_inplace(Q: np.ndarray) -> np.ndarray:
    """linalg::qr(Q, nrow, ncol) (src/Skema_BlasLapack.hpp:168-186):
    dgeqrf then dorgqr with lwork = max(1, n); returns thin Q (m x n)."""
    m, n = Q.shape
    lw = max(1, n)
    qr, tau, _, info = lapack.dgeqrf(np.asfortranarray(Q), lwork=lw)
    q, _, info2 = lapack.dorgqr(qr, tau, lwork=lw)
    return np.ascontiguousarray(q)


def qr_with_r(Q: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """linalg::qr(Q, R, nrow, ncol) (src/Skema_BlasLapack.hpp:188-217):
    R = upper triangle of dgeqrf's output (n x n), then dorgqr."""
    m, n = Q.shape
    lw = max(1, n)
    qr, tau, _, info = lapack.dgeqrf(np.asfortranarray(Q), lwork=lw)
    R = np.triu(qr[:n, :n])
    q, _, info2 = lapack.dorgqr(qr, tau, lwork=lw)
    return np.ascontiguousarray(q), R


def ls(A: np.ndarray, B: np.ndarray) -> np.ndarray:
    """linalg::ls(&'N', A, B, m, n, nrhs) (src/Skema_BlasLapack.hpp:219-243):
    dgels 'N'; info is ignored as in the reference."""
    m, n = A.shape
    p = B.shape[1]
    lw = max(1, m * n + max(m * n, p))
    _, x, info = lapack.dgels(np.asfortranarray(A), np.asfortranarray(B), trans="N", lwork=lw)
    return np.ascontiguousarray(x[:n, :])


def ls_trans(A: np.ndarray, B: np.ndarray, trans: str) -> np.ndarray:
    """linalg::ls(&trans, A, B, m, n, nrhs) with either flag (dgels)."""
    m, n = A.shape
    p = B.shape[1]
    lw = max(1, m * n + max(m * n, p))
    _, x, info = lapack.dgels(np.asfortranarray(A), np.asfortranarray(B), trans=trans, lwork=lw)
    rows = n if trans == "N" else m
    return np.ascontiguousarray(x[:rows, :])


def chol_upper(A: np.ndarray) -> np.ndarray:
    """linalg::chol (src/Skema_BlasLapack.hpp:245-272): dpotrf 'U', the strict
    lower triangle zeroed; info > 0 raises (not positive definite)."""
    c, info = lapack.dpotrf(np.asfortranarray(A), lower=0, clean=1)
    if info > 0:
        raise np.linalg.LinAlgError(f"dpotrf: the leading minor of order {info} is not positive definite")
    return np.triu(np.ascontiguousarray(c))


def svd(A: np.ndarray) -> Tuple[np.ndarray, np.ndarray, np.ndarray]:
    """linalg::svd(A, m, n, U, S, V) (src/Skema_BlasLapack.hpp:71-97):
    dgesvd jobu='S', jobvt='S'; the third output holds V^T (ldv = min(m,n))."""
    m, n = A.shape
    mn, mx = min(m, n), max(m, n)
    lw = max(1, 3 * mn + mx, 5 * mn)
    u, s, vt, info = lapack.dgesvd(np.asfortranarray(A), compute_uv=1, full_matrices=0, lwork=lw)
    return np.ascontiguousarray(u), s.copy(), np.ascontiguousarray(vt)


# --------------------------------------------------------------------------
# Impl:: helpers: src/Skema_Common.hpp
# --------------------------------------------------------------------------
def _op(M, t: str):
    return M.T if t == "T" else M


def _dense(M):
    return M.toarray() if sp.issparse(M) else M


def col_subview_csr(data: sp.csr_matrix, idx: Tuple[int, int]) -> sp.csr_matrix:
    """SparseSignDimRedux::col_subview (src/Skema_DimRedux_SparseSign.cpp:125-155):
    keep entries with idx.first <= col < idx.second, re-based columns."""
    d = data.tocsr()
    return sp.csr_matrix(d[:, idx[0]:idx[1]])


# --------------------------------------------------------------------------
# DimRedux maps
# --------------------------------------------------------------------------
class GaussDimRedux:
    """src/Skema_DimRedux.hpp:95-162 with lmap/rmap from
    src/Skema_DimRedux_Gauss.cpp.  ``data`` is the stored (nrow x ncol) map."""

    issparse = False

    def __init__(self, data: np.ndarray, init_transposed: bool = False):
        self.data = np.asarray(data, dtype=np.float64)
        self.nrow, self.ncol = self.data.shape
        self.init_transposed = init_transposed

    def lmap(self, alpha, B, beta, transA="N", transB="N", idx=(0, 0)):
        if sp.issparse(B):
            # Gauss.cpp:54-76: C = op_{transB}(B) * data[:,idx]  (the spmv mode
            # is transB; Impl::mm ignores the second flag), returned transposed.
            data_ = self.data[:, idx[0]:idx[1]] if idx[0] != idx[1] else self.data
            C = alpha * (_op(B.tocsr(), transB) @ data_)
            return np.ascontiguousarray(np.asarray(C).T)
        # Gauss.cpp:10-33
        if self.init_transposed:
            transA = "T" if transA == "N" else "N"
        data_ = self.data[:, idx[0]:idx[1]] if idx[0] != idx[1] else self.data
        return alpha * (_op(data_, transA) @ _op(B, transB))

    def rmap(self, alpha, A, beta, transA="N", transB="T", idx=(0, 0)):
        if sp.issparse(A):
            # Gauss.cpp:78-99: C = op_{transA}(A) * data[idx,:]
            data_ = self.data[idx[0]:idx[1], :] if idx[0] != idx[1] else self.data
            return np.asarray(alpha * (_op(A.tocsr(), transA) @ data_))
        # Gauss.cpp:35-52 (idx ignored)
        return alpha * (_op(A, transA) @ _op(self.data, transB))


class SparseSignDimRedux:
    """src/Skema_DimRedux.hpp:176-308 with lmap/rmap from
    src/Skema_DimRedux_SparseSign.cpp.  ``data`` is the stored CSR map."""

    issparse = True

    def __init__(self, data: sp.csr_matrix, init_transposed: bool = False):
        self.data = sp.csr_matrix(data, dtype=np.float64)
        self.nrow, self.ncol = self.data.shape
        self.init_transposed = init_transposed

    def lmap(self, alpha, B, beta, transA="N", transB="N", idx=(0, 0)):
        data_ = col_subview_csr(self.data, idx) if idx[0] != idx[1] else self.data
        if sp.issparse(B):
            # SparseSign.cpp:53-82: spgemm data_*B (modes ignored), densified.
            return np.asarray((alpha * (data_ @ B.tocsr())).toarray())
        # SparseSign.cpp:9-31: spmv(transA) -- transB is ignored.
        if self.init_transposed:
            transA = "T" if transA == "N" else "N"
        return np.asarray(alpha * (_op(data_, transA) @ B))

    def rmap(self, alpha, A, beta, transA="N", transB="T", idx=(0, 0)):
        if sp.issparse(A):
            # SparseSign.cpp:84-111: spgemm A*data (modes ignored), densified.
            return np.asarray((alpha * (A.tocsr() @ self.data)).toarray())
        # SparseSign.cpp:33-51: C = op_{swap(transB)}(data) * A^T; return C^T.
        transB = "N" if transB == "T" else "T"
        At = np.ascontiguousarray(A.T)
        C = alpha * (_op(self.data, transB) @ At)
        return np.ascontiguousarray(np.asarray(C).T)


# --------------------------------------------------------------------------
# Windows and the RBF kernel generator
# --------------------------------------------------------------------------
def gauss_rbf_compute(X: np.ndarray, Y: np.ndarray, gamma: float, offsets: Tuple[int, int]) -> np.ndarray:
    """GaussRBF<matrix_type>::compute (src/Skema_Kernel.cpp:5-81):
    K[i,j] = exp(-gamma * sum_f (X[i,f]-Y[j,f])^2) by direct differences, then
    K[jj, offsets.first+jj] := 1 for the window's own diagonal (transp=false
    branch :63-72, or the mx==1 special case :68-72)."""
    mx = X.shape[0]
    D = np.empty((mx, Y.shape[0]))
    for i in range(mx):
        diff = X[i][None, :] - Y
        D[i] = np.einsum("ij,ij->i", diff, diff)
    K = np.exp(-1.0 * gamma * D)
    if mx > 1:
        for jj, ii in enumerate(range(offsets[0], offsets[1])):
            K[jj, ii] = 1.0
    else:
        for ii in range(offsets[0], offsets[1]):
            K[0, ii] = 1.0
    return K


class Window:
    """Window<M>::get (src/Skema_Window.hpp:36-78): a row slice [first, second)."""

    def get(self, A, idx):
        if sp.issparse(A):
            return sp.csr_matrix(A.tocsr()[idx[0]:idx[1], :])
        return A[idx[0]:idx[1], :]


class GaussRBFWindow:
    """GaussRBFWindow<M>::get (src/Skema_Window.hpp:90-107): the slice of the
    point set, then the materialised RBF block against all points."""

    def __init__(self, gamma: float):
        self.gamma = gamma

    def get(self, A, idx):
        if sp.issparse(A):
            raise NotImplementedError("Gauss RBF window on a sparse matrix is not available")
        return gauss_rbf_compute(A[idx[0]:idx[1], :], A, self.gamma, idx)


# --------------------------------------------------------------------------
# Parameters and SketchySVD
# --------------------------------------------------------------------------
@dataclass
class AlgParams:
    """The sketch fields of AlgParams (src/Skema_AlgParams.hpp:12-90, defaults
    src/Skema_AlgParams.cpp:10-57; eta/nu default 1.0 via parse :275-278)."""

    matrix_m: int = 0
    matrix_n: int = 0
    rank: int = 1
    sketch_range: int = 1
    sketch_core: int = 1
    window: int = 1
    seeds: Sequence[int] = field(default_factory=lambda: [0, 1, 2, 3])
    sketch_eta: float = 1.0
    sketch_nu: float = 1.0
    dim_redux: str = "gauss"  # "gauss" | "sparse-sign"
    issparse: bool = False
    kernel_func: str = "none"  # "none" | "gaussrbf"
    kernel_gamma: float = 1.0


def sketch_sizes(p: AlgParams) -> Tuple[int, int]:
    """range/core rule of src/Skema_SketchySVD.cpp:20-24."""
    k = 4 * p.rank + 1 if p.sketch_range < p.rank else p.sketch_range
    s = 2 * k + 1 if p.sketch_core < p.rank else p.sketch_core
    return k, s


def map_shapes(p: AlgParams):
    """Stored shape and init_transposed flag of Upsilon, Omega, Phi, Psi
    (src/Skema_SketchySVD.cpp:30-95)."""
    k, s = sketch_sizes(p)
    m, n = p.matrix_m, p.matrix_n
    g = p.dim_redux == "gauss"
    tU = p.issparse and g
    tO = p.issparse
    return {
        "Upsilon": ((m, k) if tU else (k, m), tU),
        "Omega": ((n, k) if tO else (k, n), tO),
        "Phi": ((m, s) if tU else (s, m), tU),
        "Psi": ((n, s) if tO else (s, n), tO),
    }


def make_maps(p: AlgParams, sparse_convention: str = "rows", dist: str = "uniform"):
    """Materialise the four maps from p.seeds with the stored orientation.
    Gauss: gauss_map on the stored shape, U(-a, a) by default like the
    reference's fill (src/Skema_DimRedux.hpp:112-115), N(0,1) with
    dist="normal".  Sparse-sign: 'rows' = reference
    per-stored-row convention; 'cols' = per-column of the LOGICAL map."""
    shapes = map_shapes(p)
    out = {}
    for name, seed in zip(("Upsilon", "Omega", "Phi", "Psi"), p.seeds):
        (r, c), t = shapes[name]
        if p.dim_redux == "gauss":
            out[name] = GaussDimRedux(gauss_map(r, c, seed, dist), t)
        elif sparse_convention == "rows":
            out[name] = SparseSignDimRedux(sparse_sign_rows(r, c, seed), t)
        else:
            # logical map is (c x r) if stored transposed, else (r x c)
            if t:
                out[name] = SparseSignDimRedux(sparse_sign_cols(c, r, seed).T.tocsr(), t)
            else:
                out[name] = SparseSignDimRedux(sparse_sign_cols(r, c, seed), t)
    return out


def logical_map(mp) -> np.ndarray:
    """The logical (k x N) map a stored DimRedux represents."""
    d = _dense(mp.data)
    return np.ascontiguousarray(d.T) if mp.init_transposed else np.asarray(d)


class SketchySVD:
    """SketchySVD<M,D> (src/Skema_SketchySVD.hpp:11-64)."""

    def __init__(self, p: AlgParams, maps=None, sparse_convention: str = "rows", dist: str = "uniform"):
        self.p = p
        self.nrow, self.ncol = p.matrix_m, p.matrix_n
        self.rank = p.rank
        self.range, self.core = sketch_sizes(p)
        self.eta, self.nu = p.sketch_eta, p.sketch_nu
        maps = maps if maps is not None else make_maps(p, sparse_convention, dist)
        self.Upsilon, self.Omega = maps["Upsilon"], maps["Omega"]
        self.Phi, self.Psi = maps["Phi"], maps["Psi"]
        self.window = GaussRBFWindow(p.kernel_gamma) if p.kernel_func == "gaussrbf" else Window()
        self.X = self.Y = self.Z = None
        self.checkpoints = {}

    # src/Skema_SketchySVD.cpp:666-716 -- first argument scales the state.
    @staticmethod
    def axpy(eta, Y, nu, A, idx=(0, 0)):
        if idx[0] == idx[1]:
            Y[...] = eta * Y + nu * A
        else:
            Y[idx[0]:idx[1], :] = eta * Y[idx[0]:idx[1], :] + nu * A

    def update(self, H, idx=(0, 0)):
        """The four update specialisations, src/Skema_SketchySVD.cpp:255-335."""
        one, zero = 1.0, 0.0
        g = isinstance(self.Upsilon, GaussDimRedux)
        if not sp.issparse(H) and g:  # :255-268
            x = self.Upsilon.lmap(one, H, zero, "N", "N", idx)
            y = self.Omega.rmap(one, H, zero, "N", "T", idx)
            w = self.Phi.lmap(one, H, zero, "N", "N", idx)
            z = self.Psi.rmap(one, w, zero, "N", "T")
        elif not sp.issparse(H):  # :270-293
            At = np.ascontiguousarray(H.T)
            yt = self.Omega.lmap(one, At, zero, "N", "N")
            y = np.ascontiguousarray(yt.T)
            x = self.Upsilon.lmap(one, H, zero, "N", "N", idx)
            w = self.Phi.lmap(one, H, zero, "N", "N", idx)
            wt = np.ascontiguousarray(w.T)
            zt = self.Psi.lmap(one, wt, zero, "N", "N")
            z = np.ascontiguousarray(zt.T)
        elif g:  # :295-316
            xt = self.Upsilon.rmap(one, H, zero, "T", "N", idx)
            x = np.ascontiguousarray(xt.T)
            y = self.Omega.rmap(one, H, zero, "N", "N")
            wt = self.Phi.rmap(one, H, zero, "T", "N", idx)
            z = self.Psi.rmap(one, wt, zero, "T", "N")
        else:  # :318-335
            x = self.Upsilon.lmap(one, H, zero, "N", "N", idx)
            y = self.Omega.rmap(one, H, zero, "N", "N")
            w = self.Phi.lmap(one, H, zero, "N", "N", idx)
            z = self.Psi.rmap(one, w, zero, "N", "N")
        return x, y, z

    def init_state(self):
        """Zeroed X, Y, Z (src/Skema_SketchySVD.cpp:131-133)."""
        self.X = np.zeros((self.range, self.ncol))
        self.Y = np.zeros((self.nrow, self.range))
        self.Z = np.zeros((self.core, self.core))

    def update_window(self, H, idx):
        """One pass of the window loop body (:179-189) on a given window H."""
        x, y, z = self.update(H, idx)
        self.axpy(self.nu, self.X, self.eta, x)
        self.axpy(self.nu, self.Z, self.eta, z)
        self.axpy(self.nu, self.Y, self.eta, y, idx)

    def linear_update(self, A, windows: Optional[int] = None):
        """src/Skema_SketchySVD.cpp:119-246.  ``windows`` (oracle-only) stops
        after that many windows -- used to time a bounded CPU sample."""
        wsize = self.p.window
        k, s = self.range, self.core
        self.X = np.zeros((k, self.ncol))
        self.Y = np.zeros((self.nrow, k))
        self.Z = np.zeros((s, s))
        if wsize == self.nrow:  # :137-158
            H = self.window.get(A, (0, self.nrow))
            x, y, z = self.update(H)
            self.axpy(self.nu, self.X, self.eta, x)
            self.axpy(self.nu, self.Z, self.eta, z)
            self.axpy(self.nu, self.Y, self.eta, y)
            return
        irow, done = 0, 0
        while irow < self.nrow:  # :167-233
            if irow + wsize < self.nrow:
                idx = (irow, irow + wsize)
            else:
                idx = (irow, self.nrow)
                wsize = idx[1] - idx[0]
            H = self.window.get(A, idx)
            x, y, z = self.update(H, idx)
            self.axpy(self.nu, self.X, self.eta, x)
            self.axpy(self.nu, self.Z, self.eta, z)
            self.axpy(self.nu, self.Y, self.eta, y, idx)
            irow += wsize
            done += 1
            if windows is not None and done >= windows:
                break

    def initial_approx(self):
        """src/Skema_SketchySVD.cpp:337-566."""
        one, zero = 1.0, 0.0
        P = qr_inplace(np.ascontiguousarray(self.X.T))  # :355-357
        Q = qr_inplace(self.Y.copy())  # :381-383 (the reference overwrites Y)
        U1 = self.Phi.lmap(one, Q, zero, "N", "N")  # :409
        U2 = self.Psi.lmap(one, P, zero, "N", "N")  # :430
        self.checkpoints.update(P=P.copy(), Q=Q.copy(), PhiQ=U1.copy(), PsiP=U2.copy())
        U1, T1 = qr_with_r(U1)  # :453
        U2, T2 = qr_with_r(U2)  # :476
        Z1 = U1.T @ self.Z  # :500
        Z2 = Z1 @ U2  # :515
        Z2 = ls(T1, Z2)  # :532
        Z2t = ls(T2, np.ascontiguousarray(Z2.T))  # :549-551
        C = np.ascontiguousarray(Z2t.T)  # :562
        self.checkpoints.update(T1=T1, T2=T2, C=C.copy())
        return Q, C, P

    def low_rank_approx(self):
        """src/Skema_SketchySVD.cpp:568-664."""
        Q, C, P = self.initial_approx()
        U, S, Vt = svd(C)  # :609 ("V" holds V^T)
        QU = Q @ U  # :627
        PV = P @ Vt.T  # :645 (mm 'N','T')
        r = self.rank
        self.uvecs, self.svals, self.vvecs = QU[:, :r].copy(), S[:r].copy(), PV[:, :r].copy()
        return self.uvecs, self.svals, self.vvecs


# --------------------------------------------------------------------------
# SketchySPD: the Nystrom sketch of a symmetric PSD matrix
# --------------------------------------------------------------------------
class SketchySPD:
    """SketchySPD<M,D> (src/Skema_SketchySVD.cpp:757-1151).

    Y = A Omega (n x k) by row windows, then the fixed-rank Nystrom
    approximation of Tropp, Yurtsever, Udell and Cevher (2017, Alg. 3): a
    shift eps ||Y||_2 Omega, B = Omega^T Y_shift, C = chol((B + B^T) / 2), E = Y
    C^{-1} and the SVD of E.  Two reference defects are not reproduced:
    the window loop never resets its ``pass`` counter (:795-858), so only
    window 0 would be sketched -- here every window is; and the shift removal
    subtracts ||Y||_2 (the local ``nu``, :1141-1145) rather than the shift
    eps ||Y||_2 -- ``shift_removal="reference"`` reproduces it.  The
    sparse-sign ``axpy`` multiplies (``*=``, src/Skema_DimRedux_SparseSign.cpp:
    114-123); here the shift is added for both map families."""

    def __init__(self, p: AlgParams, omega=None, shift_removal: str = "textbook"):
        self.p = p
        self.nrow, self.ncol, self.rank = p.matrix_m, p.matrix_n, p.rank
        self.range = 4 * p.rank + 1 if p.sketch_range < p.rank else p.sketch_range  # :763-765
        self.eta, self.nu = p.sketch_eta, p.sketch_nu
        if omega is None:  # DimReduxT(ncol, range, seeds[0]) (:768-773)
            if p.dim_redux == "gauss":
                omega = GaussDimRedux(gauss_map(self.ncol, self.range, p.seeds[0]))
            else:
                omega = SparseSignDimRedux(sparse_sign_rows(self.ncol, self.range, p.seeds[0]))
        self.Omega = omega
        self.window = GaussRBFWindow(p.kernel_gamma) if p.kernel_func == "gaussrbf" else Window()
        self.shift_removal = shift_removal
        self.Y = None

    def update(self, H):
        """y = H Omega: Gauss rmap (:896-901), sparse-sign lmap of H^T in 'T'
        mode transposed back (:903-912), CRS rmap (:914-931)."""
        one, zero = 1.0, 0.0
        if sp.issparse(H) or not self.Omega.issparse:
            return self.Omega.rmap(one, H, zero, "N", "N")
        At = np.ascontiguousarray(H.T)
        return np.ascontiguousarray(self.Omega.lmap(one, At, zero, "T", "N").T)

    def linear_update(self, A):
        """src/Skema_SketchySVD.cpp:790-880 with axpy (:1153-1200):
        Y[idx] = nu Y[idx] + eta y for every window."""
        m, w = self.nrow, self.p.window
        self.Y = np.zeros((m, self.range))
        if w == m:
            self.Y = self.nu * self.Y + self.eta * self.update(self.window.get(A, (0, m)))
            return
        for r0 in range(0, m, w):
            idx = (r0, min(r0 + w, m))
            y = self.update(self.window.get(A, idx))
            self.Y[idx[0]:idx[1], :] = self.nu * self.Y[idx[0]:idx[1], :] + self.eta * y

    def low_rank_approx(self):
        """src/Skema_SketchySVD.cpp:934-1151 -> (U n x r, lambda r)."""
        mu = np.finfo(np.float64).eps
        Y = self.Y
        nrm = svd(Y.copy())[1][0]  # linalg::nrm2 = largest singular value (BlasLapack.hpp:159-166)
        shift = mu * nrm
        Om = _dense(self.Omega.data)
        Yv = Y + shift * Om  # Omega.axpy (DimRedux_Gauss.cpp:102-130)
        B = np.asarray(Om.T @ Yv)  # Omega.lmap(1, Yv, 0, 'T', 'N')
        C = 0.5 * (B + B.T)  # :1004-1022
        R = chol_upper(C)  # :1048-1058
        E = ls_trans(R, np.ascontiguousarray(Y.T), "T").T  # E = Y R^{-1} (:1064-1082)
        U, S, _ = svd(np.ascontiguousarray(E))  # :1094-1107
        r = self.rank
        lam = S[:r] ** 2 - (nrm if self.shift_removal == "reference" else shift)  # :1141-1145
        self.uvecs, self.svals = U[:, :r].copy(), np.maximum(lam, 0.0)
        return self.uvecs, self.svals


# --------------------------------------------------------------------------
# Residual norms of the final factors (src/Skema_Residuals.hpp)
# --------------------------------------------------------------------------
def _row_windows(m: int, window: Optional[int]):
    w = m if not window else window
    for r0 in range(0, m, w):
        yield r0, min(r0 + w, m)


def map_A_by_window(A, U, V, window: Optional[int] = None, rbf_gamma: Optional[float] = None):
    """(A V, A^T U) window by window (src/Skema_Residuals.hpp:60-121): each
    window A_w = window->get(A, idx) contributes av = A_w V to the rows idx of
    AV and atu = A_w^T U[idx, :] added into AtU.  With rbf_gamma, A is the
    point set and A_w the Gaussian RBF rows of the window
    (GaussRBFWindow::get, src/Skema_Window.hpp:90-107)."""
    if rbf_gamma is not None:
        X = np.asarray(A, dtype=np.float64)
        m = n = X.shape[0]
        win = GaussRBFWindow(rbf_gamma)
    else:
        m, n = A.shape
        win = Window()
    r = V.shape[1]
    AV = np.zeros((m, r))
    AtU = np.zeros((n, r))
    for r0, r1 in _row_windows(m, window):
        Aw = win.get(A if rbf_gamma is None else X, (r0, r1))
        AtU += _dense_or_sparse_T_mul(Aw, U[r0:r1, :])
        AV[r0:r1, :] = _dense(Aw @ V) if sp.issparse(Aw) else Aw @ V
    return AV, AtU


def _dense_or_sparse_T_mul(Aw, B):
    out = Aw.T @ B
    return np.asarray(out.todense()) if sp.issparse(out) else np.asarray(out)


def residual_norms(A, U, S, V, window: Optional[int] = None, rbf_gamma: Optional[float] = None) -> np.ndarray:
    """SketchySVD triplet residuals (src/Skema_Residuals.hpp:160-217):
    rnorms[i] = || [A^T u_i - s_i v_i ; A v_i - s_i u_i] ||_2 (the stacked
    vector of length n + m, ``rvec`` in the reference)."""
    AV, AtU = map_A_by_window(A, U, V, window, rbf_gamma)
    dAV = AV - U * S
    dAtU = AtU - V * S
    return np.sqrt((dAtU ** 2).sum(axis=0) + (dAV ** 2).sum(axis=0))


def residual_norms_sym(A, V, S, window: Optional[int] = None, rbf_gamma: Optional[float] = None) -> np.ndarray:
    """SketchySPD eigen-residuals (src/Skema_Residuals.hpp:123-158):
    rnorms[i] = || A v_i - s_i v_i ||_2 with A V by windows (:17-56)."""
    AV, _ = map_A_by_window(A, V, V, window, rbf_gamma)
    return np.sqrt(((AV - V * S) ** 2).sum(axis=0))


# --------------------------------------------------------------------------
# Comparison helpers used by the tests
# --------------------------------------------------------------------------
def rel_err(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))


def align_signs(U, V, Uref, Vref):
    """Joint per-column sign alignment of (u_i, v_i) onto the reference."""
    sgn = np.sign(np.einsum("ij,ij->j", U, Uref))
    sgn[sgn == 0] = 1.0
    return U * sgn, V * sgn


def rel_fro_error(A, U, S, V) -> float:
    """||A - U diag(S) V^T||_F / ||A||_F (dense A)."""
    A = _dense(A)
    return float(np.linalg.norm(A - (U * S) @ V.T) / np.linalg.norm(A))
