"""Pins the CPU oracle (CPU only, no GPU).

What fixes each oracle function:
  * the reference's own fixtures test/data/*_svals.txt (exact-recovery regime,
    k = min(m, n)) for the whole path, dense and CSR, both map families;
  * closed forms computed by an independent route (logical maps densified
    and multiplied directly) for all four update() forms, the nu/eta
    recurrence and window-size invariance;
  * brute-force index loops on tiny inputs;
  * LAPACK conventions: Q^T Q = I, QR = A, dlarfg's sign rule, dgels on a
    triangular matrix = triangular solve, SVD invariants;
  * scipy.spatial.distance for the RBF window;
  * planted mistakes: mutated copies of oracle functions must fail the checks.
"""
import math

import numpy as np
import pytest
import scipy.linalg as sla
import scipy.sparse as sp
from scipy.spatial.distance import cdist

from oracle import skema_oracle as O


def load(golden_dir, name):
    return np.loadtxt(f"{golden_dir}/{name}")


# ---------------------------------------------------------------- fixtures ---
@pytest.mark.parametrize("model", ["gauss", "sparse-sign"])
@pytest.mark.parametrize("issparse", [False, True])
@pytest.mark.parametrize("window", [1, 5, 7, 32])
def test_rectangular_fixture_exact_recovery(golden_dir, model, issparse, window):
    A = load(golden_dir, "rectangular_dense.txt")
    sv = load(golden_dir, "rectangular_svals.txt")
    p = O.AlgParams(matrix_m=32, matrix_n=14, rank=10, sketch_range=14, sketch_core=29, window=window,
                    dim_redux=model, issparse=issparse)
    sk = O.SketchySVD(p)
    sk.linear_update(sp.csr_matrix(A) if issparse else A)
    U, S, V = sk.low_rank_approx()
    assert np.max(np.abs(S - sv[:10])) <= 1e-13 * sv[0]
    assert O.rel_err((U * S) @ V.T, A) <= 1e-13


def test_sparse_fixture_equals_dense_fixture(golden_dir):
    """rectangular_sparse.mtx (MatrixMarket coordinate) is the dense fixture."""
    from scipy.io import mmread

    A = load(golden_dir, "rectangular_dense.txt")
    M = mmread(f"{golden_dir}/rectangular_sparse.mtx").toarray()
    # the .txt copy carries 15 significant digits, the .mtx 17
    assert np.allclose(M, A, rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("model", ["gauss", "sparse-sign"])
def test_symmetric_fixture_exact_recovery(golden_dir, model):
    A = load(golden_dir, "symmetric_dense.txt")
    sv = load(golden_dir, "symmetric_svals.txt")
    p = O.AlgParams(matrix_m=27, matrix_n=27, rank=27, sketch_range=27, sketch_core=55, window=5, dim_redux=model)
    sk = O.SketchySVD(p)
    sk.linear_update(A)
    U, S, V = sk.low_rank_approx()
    assert np.max(np.abs(S - sv)) <= 1e-13 * sv[0]


def test_symmetric_bin_fixture(golden_dir):
    """symmetric_sparse.bin: Kokkos-Kernels binary CRS (int32 nrows, uint64
    nnz, uint64 rowptr[n+1], int32 col[nnz], f64 val[nnz])."""
    raw = open(f"{golden_dir}/symmetric_sparse.bin", "rb").read()
    n = int(np.frombuffer(raw, np.int32, 1, 0)[0])
    nnz = int(np.frombuffer(raw, np.uint64, 1, 4)[0])
    rp = np.frombuffer(raw, np.uint64, n + 1, 12).astype(np.int64)
    off = 12 + 8 * (n + 1)
    col = np.frombuffer(raw, np.int32, nnz, off)
    val = np.frombuffer(raw, np.float64, nnz, off + 4 * nnz)
    M = sp.csr_matrix((val, col, rp), shape=(n, n)).toarray()
    assert np.allclose(M, load(golden_dir, "symmetric_dense.txt"), rtol=1e-14, atol=1e-9)


@pytest.mark.parametrize("window", [4, 25])
def test_kernel_fixture_exact_recovery(golden_dir, window):
    """kernel_data.txt with gamma = 1 (the --gamma default) -> kernel_svals.txt."""
    X = load(golden_dir, "kernel_data.txt")
    sv = load(golden_dir, "kernel_svals.txt")
    p = O.AlgParams(matrix_m=25, matrix_n=25, rank=10, sketch_range=25, sketch_core=51, window=window,
                    kernel_func="gaussrbf", kernel_gamma=1.0)
    sk = O.SketchySVD(p)
    sk.linear_update(X)
    U, S, V = sk.low_rank_approx()
    assert np.max(np.abs(S - sv[:10])) <= 1e-13 * sv[0]


# ------------------------------------------------------------ closed forms ---
def closed_form_sketch(A, maps, nu=1.0, eta=1.0, window=None):
    """X, Y, Z from densified logical maps, windowed recurrence written out."""
    Ad = A.toarray() if sp.issparse(A) else A
    m = Ad.shape[0]
    Ups, Om = O.logical_map(maps["Upsilon"]), O.logical_map(maps["Omega"])
    Ph, Ps = O.logical_map(maps["Phi"]), O.logical_map(maps["Psi"])
    w = window or m
    X = np.zeros((Ups.shape[0], Ad.shape[1]))
    Y = np.zeros((m, Om.shape[0]))
    Z = np.zeros((Ph.shape[0], Ph.shape[0]))
    for r0 in range(0, m, w):
        r1 = min(m, r0 + w)
        Aw = Ad[r0:r1]
        X = nu * X + eta * (Ups[:, r0:r1] @ Aw)
        Y[r0:r1] = nu * Y[r0:r1] + eta * (Aw @ Om.T)
        Z = nu * Z + eta * (Ph[:, r0:r1] @ (Aw @ Ps.T))
    return X, Y, Z


FORMS = [("gauss", False), ("sparse-sign", False), ("gauss", True), ("sparse-sign", True)]


@pytest.mark.parametrize("model,issparse", FORMS)
@pytest.mark.parametrize("conv", ["rows", "cols"])
def test_update_forms_match_closed_form(model, issparse, conv):
    rng = np.random.default_rng(0)
    m, n = 61, 23
    A = rng.standard_normal((m, n))
    if issparse:
        A[rng.random((m, n)) < 0.7] = 0.0
        Ain = sp.csr_matrix(A)
    else:
        Ain = A
    p = O.AlgParams(matrix_m=m, matrix_n=n, rank=3, window=9, dim_redux=model, issparse=issparse,
                    sketch_nu=0.7, sketch_eta=1.9)
    sk = O.SketchySVD(p, sparse_convention=conv)
    sk.linear_update(Ain)
    X, Y, Z = closed_form_sketch(A, {"Upsilon": sk.Upsilon, "Omega": sk.Omega, "Phi": sk.Phi, "Psi": sk.Psi},
                                 nu=0.7, eta=1.9, window=9)
    assert O.rel_err(sk.X, X) <= 1e-13
    assert O.rel_err(sk.Y, Y) <= 1e-13
    assert O.rel_err(sk.Z, Z) <= 1e-13


def test_map_orientation_follows_reference():
    """Stored shapes / init_transposed of src/Skema_SketchySVD.cpp:30-95."""
    p = O.AlgParams(matrix_m=50, matrix_n=20, rank=2)
    k, s = O.sketch_sizes(p)
    assert (k, s) == (9, 19)
    assert O.map_shapes(p)["Upsilon"] == ((9, 50), False)
    p.issparse, p.dim_redux = True, "gauss"
    sh = O.map_shapes(p)
    assert sh["Upsilon"] == ((50, 9), True) and sh["Omega"] == ((20, 9), True)
    assert sh["Phi"] == ((50, 19), True) and sh["Psi"] == ((20, 19), True)
    p.dim_redux = "sparse-sign"
    sh = O.map_shapes(p)
    assert sh["Upsilon"] == ((9, 50), False) and sh["Omega"] == ((20, 9), True)
    assert sh["Phi"] == ((19, 50), False) and sh["Psi"] == ((20, 19), True)
    p2 = O.AlgParams(rank=3, sketch_range=7, sketch_core=5)
    assert O.sketch_sizes(p2) == (7, 5)


@pytest.mark.parametrize("model", ["gauss", "sparse-sign"])
def test_window_size_invariance(model):
    rng = np.random.default_rng(1)
    A = rng.standard_normal((40, 17))
    ref = None
    for w in [1, 3, 16, 40, 100]:
        p = O.AlgParams(matrix_m=40, matrix_n=17, rank=2, window=w, dim_redux=model)
        sk = O.SketchySVD(p)
        sk.linear_update(A)
        if ref is None:
            ref = (sk.X, sk.Y, sk.Z)
        else:
            for a, b in zip((sk.X, sk.Y, sk.Z), ref):
                assert O.rel_err(a, b) <= 1e-14


def test_brute_force_tiny():
    """Index loops on a 6 x 5 matrix, two windows, nu/eta != 1."""
    rng = np.random.default_rng(2)
    m, n = 6, 5
    A = rng.standard_normal((m, n))
    p = O.AlgParams(matrix_m=m, matrix_n=n, rank=1, sketch_range=2, sketch_core=3, window=4,
                    sketch_nu=0.5, sketch_eta=2.0)
    sk = O.SketchySVD(p)
    sk.linear_update(A)
    Ups, Om, Ph, Ps = (sk.Upsilon.data, sk.Omega.data, sk.Phi.data, sk.Psi.data)
    X = [[0.0] * n for _ in range(2)]
    Y = [[0.0] * 2 for _ in range(m)]
    Z = [[0.0] * 3 for _ in range(3)]
    for r0, r1 in [(0, 4), (4, 6)]:
        for pp in range(2):
            for j in range(n):
                X[pp][j] = 0.5 * X[pp][j] + 2.0 * sum(Ups[pp][i] * A[i][j] for i in range(r0, r1))
        for i in range(r0, r1):
            for pp in range(2):
                Y[i][pp] = 0.5 * Y[i][pp] + 2.0 * sum(A[i][j] * Om[pp][j] for j in range(n))
        for a in range(3):
            for b in range(3):
                zab = 0.0
                for i in range(r0, r1):
                    for j in range(n):
                        zab += Ph[a][i] * A[i][j] * Ps[b][j]
                Z[a][b] = 0.5 * Z[a][b] + 2.0 * zab
    assert O.rel_err(sk.X, np.array(X)) <= 1e-14
    assert O.rel_err(sk.Y, np.array(Y)) <= 1e-14
    assert O.rel_err(sk.Z, np.array(Z)) <= 1e-14


# ------------------------------------------------------- LAPACK conventions ---
def test_qr_conventions():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((40, 7))
    Q, R = O.qr_with_r(A.copy())
    assert np.allclose(Q.T @ Q, np.eye(7), atol=1e-14)
    assert O.rel_err(Q @ R, A) <= 1e-14
    assert np.allclose(np.tril(R, -1), 0.0)
    # dlarfg: beta = -sign(alpha) * ||x||  ->  R[0,0] = -sign(A[0,0]) ||A[:,0]||
    assert R[0, 0] == pytest.approx(-math.copysign(np.linalg.norm(A[:, 0]), A[0, 0]), rel=1e-14)
    assert O.rel_err(O.qr_inplace(A.copy()), Q) <= 1e-15
    # rank-deficient input still gives an orthonormal Q
    B = np.c_[A[:, :3], A[:, :3] @ rng.standard_normal((3, 4))]
    Qb = O.qr_inplace(B)
    assert np.allclose(Qb.T @ Qb, np.eye(7), atol=1e-13)


def test_ls_on_triangular_is_back_substitution():
    rng = np.random.default_rng(4)
    T = np.triu(rng.standard_normal((9, 9))) + 5 * np.eye(9)
    B = rng.standard_normal((9, 9))
    assert O.rel_err(O.ls(T, B), sla.solve_triangular(T, B, lower=False)) <= 1e-14


def test_svd_invariants():
    rng = np.random.default_rng(5)
    A = rng.standard_normal((12, 12))
    U, S, Vt = O.svd(A)
    assert np.all(np.diff(S) <= 0)
    assert O.rel_err((U * S) @ Vt, A) <= 1e-14
    assert np.allclose(U.T @ U, np.eye(12), atol=1e-14)
    assert np.allclose(S, np.linalg.svd(A, compute_uv=False), rtol=1e-14)


# -------------------------------------------------------------------- RBF ---
def test_rbf_window_closed_form():
    rng = np.random.default_rng(6)
    X = rng.standard_normal((30, 4))
    gamma = 0.37
    K = O.gauss_rbf_compute(X[10:17], X, gamma, (10, 17))
    ref = np.exp(-gamma * cdist(X[10:17], X, "sqeuclidean"))
    assert np.max(np.abs(K - ref)) <= 1e-15
    for jj in range(7):
        assert K[jj, 10 + jj] == 1.0
    full = O.gauss_rbf_compute(X, X, gamma, (0, 30))
    assert np.array_equal(full, full.T) or np.max(np.abs(full - full.T)) <= 1e-16
    assert np.all(O.gauss_rbf_compute(X[:3], X, 0.0, (0, 3)) == 1.0)


# ------------------------------------------------------------- generators ---
def test_gauss_map_moments_and_uniform_bounds():
    G = O.gauss_map(100, 2000, 0)
    assert abs(G.mean()) < 0.01 and abs(G.var() - 1.0) < 0.01
    assert np.array_equal(G, O.gauss_map(100, 2000, 0))
    assert not np.array_equal(G, O.gauss_map(100, 2000, 1))
    U = O.gauss_map(50, 40, 3, "uniform")
    a = math.sqrt(2 * math.log(50 * 40))
    assert np.all(np.abs(U) < a) and U.max() > 0.9 * a and U.min() < -0.9 * a


def test_sparse_sign_conventions():
    R = O.sparse_sign_rows(40, 300, 1)
    assert np.all(np.diff(R.indptr) == 8)
    assert set(np.unique(R.data)) <= {-1.0, 1.0}
    for i in range(40):
        row = R.indices[R.indptr[i]:R.indptr[i + 1]]
        assert np.all(np.diff(row) >= 0)
    Cc = O.sparse_sign_cols(101, 500, 2).tocsc()
    assert np.all(np.diff(Cc.indptr) == 8)
    for j in range(500):
        rows = Cc.indices[Cc.indptr[j]:Cc.indptr[j + 1]]
        assert len(set(rows)) == 8
    assert np.allclose(np.abs(Cc.data), 1 / math.sqrt(8))
    assert abs(np.mean(Cc.data > 0) - 0.5) < 0.05
    small = O.sparse_sign_rows(5, 3, 0)  # zeta = max(2, min(ncol, 8)) = 3
    assert np.all(np.diff(small.indptr) == 3)


# ------------------------------------------------------- planted mistakes ---
def _mut_ignore_transposed_flag(monkeypatch):
    orig = O.GaussDimRedux.lmap

    def lmap(self, alpha, B, beta, transA="N", transB="N", idx=(0, 0)):
        if sp.issparse(B):
            return orig(self, alpha, B, beta, transA, transB, idx)
        return alpha * (O._op(self.data, transA) @ O._op(B, transB))  # flip dropped

    monkeypatch.setattr(O.GaussDimRedux, "lmap", lmap)


def _mut_axpy_swapped(monkeypatch):
    def axpy(eta, Y, nu, A, idx=(0, 0)):  # scales the new term instead of the state
        if idx[0] == idx[1]:
            Y[...] = nu * Y + eta * A
        else:
            Y[idx[0]:idx[1], :] = nu * Y[idx[0]:idx[1], :] + eta * A

    monkeypatch.setattr(O.SketchySVD, "axpy", staticmethod(axpy))


def _mut_window_off_by_one(monkeypatch):
    orig = O.Window.get

    def get(self, A, idx):
        return orig(self, A, (idx[0], idx[1]) if idx[0] == 0 else (idx[0] - 1, idx[1] - 1))

    monkeypatch.setattr(O.Window, "get", get)


def _mut_rbf_sign(monkeypatch):
    def compute(X, Y, gamma, offsets):
        return np.exp(+gamma * cdist(X, Y, "sqeuclidean") * 0.0 - gamma * cdist(X, Y, "euclidean"))

    monkeypatch.setattr(O, "gauss_rbf_compute", compute)


def _mut_qr_no_sign_rule(monkeypatch):
    def qr_with_r(Q):
        q, r = np.linalg.qr(Q)
        d = np.sign(np.diag(r))
        return q * d, (r.T * d).T  # positive-diagonal convention

    monkeypatch.setattr(O, "qr_with_r", qr_with_r)


@pytest.mark.parametrize("mutant,check", [
    (_mut_ignore_transposed_flag, lambda gd: test_rectangular_fixture_exact_recovery(gd, "gauss", True, 5)),
    (_mut_axpy_swapped, lambda gd: test_update_forms_match_closed_form("gauss", False, "rows")),
    (_mut_window_off_by_one, lambda gd: test_update_forms_match_closed_form("gauss", False, "rows")),
    (_mut_rbf_sign, lambda gd: test_kernel_fixture_exact_recovery(gd, 4)),
    (_mut_qr_no_sign_rule, lambda gd: test_qr_conventions()),
])
def test_planted_mistakes_are_caught(monkeypatch, golden_dir, mutant, check):
    mutant(monkeypatch)
    with pytest.raises((AssertionError, ValueError)):
        check(golden_dir)


# ---------------------------------------------------------- golden vectors ---
def test_oracle_reproduces_committed_golden(golden_dir):
    """toy_sketch.npz (written by tests/golden/make_golden.py) guards the
    oracle against drift: maps, state and the factors must reproduce."""
    g = np.load(f"{golden_dir}/toy_sketch.npz")
    m, n, r, w, nu, eta = g["params"]
    p = O.AlgParams(matrix_m=int(m), matrix_n=int(n), rank=int(r), window=int(w), sketch_nu=nu, sketch_eta=eta)
    sk = O.SketchySVD(p, dist="normal")
    for name in ("Upsilon", "Omega", "Phi", "Psi"):
        assert np.array_equal(getattr(sk, name).data, g[name])
    sk.linear_update(g["A"])
    for name in ("X", "Y", "Z"):
        assert O.rel_err(getattr(sk, name), g[name]) <= 1e-14
    U, S, V = sk.low_rank_approx()
    assert O.rel_err(S, g["S"]) <= 1e-14
    Ua, Va = O.align_signs(U, V, g["U"], g["V"])
    assert O.rel_err(Ua, g["U"]) <= 1e-13 and O.rel_err(Va, g["V"]) <= 1e-13


# --------------------------------------------------------------------------
# Residual norms (src/Skema_Residuals.hpp:17-217), pinned by what the
# singular value decomposition fixes: exact triplets have zero residual, a
# shifted singular value by delta gives sqrt(2)|delta|, a flipped u_i gives
# 2 sqrt(2) s_i, plus a loop restatement, window invariance and the RBF path
# equal to the dense path on the materialised kernel.
# --------------------------------------------------------------------------
def _exact_triplets(m, n, r, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n))
    U, S, Vt = np.linalg.svd(A, full_matrices=False)
    return A, U[:, :r].copy(), S[:r].copy(), Vt[:r, :].T.copy()


@pytest.mark.parametrize("window", [None, 7, 64])
def test_residuals_zero_for_exact_triplets(window):
    A, U, S, V = _exact_triplets(60, 40, 6, 1)
    rn = O.residual_norms(A, U, S, V, window=window)
    assert rn.shape == (6,)
    assert rn.max() <= 1e-12 * S[0]


def test_residuals_closed_forms_of_planted_errors():
    A, U, S, V = _exact_triplets(50, 30, 5, 2)
    delta = np.array([1e-3, -2e-2, 0.5, 0.0, 3.0])
    rn = O.residual_norms(A, U, S + delta, V)
    assert np.allclose(rn, math.sqrt(2.0) * np.abs(delta), rtol=1e-10, atol=1e-12 * S[0])
    Uf = U.copy()
    Uf[:, 2] *= -1.0  # a flipped left vector: A v - s u' = 2 s u, A^T u' - s v = -2 s v
    rn = O.residual_norms(A, Uf, S, V)
    assert abs(rn[2] - 2.0 * math.sqrt(2.0) * S[2]) <= 1e-10 * S[2]
    assert np.delete(rn, 2).max() <= 1e-12 * S[0]


def test_residuals_loop_restatement_and_sparse():
    rng = np.random.default_rng(3)
    m, n, r = 9, 7, 3
    A = rng.standard_normal((m, n))
    U = rng.standard_normal((m, r))
    V = rng.standard_normal((n, r))
    S = rng.standard_normal(r)
    ref = np.zeros(r)
    for c in range(r):
        acc = 0.0
        for j in range(n):
            t = sum(A[i, j] * U[i, c] for i in range(m)) - S[c] * V[j, c]
            acc += t * t
        for i in range(m):
            t = sum(A[i, j] * V[j, c] for j in range(n)) - S[c] * U[i, c]
            acc += t * t
        ref[c] = math.sqrt(acc)
    assert np.allclose(O.residual_norms(A, U, S, V, window=4), ref, rtol=1e-13)
    As = sp.csr_matrix(np.where(rng.random((m, n)) < 0.4, A, 0.0))
    assert np.allclose(O.residual_norms(As, U, S, V, window=2),
                       O.residual_norms(As.toarray(), U, S, V), rtol=1e-13)


def test_residuals_rbf_path_equals_materialised_kernel():
    rng = np.random.default_rng(4)
    X = rng.standard_normal((40, 3))
    gamma = 0.3
    K = O.gauss_rbf_compute(X, X, gamma, (0, 40))
    U = rng.standard_normal((40, 4))
    V = rng.standard_normal((40, 4))
    S = rng.random(4)
    rn = O.residual_norms(X, U, S, V, window=9, rbf_gamma=gamma)
    assert np.allclose(rn, O.residual_norms(K, U, S, V), rtol=1e-12)


def test_residuals_sym_eigenpairs():
    rng = np.random.default_rng(5)
    B = rng.standard_normal((30, 30))
    A = B @ B.T
    w, Q = np.linalg.eigh(A)
    V, S = Q[:, ::-1][:, :4].copy(), w[::-1][:4].copy()
    assert O.residual_norms_sym(A, V, S, window=8).max() <= 1e-11 * S[0]
    d = np.array([0.0, 1e-2, -3.0, 0.25])
    assert np.allclose(O.residual_norms_sym(A, V, S + d), np.abs(d), rtol=1e-9, atol=1e-11 * S[0])


# --------------------------------------------------------------------------
# SketchySPD (src/Skema_SketchySVD.cpp:757-1151), pinned by the reference's
# own symmetric and kernel fixtures (exact recovery at full range), the
# closed form Y = eta A Omega, exact recovery of a rank-5 PSD matrix, and the
# reference's shift removal written out directly.
# --------------------------------------------------------------------------
def test_spd_symmetric_fixture_exact_recovery(golden_dir):
    A = np.loadtxt(f"{golden_dir}/symmetric_dense.txt")
    sv = np.loadtxt(f"{golden_dir}/symmetric_svals.txt")
    p = O.AlgParams(matrix_m=27, matrix_n=27, rank=27, sketch_range=27, window=4)
    s = O.SketchySPD(p)
    s.linear_update(A)
    U, lam = s.low_rank_approx()
    assert np.abs(lam - sv).max() <= 1e-12 * sv[0]
    assert O.residual_norms_sym(A, U, lam).max() <= 1e-12 * sv[0]


def test_spd_kernel_fixture_exact_recovery(golden_dir):
    X = np.loadtxt(f"{golden_dir}/kernel_data.txt")
    ks = np.loadtxt(f"{golden_dir}/kernel_svals.txt")
    n = X.shape[0]
    p = O.AlgParams(matrix_m=n, matrix_n=n, rank=n, sketch_range=n, window=7, kernel_func="gaussrbf",
                    kernel_gamma=1.0)
    s = O.SketchySPD(p)
    s.linear_update(X)
    _, lam = s.low_rank_approx()
    assert np.abs(lam - ks[:n]).max() <= 1e-12 * ks[0]


@pytest.mark.parametrize("window", [1, 7, 40])
def test_spd_sketch_closed_form(window):
    rng = np.random.default_rng(21)
    B = rng.standard_normal((40, 40))
    A = B @ B.T
    p = O.AlgParams(matrix_m=40, matrix_n=40, rank=3, window=window, sketch_eta=1.7, sketch_nu=0.4)
    s = O.SketchySPD(p)
    s.linear_update(A)
    # every row sits in one window and Y starts at zero: Y = eta A Omega
    assert np.allclose(s.Y, 1.7 * A @ s.Omega.data, rtol=1e-13, atol=1e-12)
    assert s.range == 13  # range < rank selects 4r + 1


def test_spd_low_rank_exact_and_reference_shift():
    rng = np.random.default_rng(22)
    Q, _ = np.linalg.qr(rng.standard_normal((60, 5)))
    lam_true = np.array([50.0, 20.0, 7.0, 3.0, 1.0])
    A = (Q * lam_true) @ Q.T
    p = O.AlgParams(matrix_m=60, matrix_n=60, rank=5, sketch_range=12, window=16)
    s = O.SketchySPD(p)
    s.linear_update(A)
    U, lam = s.low_rank_approx()
    assert np.allclose(lam, lam_true, rtol=1e-10)
    assert np.allclose(U @ U.T, Q @ Q.T, atol=1e-10)
    # the reference's removal subtracts ||Y||_2 from S^2 (its local `nu`)
    s2 = O.SketchySPD(p, shift_removal="reference")
    s2.linear_update(A)
    _, lam2 = s2.low_rank_approx()
    nrm = np.linalg.norm(s2.Y, 2)
    assert np.allclose(lam2, np.maximum(lam_true + np.finfo(float).eps * nrm - nrm, 0.0), atol=1e-8 * nrm)


def test_spd_sparse_sign_map_equals_dense_map():
    rng = np.random.default_rng(23)
    B = rng.standard_normal((30, 30))
    A = B @ B.T
    p = O.AlgParams(matrix_m=30, matrix_n=30, rank=4, sketch_range=9, window=8, dim_redux="sparse-sign")
    s = O.SketchySPD(p)
    s.linear_update(A)
    d = O.SketchySPD(p, omega=O.GaussDimRedux(s.Omega.data.toarray()))
    d.linear_update(A)
    assert np.allclose(s.Y, d.Y, rtol=1e-13, atol=1e-12)
    assert np.allclose(s.low_rank_approx()[1], d.low_rank_approx()[1], rtol=1e-10)
