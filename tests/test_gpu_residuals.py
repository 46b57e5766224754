"""GPU checks of the residual norms of the final factors against the oracle
(src/Skema_Residuals.hpp:17-217): dense, CSR and RBF inputs, the symmetric
(eigenpair) form, and SketchySVD.compute_residuals end to end.

The residuals are norms of differences of O(sigma_1) terms, so the bound is
absolute: 1e-12 times ||A||_F (FP64 rounding of A V and A^T U is ~1e-16
relative to those terms)."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import skema_oracle as O
from tests.gpu_helpers import dev, gpu_maps, to_gpu_params

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import skema_amd

    skema_amd.load()
    return skema_amd


def _factors(m, n, r, seed, A=None):
    """Approximate factors (truncated SVD of a perturbed A) so that the
    residuals are well away from zero."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n)) if A is None else A
    U, s, Vt = np.linalg.svd(A + 0.3 * rng.standard_normal(A.shape), full_matrices=False)
    return A, U[:, :r].copy(), s[:r].copy(), Vt[:r, :].T.copy()


def _close(rn, ref, A):
    scale = np.linalg.norm(A.toarray() if sp.issparse(A) else A)
    assert np.abs(rn - ref).max() <= 1e-12 * scale, (rn, ref)


@pytest.mark.parametrize("m,n,r", [(300, 200, 7), (1000, 37, 37), (64, 500, 1)])
def test_residuals_dense_against_oracle(S, m, n, r):
    print(f"case residuals dense {m}x{n} r={r}", flush=True)
    A, U, s, V = _factors(m, n, r, m + n)
    rn = S.residual_norms(dev(A), dev(U), dev(s), dev(V))
    _close(rn, O.residual_norms(A, U, s, V), A)


def test_residuals_csr_against_oracle(S):
    print("case residuals csr", flush=True)
    rng = np.random.default_rng(7)
    A = sp.random(800, 300, density=0.02, random_state=7, format="csr")
    A.data = rng.standard_normal(A.nnz)
    _, U, s, V = _factors(800, 300, 9, 8, A=A.toarray())
    rn = S.residual_norms(S.CrsMatrix.from_scipy(A), dev(U), dev(s), dev(V))
    _close(rn, O.residual_norms(A, U, s, V, window=100), A)


@pytest.mark.parametrize("n,d,window", [(500, 5, 128), (777, 33, 100)])
def test_residuals_rbf_against_oracle(S, n, d, window):
    print(f"case residuals rbf n={n} d={d}", flush=True)
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, d))
    gamma = 1.0 / d
    K = O.gauss_rbf_compute(X, X, gamma, (0, n))
    _, U, s, V = _factors(n, n, 6, n + 1, A=K)
    rn = S.residual_norms_rbf(dev(X), gamma, dev(U), dev(s), dev(V), window=window)
    _close(rn, O.residual_norms(X, U, s, V, window=window, rbf_gamma=gamma), K)


def test_residuals_sym_forms(S):
    print("case residuals sym", flush=True)
    rng = np.random.default_rng(11)
    B = rng.standard_normal((300, 300))
    A = B @ B.T / 300
    w, Q = np.linalg.eigh(A + 0.01 * np.diag(rng.standard_normal(300)))
    V, s = Q[:, ::-1][:, :5].copy(), w[::-1][:5].copy()
    rn = S.residual_norms_sym(dev(A), dev(V), dev(s))
    _close(rn, O.residual_norms_sym(A, V, s), A)
    X = rng.standard_normal((400, 4))
    gamma = 0.25
    K = O.gauss_rbf_compute(X, X, gamma, (0, 400))
    w, Q = np.linalg.eigh(K)
    V, s = Q[:, ::-1][:, :6].copy(), w[::-1][:6].copy() * 1.001
    rn = S.residual_norms_sym(dev(X), dev(V), dev(s), window=97, gamma=gamma)
    _close(rn, O.residual_norms_sym(X, V, s, window=97, rbf_gamma=gamma), K)


def test_compute_residuals_end_to_end(S):
    """linear_update -> low_rank_approx -> compute_residuals on the GPU against
    the oracle's own factors and residuals."""
    print("case residuals end to end", flush=True)
    rng = np.random.default_rng(12)
    m, n = 400, 250
    A = rng.standard_normal((m, 20)) @ rng.standard_normal((20, n)) + 0.05 * rng.standard_normal((m, n))
    p = O.AlgParams(matrix_m=m, matrix_n=n, rank=10, window=64)
    osk = O.SketchySVD(p)
    osk.linear_update(A)
    Ur, sr, Vr = osk.low_rank_approx()
    ref = O.residual_norms(A, Ur, sr, Vr)
    maps = gpu_maps({"Upsilon": osk.Upsilon, "Omega": osk.Omega, "Phi": osk.Phi, "Psi": osk.Psi})
    g = S.SketchySVD(to_gpu_params(p), maps=maps)
    Ad = dev(A)
    g.linear_update(Ad)
    rn = g.compute_residuals(Ad)
    _close(rn, ref, A)


def test_residuals_rejects_bad_rank(S):
    A, U, s, V = _factors(50, 40, 3, 1)
    with pytest.raises(S.SkemaError) as e:
        S.residual_norms(dev(A), dev(np.zeros((50, 41))), dev(np.zeros(41)), dev(np.zeros((40, 41))))
    assert e.value.code == 2
