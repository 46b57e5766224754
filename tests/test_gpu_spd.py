"""GPU checks of SketchySPD (src/Skema_SketchySVD.cpp:757-1151) against the
oracle: the sketch Y at 1e-12, the eigenvalues at 1e-12 of lambda_1 and the
eigenvectors after sign alignment, for dense, CSR and RBF input with Gauss
and sparse-sign maps, the reference's shift removal, and exact recovery on
the reference's own symmetric and kernel fixtures."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import skema_oracle as O
from tests.gpu_helpers import TOL, dev, to_gpu_params

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import skema_amd

    skema_amd.load()
    return skema_amd


def _omega(S, osk):
    om = osk.Omega
    if om.issparse:
        d = om.data.tocsr()
        d.sort_indices()
        return S.SparseSignDimRedux(om.nrow, om.ncol, csr=(d.indptr, d.indices, d.data))
    return S.GaussDimRedux(om.nrow, om.ncol, data=om.data)


def _compare(g, osk, A_gpu, shift="eps"):
    Y = g.state()
    assert O.rel_err(Y, osk.Y) <= TOL, O.rel_err(Y, osk.Y)
    U, lam = g.low_rank_approx(shift_removal=shift)
    Ur, lr = osk.low_rank_approx()
    assert np.abs(lam - lr).max() <= 1e-12 * max(lr[0], 1e-300), (lam, lr)
    if shift == "eps":
        sgn = np.sign(np.einsum("ij,ij->j", U, Ur))
        sgn[sgn == 0] = 1.0
        assert O.rel_err(U * sgn, Ur) <= 1e-10, O.rel_err(U * sgn, Ur)


def _psd(n, rank, seed, decay=0.7):
    rng = np.random.default_rng(seed)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = decay ** np.arange(n)
    lam[rank:] *= 1e-3
    return (Q * lam) @ Q.T


@pytest.mark.parametrize("dim_redux", ["gauss", "sparse-sign"])
@pytest.mark.parametrize("window", [64, 300])
def test_spd_dense_against_oracle(S, dim_redux, window):
    print(f"case spd dense {dim_redux} w={window}", flush=True)
    n = 300
    A = _psd(n, 12, 1)
    p = O.AlgParams(matrix_m=n, matrix_n=n, rank=6, sketch_range=25, window=window, sketch_eta=1.3,
                    sketch_nu=0.8, dim_redux=dim_redux)
    osk = O.SketchySPD(p)
    osk.linear_update(A)
    g = S.SketchySPD(to_gpu_params(p), omega=_omega(S, osk))
    g.linear_update(dev(A))
    _compare(g, osk, dev(A))


def test_spd_reference_shift_removal(S):
    print("case spd reference shift", flush=True)
    n = 200
    A = _psd(n, 8, 2, decay=0.9)
    p = O.AlgParams(matrix_m=n, matrix_n=n, rank=5, sketch_range=20, window=50)
    osk = O.SketchySPD(p, shift_removal="reference")
    osk.linear_update(A)
    g = S.SketchySPD(to_gpu_params(p), omega=_omega(S, osk))
    g.linear_update(dev(A))
    U, lam = g.low_rank_approx(shift_removal="reference")
    _, lr = osk.low_rank_approx()
    nrm = np.linalg.norm(osk.Y, 2)
    assert np.abs(lam - lr).max() <= 1e-12 * nrm


def test_spd_csr_against_oracle(S):
    print("case spd csr", flush=True)
    rng = np.random.default_rng(3)
    B = sp.random(400, 400, density=0.01, random_state=3, format="csr")
    A = (B @ B.T + sp.identity(400) * 0.1).tocsr()
    A.sort_indices()
    p = O.AlgParams(matrix_m=400, matrix_n=400, rank=8, sketch_range=30, window=77, issparse=True)
    osk = O.SketchySPD(p)
    osk.linear_update(A)
    g = S.SketchySPD(to_gpu_params(p), omega=_omega(S, osk))
    g.linear_update(S.CrsMatrix.from_scipy(A))
    _compare(g, osk, None)


@pytest.mark.parametrize("n,d,window", [(600, 6, 128), (257, 33, 64)])
def test_spd_rbf_against_oracle(S, n, d, window):
    print(f"case spd rbf n={n} d={d}", flush=True)
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, d))
    p = O.AlgParams(matrix_m=n, matrix_n=n, rank=10, sketch_range=41, window=window, kernel_func="gaussrbf",
                    kernel_gamma=1.0 / d)
    osk = O.SketchySPD(p)
    osk.linear_update(X)
    g = S.SketchySPD(to_gpu_params(p, nfeat=d), omega=_omega(S, osk))
    g.linear_update(dev(X))
    _compare(g, osk, dev(X))


def test_spd_reference_fixtures_exact_recovery(S, golden_dir):
    print("case spd fixtures", flush=True)
    A = np.loadtxt(f"{golden_dir}/symmetric_dense.txt")
    sv = np.loadtxt(f"{golden_dir}/symmetric_svals.txt")
    p = S.AlgParams(matrix_m=27, matrix_n=27, rank=27, sketch_range=27, window=4)
    g = S.SketchySPD(p)
    g.linear_update(dev(A))
    U, lam = g.low_rank_approx()
    assert np.abs(lam - sv).max() <= 1e-12 * sv[0]
    rn = g.compute_residuals(dev(A), dev(U), dev(lam))
    assert rn.max() <= 1e-10 * sv[0], rn
    X = np.loadtxt(f"{golden_dir}/kernel_data.txt")
    ks = np.loadtxt(f"{golden_dir}/kernel_svals.txt")
    n = X.shape[0]
    pk = S.AlgParams(matrix_m=n, matrix_n=n, rank=n, sketch_range=n, window=7, kernel_func="gaussrbf",
                     kernel_gamma=1.0, nfeat=X.shape[1])
    gk = S.SketchySPD(pk)
    gk.linear_update(dev(X))
    _, lk = gk.low_rank_approx()
    assert np.abs(lk - ks[:n]).max() <= 1e-12 * ks[0]


def test_spd_rejects_bad_shapes(S):
    with pytest.raises(S.SkemaError) as e:
        S.SketchySPD(S.AlgParams(matrix_m=50, matrix_n=40, rank=3, sketch_range=8, window=10))
    assert e.value.code in (2, 1)
