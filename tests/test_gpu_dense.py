"""GPU parity of the plain FP64 path against the oracle (dense windows).

Each case prints its name before touching the GPU so a failing run names it.
"""
import numpy as np
import pytest

from oracle import skema_oracle as O
from tests.gpu_helpers import TOL, compare_usv, dev, gpu_maps, host, to_gpu_params

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import skema_amd

    skema_amd.load()
    return skema_amd


def _rand(m, n, seed):
    return np.random.default_rng(seed).standard_normal((m, n))


@pytest.mark.parametrize("tA", ["N", "T"])
@pytest.mark.parametrize("tB", ["N", "T"])
@pytest.mark.parametrize("transposed", [False, True])
def test_gauss_lmap_rmap(S, tA, tB, transposed):
    print(f"case gauss lmap/rmap tA={tA} tB={tB} transposed={transposed}", flush=True)
    rng = np.random.default_rng(1)
    nrow, ncol = 37, 53
    D = rng.standard_normal((nrow, ncol))
    om = O.GaussDimRedux(D, transposed)
    gm = S.GaussDimRedux(nrow, ncol, init_transposed=transposed, data=D)
    eff = tA if not transposed else ("T" if tA == "N" else "N")
    for idx in [(0, 0), (5, 29)]:
        wc = idx[1] - idx[0] if idx[0] != idx[1] else ncol
        kd = wc if eff == "N" else nrow
        nb = 19
        B = rng.standard_normal((kd, nb) if tB == "N" else (nb, kd))
        ref = om.lmap(1.5, B, 0.0, tA, tB, idx)
        got = host(gm.lmap(1.5, dev(B), 0.0, tA, tB, idx))
        assert O.rel_err(got, ref) <= TOL
    # rmap: op(A) * op(map)
    kd = ncol if tB == "T" else nrow
    A = rng.standard_normal((23, kd) if tA == "N" else (kd, 23))
    ref = om.rmap(0.5, A, 0.0, tA, tB)
    got = host(gm.rmap(0.5, dev(A), 0.0, tA, tB))
    assert O.rel_err(got, ref) <= TOL


def test_gemm_beta_and_large_k(S):
    """beta accumulation and the split-K path (small output, long K)."""
    print("case gemm beta/split-K", flush=True)
    rng = np.random.default_rng(2)
    D = rng.standard_normal((83, 20011))
    gm = S.GaussDimRedux(83, 20011, data=D)
    B = rng.standard_normal((20011, 41))
    C0 = rng.standard_normal((83, 41))
    Cd = dev(C0)
    gm.lmap(2.0, dev(B), 0.25, "N", "N", (0, 0), C_out=Cd)
    assert O.rel_err(host(Cd), 2.0 * D @ B + 0.25 * C0) <= TOL


def test_gauss_seeded_matches_oracle(S):
    print("case gauss seeded generator", flush=True)
    for dist in ["normal", "uniform"]:
        gm = S.GaussDimRedux(41, 997, seed=7, dist=dist)
        ref = O.gauss_map(41, 997, 7, dist)
        assert np.max(np.abs(gm.download() - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("conv", ["rows", "cols"])
@pytest.mark.parametrize("transposed", [False, True])
def test_sparse_sign_seeded_matches_oracle(S, conv, transposed):
    print(f"case sparse-sign seeded {conv} transposed={transposed}", flush=True)
    nrow, ncol = (300, 23) if transposed else (23, 300)
    sm = S.SparseSignDimRedux(nrow, ncol, seed=11, init_transposed=transposed, convention=conv)
    rp, ci, v = sm.download()
    if conv == "rows":
        ref = O.sparse_sign_rows(nrow, ncol, 11)
    else:
        ref = O.sparse_sign_cols(ncol, nrow, 11).T.tocsr() if transposed else O.sparse_sign_cols(nrow, ncol, 11)
    ref.sort_indices()
    assert np.array_equal(rp, ref.indptr)
    assert np.array_equal(ci, ref.indices)
    assert np.array_equal(v, ref.data)


@pytest.mark.parametrize("model", ["gauss", "sparse-sign"])
@pytest.mark.parametrize("window,nu,eta", [(37, 1.0, 1.0), (300, 0.9, 1.3), (1, 1.0, 1.0), (64, 0.0, 2.0)])
def test_dense_sketch_parity(S, model, window, nu, eta):
    print(f"case dense sketch {model} w={window} nu={nu} eta={eta}", flush=True)
    m, n = 300, 200
    if window == 1:
        m, n = 40, 30
    A = _rand(m, n, 3) @ np.diag(0.9 ** np.arange(n)) + 1e-3 * _rand(m, n, 4)
    p = O.AlgParams(matrix_m=m, matrix_n=n, rank=5, window=window, dim_redux=model, sketch_eta=eta, sketch_nu=nu)
    if window == 1:
        p.sketch_range, p.sketch_core = 12, 25
        p.rank = 4
    osk = O.SketchySVD(p, sparse_convention="rows")
    osk.linear_update(A)
    Xr, Yr, Zr = osk.X.copy(), osk.Y.copy(), osk.Z.copy()
    Ur, Sr, Vr = osk.low_rank_approx()

    gsk = S.SketchySVD(to_gpu_params(p), maps=gpu_maps({"Upsilon": osk.Upsilon, "Omega": osk.Omega,
                                                       "Phi": osk.Phi, "Psi": osk.Psi}))
    gsk.linear_update(dev(A))
    X, Y, Z = gsk.state()
    assert O.rel_err(X, Xr) <= TOL
    assert O.rel_err(Y, Yr) <= TOL
    assert O.rel_err(Z, Zr) <= TOL
    U, Sv, V = gsk.low_rank_approx()
    compare_usv(U, Sv, V, Ur, Sr, Vr)
    # stage checkpoints (the QR sign rule is LAPACK's, so no alignment needed)
    for name in ["P", "Q", "PhiQ", "PsiP", "T1", "T2", "C"]:
        assert O.rel_err(gsk.checkpoint(name), osk.checkpoints[name]) <= 1e-11, name


def test_per_window_update_equals_linear_update(S):
    print("case per-window update", flush=True)
    m, n = 257, 129
    A = _rand(m, n, 5)
    p = O.AlgParams(matrix_m=m, matrix_n=n, rank=4, window=50)
    osk = O.SketchySVD(p)
    maps = gpu_maps({"Upsilon": osk.Upsilon, "Omega": osk.Omega, "Phi": osk.Phi, "Psi": osk.Psi})
    g1 = S.SketchySVD(to_gpu_params(p), maps=maps)
    Ad = dev(A)
    g1.linear_update(Ad)
    g2 = S.SketchySVD(to_gpu_params(p), maps=maps)
    for r0 in range(0, m, 50):
        g2.update(Ad, (r0, min(m, r0 + 50)))
    for a, b in zip(g1.state(), g2.state()):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("name,m,n,r,k,s", [("rectangular", 32, 14, 10, 14, 29), ("symmetric", 27, 27, 27, 27, 55)])
@pytest.mark.parametrize("window", [5, 7])
def test_fixture_exact_recovery(S, golden_dir, name, m, n, r, k, s, window):
    """The reference's fixtures (test/data/*_dense.txt, *_svals.txt): with
    k = min(m, n) the three-sketch reconstruction is exact."""
    print(f"case fixture {name} w={window}", flush=True)
    A = np.loadtxt(f"{golden_dir}/{name}_dense.txt")
    sv = np.loadtxt(f"{golden_dir}/{name}_svals.txt")
    p = O.AlgParams(matrix_m=m, matrix_n=n, rank=r, sketch_range=k, sketch_core=s, window=window)
    g = S.SketchySVD(to_gpu_params(p))  # seeded device maps
    g.linear_update(dev(A))
    U, Sv, V = g.low_rank_approx()
    assert np.max(np.abs(Sv - sv[:r])) <= 1e-13 * sv[0]
    assert O.rel_err((U * Sv) @ V.T, A) <= 1e-13


def test_residual_pass(S):
    print("case residual pass", flush=True)
    m, n, r = 301, 77, 6
    rng = np.random.default_rng(9)
    A, U, V = rng.standard_normal((m, n)), rng.standard_normal((m, r)), rng.standard_normal((n, r))
    sv = rng.random(r)
    e2, a2 = S.residual_fro(dev(A), dev(U), dev(sv), dev(V), window=64)
    assert abs(e2 - np.linalg.norm(A - (U * sv) @ V.T) ** 2) <= 1e-12 * e2
    assert abs(a2 - np.linalg.norm(A) ** 2) <= 1e-12 * a2


def test_shape_errors_are_reported(S):
    """Host-side validation: no kernel is launched on bad shapes."""
    print("case error paths", flush=True)
    p = S.AlgParams(matrix_m=50, matrix_n=40, rank=3, window=10)
    g = S.SketchySVD(p)
    small = dev(np.zeros((10, 40)))
    with pytest.raises(S.SkemaError) as e:
        g.linear_update(small)  # 50 x 40 expected
    assert "SKEMA_ERR_SHAPE" in str(e.value)
    with pytest.raises(S.SkemaError):
        g.update(dev(np.zeros((50, 40))), (45, 60))
    with pytest.raises(S.SkemaError):
        S.SketchySVD(S.AlgParams(matrix_m=50, matrix_n=40, rank=3, window=0))
