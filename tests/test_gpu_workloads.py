"""GPU parity against the oracle at the sizes of the benchmark workloads.

Each workload is built by bench.py's own generator (same seeds), its maps are
drawn on the device and downloaded for the oracle, so both sides see the
same inputs:

* dense 2000 x 1000, rank 10, range 41, core 83, windows of 100 rows
  (sigma_i = 0.8^i plus N(0, 1e-6) noise): the FP64 form at 1e-12 on X, Y, Z,
  U, S, V and the finalize checkpoints, every reduced form at its bound;
* CSR 200000 x 20000 (density 1e-3), sparse-sign maps with 8 distinct rows
  per column, rank 25, range 101, core 203, windows of 10000 rows: the
  whole-matrix pass over the row-blocked CSC with the ELL map kernels;
* Gaussian RBF of 8192 points in d = 128 (10-cluster mixture), gamma = 1/128,
  rank 50, range 201, core 403, windows of 2048 rows: FP64 and BF16 fused
  generators;
* dense 65536 x 16384 (sigma_i = 0.95^i plus N(0, 1e-4) noise), rank 100,
  range 401, core 803, windows of 4096 rows: FP64 at 1e-12 on X, Y, Z and
  every reduced form, row- and column-wise as well as normwise.

Reduced-form bounds are about twice the errors measured at these sizes.
"""
import math

import numpy as np
import pytest
import scipy.sparse as sp

import bench as B
from oracle import skema_oracle as O
from tests.gpu_helpers import TOL, compare_usv, dev, host

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import skema_amd

    skema_amd.load()
    return skema_amd


def _oracle_maps(maps):
    """The device maps, downloaded, as the oracle's DimRedux objects."""
    out = {}
    for name, mp in maps.items():
        if mp.issparse_:
            rp, ci, v = mp.download()
            out[name] = O.SparseSignDimRedux(sp.csr_matrix((v, ci, rp), shape=(mp.nrow, mp.ncol)), mp.init_transposed)
        else:
            out[name] = O.GaussDimRedux(mp.download(), mp.init_transposed)
    return out


def _oparams(cfg, **kw):
    return O.AlgParams(matrix_m=cfg["m"], matrix_n=cfg["n"], rank=cfg["rank"], sketch_range=cfg["k"],
                       sketch_core=cfg["s"], window=cfg["window"], **kw)


def _gparams(S, cfg, **kw):
    return S.AlgParams(matrix_m=cfg["m"], matrix_n=cfg["n"], rank=cfg["rank"], sketch_range=cfg["k"],
                       sketch_core=cfg["s"], window=cfg["window"], **kw)


def _worst_line(G, R, axis):
    """Largest relative error of one row (axis=1) or column (axis=0)."""
    nr = np.linalg.norm(R, axis=axis)
    nr[nr == 0] = 1.0
    return float((np.linalg.norm(G - R, axis=axis) / nr).max())


def _rank_r_distance(U1, S1, V1, U2, S2, V2):
    """||U1 S1 V1^T - U2 S2 V2^T||_F / ||U2 S2 V2^T||_F from r x r products
    (the m x n matrices are never formed)."""
    G1 = (U1.T @ U1) * np.outer(S1, S1) * (V1.T @ V1)
    G2 = (U2.T @ U2) * np.outer(S2, S2) * (V2.T @ V2)
    X12 = (U1.T @ U2) * np.outer(S1, S2) * (V1.T @ V2)
    d2 = G1.sum() + G2.sum() - 2.0 * X12.sum()
    return math.sqrt(max(d2, 0.0) / G2.sum())


# Reduced-form bounds on X, Y (Z: twice) at the benchmark sizes.
LOWP_TOLS = {"bf16": 5e-3, "bf16x3": 1e-5, "fp32": 2e-6, "fp8": 8e-2}


# ------------------------------------------------------------ dense 2000 x 1000
@pytest.fixture(scope="module")
def dense_small(S):
    cfg = B.WORKLOADS["dense-2000x1000"]
    A = B.make_dense(cfg, torch)
    ctx = S.Context(0)
    maps = S.make_maps(_gparams(S, cfg), ctx, dist="normal")
    osk = O.SketchySVD(_oparams(cfg), maps=_oracle_maps(maps))
    Ah = host(A)
    osk.linear_update(Ah)
    ref = dict(X=osk.X.copy(), Y=osk.Y.copy(), Z=osk.Z.copy())
    ref["U"], ref["S"], ref["V"] = osk.low_rank_approx()
    ref["ck"] = dict(osk.checkpoints)
    return cfg, A, Ah, ctx, maps, ref


def test_dense_2000x1000_fp64(S, dense_small):
    cfg, A, Ah, ctx, maps, ref = dense_small
    print("case dense 2000x1000 r=10 k=41 s=83 w=100 fp64", flush=True)
    g = S.SketchySVD(_gparams(S, cfg), maps=maps, ctx=ctx)
    g.linear_update(A)
    X, Y, Z = g.state()
    for name, got in zip("XYZ", (X, Y, Z)):
        assert O.rel_err(got, ref[name]) <= TOL, (name, O.rel_err(got, ref[name]))
    U, Sv, V = g.low_rank_approx()
    compare_usv(U, Sv, V, ref["U"], ref["S"], ref["V"])
    for name in ["P", "Q", "PhiQ", "PsiP", "T1", "T2", "C"]:
        assert O.rel_err(g.checkpoint(name), ref["ck"][name]) <= 1e-11, name


@pytest.mark.parametrize("form", ["fp32", "bf16x3", "bf16", "fp8"])
def test_dense_2000x1000_reduced_forms(S, dense_small, form):
    cfg, A, Ah, ctx, maps, ref = dense_small
    print(f"case dense 2000x1000 {form}", flush=True)
    g = S.SketchySVD(_gparams(S, cfg, precision=form), maps=maps, ctx=ctx)
    g.linear_update(A)
    X, Y, Z = g.state()
    tol = LOWP_TOLS[form]
    ex, ey, ez = O.rel_err(X, ref["X"]), O.rel_err(Y, ref["Y"]), O.rel_err(Z, ref["Z"])
    print(f"{form} dense 2000x1000 rel err X {ex:.3e} Y {ey:.3e} Z {ez:.3e}", flush=True)
    assert ex <= tol and ey <= tol and ez <= 2 * tol, (ex, ey, ez)
    U, Sv, V = g.low_rank_approx()
    err = O.rel_fro_error(Ah, U, Sv, V)
    err_ref = O.rel_fro_error(Ah, ref["U"], ref["S"], ref["V"])
    print(f"{form} dense 2000x1000 rank-10 error {err:.6e} (fp64 oracle {err_ref:.6e})", flush=True)
    assert err <= 1.05 * err_ref, (err, err_ref)


# --------------------------------------------------------- CSR 200000 x 20000
@pytest.fixture(scope="module")
def csr_large(S):
    cfg = B.WORKLOADS["sparse-200000x20000"]
    A = B.make_sparse(cfg)
    ctx = S.Context(0)
    p = _gparams(S, cfg, dim_redux="sparse-sign", issparse=True)
    maps = S.make_maps(p, ctx, convention="cols")
    osk = O.SketchySVD(_oparams(cfg, dim_redux="sparse-sign", issparse=True), maps=_oracle_maps(maps))
    osk.linear_update(A)
    ref = dict(X=osk.X.copy(), Y=osk.Y.copy(), Z=osk.Z.copy())
    ref["U"], ref["S"], ref["V"] = osk.low_rank_approx()
    return cfg, A, ctx, p, maps, ref


def test_csr_200000x20000_sparse_sign(S, csr_large):
    cfg, A, ctx, p, maps, ref = csr_large
    print(f"case csr 200000x20000 nnz={A.nnz} sparse-sign zeta=8 r=25 k=101 s=203 w=10000", flush=True)
    # the maps take the ELL form (8 entries of one magnitude in every column)
    g = S.SketchySVD(p, maps=maps, ctx=ctx)
    Ad = S.CrsMatrix.from_scipy(A)
    g.linear_update(Ad)
    X, Y, Z = g.state()
    for name, got in zip("XYZ", (X, Y, Z)):
        assert O.rel_err(got, ref[name]) <= TOL, (name, O.rel_err(got, ref[name]))
    U, Sv, V = g.low_rank_approx()
    compare_usv(U, Sv, V, ref["U"], ref["S"], ref["V"])
    # a second pass gives the same bits (exact fixed-point accumulation)
    g.linear_update(Ad)
    for a, b in zip((X, Y, Z), g.state()):
        assert np.array_equal(a, b)


# ------------------------------------------------------------ RBF 8192 x 128
@pytest.fixture(scope="module")
def rbf_mid(S):
    cfg = dict(B.WORKLOADS["rbf-100000x128"])
    cfg["m"] = cfg["n"] = 8192
    X = B.make_points(cfg)
    gamma = 1.0 / cfg["d"]
    ctx = S.Context(0)
    p = _gparams(S, cfg, kernel_func="gaussrbf", kernel_gamma=gamma, nfeat=cfg["d"])
    maps = S.make_maps(p, ctx, dist="normal")
    osk = O.SketchySVD(_oparams(cfg, kernel_func="gaussrbf", kernel_gamma=gamma), maps=_oracle_maps(maps))
    osk.linear_update(X)
    ref = dict(X=osk.X.copy(), Y=osk.Y.copy(), Z=osk.Z.copy())
    ref["U"], ref["S"], ref["V"] = osk.low_rank_approx()
    return cfg, X, gamma, ctx, maps, ref


@pytest.mark.parametrize("form", ["fp64", "bf16"])
def test_rbf_8192x128(S, rbf_mid, form):
    cfg, X, gamma, ctx, maps, ref = rbf_mid
    print(f"case rbf 8192 points d=128 gamma=1/128 r=50 k=201 s=403 w=2048 {form}", flush=True)
    p = _gparams(S, cfg, kernel_func="gaussrbf", kernel_gamma=gamma, nfeat=cfg["d"], precision=form)
    g = S.SketchySVD(p, maps=maps, ctx=ctx)
    g.linear_update(dev(X))
    Xg, Yg, Zg = g.state()
    errs = [O.rel_err(a, ref[nm]) for a, nm in zip((Xg, Yg, Zg), "XYZ")]
    print(f"rbf 8192 {form} rel err X {errs[0]:.3e} Y {errs[1]:.3e} Z {errs[2]:.3e}", flush=True)
    U, Sv, V = g.low_rank_approx()
    if form == "fp64":
        assert max(errs) <= TOL, errs
        compare_usv(U, Sv, V, ref["U"], ref["S"], ref["V"])
    else:
        tol = 1e-2  # BF16 Gram of the rounded points and BF16 sketch products
        assert errs[0] <= tol and errs[1] <= tol and errs[2] <= 2 * tol, errs
        es = float(np.max(np.abs(Sv - ref["S"])) / ref["S"][0])
        print(f"rbf 8192 bf16 max |dS| / S0 {es:.3e}", flush=True)
        assert es <= tol


# -------------------------------------------------------- dense 65536 x 16384
@pytest.fixture(scope="module")
def dense_large(S):
    cfg = B.WORKLOADS["dense-65536x16384"]
    A = B.make_dense(cfg, torch)
    ctx = S.Context(0)
    maps = S.make_maps(_gparams(S, cfg), ctx, dist="normal")
    Ah = A.t().cpu().numpy().T  # column-major host copy (8.6 GB)
    osk = O.SketchySVD(_oparams(cfg), maps=_oracle_maps(maps))
    osk.linear_update(Ah)
    ref = dict(X=osk.X, Y=osk.Y, Z=osk.Z)
    return cfg, A, ctx, maps, ref, osk


@pytest.mark.parametrize("form", ["fp64", "fp32", "bf16x3", "bf16", "fp8"])
def test_dense_65536x16384(S, dense_large, form):
    cfg, A, ctx, maps, ref, _ = dense_large
    print(f"case dense 65536x16384 r=100 k=401 s=803 w=4096 {form}", flush=True)
    g = S.SketchySVD(_gparams(S, cfg, precision=form), maps=maps, ctx=ctx)
    g.linear_update(A)
    X, Y, Z = g.state()
    g.close()
    tol = TOL if form == "fp64" else LOWP_TOLS[form]
    ex, ey, ez = O.rel_err(X, ref["X"]), O.rel_err(Y, ref["Y"]), O.rel_err(Z, ref["Z"])
    wx = max(_worst_line(X, ref["X"], 0), _worst_line(X, ref["X"], 1))
    wy = max(_worst_line(Y, ref["Y"], 0), _worst_line(Y, ref["Y"], 1))
    print(f"{form} dense 65536x16384 rel err X {ex:.3e} Y {ey:.3e} Z {ez:.3e}; worst row/col X {wx:.3e} "
          f"Y {wy:.3e}", flush=True)
    assert ex <= tol and ey <= tol and ez <= (tol if form == "fp64" else 2 * tol), (ex, ey, ez)
    assert wx <= 2 * tol and wy <= 2 * tol, (wx, wy)


def test_dense_65536x16384_finalize_fp64(S, dense_large):
    """The FP64 finalize at full size from the oracle's own X, Y, Z: the
    intermediates P = qr(X^T), Q = qr(Y) (65536 x 401), Phi Q, Psi P, T1, T2
    and the core C at 1e-11 against LAPACK (dgeqrf + dorgqr, dgels), the
    singular values at 1e-12.  The noise-dominated spectrum packs the
    singular values around the rank cut closely, so the rank-100 product
    U S V^T is only as well determined as those gaps allow (measured
    relative distance 1.8e-8 for rounding-level differences in C); it is
    bounded at 1e-6."""
    cfg, A, ctx, maps, ref, osk = dense_large
    print("case dense 65536x16384 finalize fp64", flush=True)
    g = S.SketchySVD(_gparams(S, cfg), maps=maps, ctx=ctx)
    g.set_state(ref["X"], ref["Y"], ref["Z"])
    U, Sv, V = g.low_rank_approx()
    Ur, Sr, Vr = osk.low_rank_approx()
    for name in ["P", "Q", "PhiQ", "PsiP", "T1", "T2", "C"]:
        e = O.rel_err(g.checkpoint(name), osk.checkpoints[name])
        print(f"finalize 65536x16384 checkpoint {name}: {e:.3e}", flush=True)
        assert e <= 1e-11, (name, e)
    g.close()
    es = O.rel_err(Sv, Sr)
    ep = _rank_r_distance(U, Sv, V, Ur, Sr, Vr)
    gap = float(np.min(np.abs(np.diff(Sr))) / Sr[0])
    print(f"finalize 65536x16384: S rel err {es:.3e}, rank-100 product rel distance {ep:.3e}, smallest relative gap "
          f"between the leading 100 singular values {gap:.3e}", flush=True)
    assert es <= TOL, es
    assert ep <= 1e-6, ep
