"""GPU parity of the CSR-input forms and of the RBF window generator."""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import skema_oracle as O
from tests.gpu_helpers import TOL, compare_usv, dev, gpu_maps, host, to_gpu_params

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import skema_amd

    skema_amd.load()
    return skema_amd


def _sparse(m, n, density, seed):
    rng = np.random.default_rng(seed)
    A = sp.random(m, n, density=density, random_state=rng, data_rvs=rng.standard_normal, format="csr")
    A.sort_indices()
    return A


def _omaps(osk):
    return {"Upsilon": osk.Upsilon, "Omega": osk.Omega, "Phi": osk.Phi, "Psi": osk.Psi}


@pytest.mark.parametrize("model,conv", [("gauss", "rows"), ("sparse-sign", "rows"), ("sparse-sign", "cols")])
@pytest.mark.parametrize("window,nu,eta", [(37, 1.0, 1.0), (400, 0.8, 1.2), (64, 0.5, 2.0)])
def test_csr_sketch_parity(S, model, conv, window, nu, eta):
    print(f"case csr sketch {model}/{conv} w={window} nu={nu} eta={eta}", flush=True)
    m, n = 400, 150
    A = _sparse(m, n, 0.05, 1)
    p = O.AlgParams(matrix_m=m, matrix_n=n, rank=5, window=window, dim_redux=model, issparse=True,
                    sketch_nu=nu, sketch_eta=eta)
    osk = O.SketchySVD(p, sparse_convention=conv)
    osk.linear_update(A)
    Xr, Yr, Zr = osk.X.copy(), osk.Y.copy(), osk.Z.copy()
    Ur, Sr, Vr = osk.low_rank_approx()
    g = S.SketchySVD(to_gpu_params(p), maps=gpu_maps(_omaps(osk)))
    g.linear_update(S.CrsMatrix.from_scipy(A))
    X, Y, Z = g.state()
    assert O.rel_err(X, Xr) <= TOL
    assert O.rel_err(Y, Yr) <= TOL
    assert O.rel_err(Z, Zr) <= TOL
    U, Sv, V = g.low_rank_approx()
    compare_usv(U, Sv, V, Ur, Sr, Vr)


def test_csr_fixture_exact_recovery(S, golden_dir):
    """rectangular_sparse.mtx through the CSR path (maps stored as the
    reference stores them for sparse input)."""
    from scipy.io import mmread

    print("case csr fixture", flush=True)
    A = mmread(f"{golden_dir}/rectangular_sparse.mtx").tocsr()
    sv = np.loadtxt(f"{golden_dir}/rectangular_svals.txt")
    for model in ["gauss", "sparse-sign"]:
        p = S.AlgParams(matrix_m=32, matrix_n=14, rank=10, sketch_range=14, sketch_core=29, window=6,
                        dim_redux=model, issparse=True)
        g = S.SketchySVD(p)
        g.linear_update(S.CrsMatrix.from_scipy(A))
        U, Sv, V = g.low_rank_approx()
        assert np.max(np.abs(Sv - sv[:10])) <= 1e-13 * sv[0], model


def test_csr_per_window_update(S):
    print("case csr per-window", flush=True)
    m, n = 300, 90
    A = _sparse(m, n, 0.1, 2)
    p = O.AlgParams(matrix_m=m, matrix_n=n, rank=3, window=70, issparse=True)
    osk = O.SketchySVD(p)
    osk.linear_update(A)
    Ad = S.CrsMatrix.from_scipy(A)
    g = S.SketchySVD(to_gpu_params(p), maps=gpu_maps(_omaps(osk)))
    for r0 in range(0, m, 70):
        g.update(Ad, (r0, min(m, r0 + 70)))
    X, Y, Z = g.state()
    assert O.rel_err(X, osk.X) <= TOL and O.rel_err(Y, osk.Y) <= TOL and O.rel_err(Z, osk.Z) <= TOL


def test_csr_empty_rows_and_columns(S):
    """Ragged input: empty rows, an empty window, empty columns."""
    print("case csr ragged", flush=True)
    m, n = 120, 40
    A = _sparse(m, n, 0.05, 3).tolil()
    A[10:50, :] = 0.0
    A[:, 5:9] = 0.0
    A = A.tocsr()
    A.eliminate_zeros()
    p = O.AlgParams(matrix_m=m, matrix_n=n, rank=3, window=20, issparse=True, dim_redux="sparse-sign")
    osk = O.SketchySVD(p)
    osk.linear_update(A)
    g = S.SketchySVD(to_gpu_params(p), maps=gpu_maps(_omaps(osk)))
    g.linear_update(S.CrsMatrix.from_scipy(A))
    X, Y, Z = g.state()
    assert O.rel_err(X, osk.X) <= TOL and O.rel_err(Y, osk.Y) <= TOL and O.rel_err(Z, osk.Z) <= TOL


@pytest.mark.parametrize("model", ["sparse-sign", "gauss"])
def test_csr_whole_pass_row_blocked_csc(S, model):
    """A matrix tall enough (>= 4096 rows) for the row-blocked CSC build of
    the whole-matrix pass (LDS column histograms per row block, nonzeros
    walked in storage order, rows found by binary search in the block's
    slice of rowptr), with an empty band of rows and empty columns."""
    print(f"case csr whole pass blocked csc {model}", flush=True)
    m, n = 20000, 3000
    A = _sparse(m, n, 2e-3, 17).tolil()
    A[7000:7600, :] = 0.0
    A[:, 100:140] = 0.0
    A = A.tocsr()
    A.eliminate_zeros()
    A.sort_indices()
    p = O.AlgParams(matrix_m=m, matrix_n=n, rank=8, window=5000, issparse=True, dim_redux=model)
    osk = O.SketchySVD(p)
    osk.linear_update(A)
    g = S.SketchySVD(to_gpu_params(p), maps=gpu_maps(_omaps(osk)))
    g.linear_update(S.CrsMatrix.from_scipy(A))
    X, Y, Z = g.state()
    assert O.rel_err(X, osk.X) <= TOL and O.rel_err(Y, osk.Y) <= TOL and O.rel_err(Z, osk.Z) <= TOL


@pytest.mark.parametrize("conv", ["cols", "rows"])
def test_csr_whole_pass_bitwise_reproducible(S, conv):
    """The sparse-map products of the whole-matrix pass accumulate in exact
    fixed point, so two passes over the same CSR give identical bits for X,
    Y and Z (the CSC is rebuilt each time, its column order varying with the
    atomics), and both equal the oracle to 1e-12.  "cols": the ELL form
    (zeta distinct rows per map column, one magnitude); "rows": the
    reference's convention (per stored row, with replacement), the general
    CSR form of the maps."""
    print(f"case csr whole pass reproducible {conv}", flush=True)
    m, n = 30000, 4000
    A = _sparse(m, n, 2e-3, 23)
    A.data[::97] *= 1e6  # a wide dynamic range within rows and columns
    p = O.AlgParams(matrix_m=m, matrix_n=n, rank=10, window=7000, issparse=True, dim_redux="sparse-sign",
                    sketch_nu=0.9, sketch_eta=1.1)
    osk = O.SketchySVD(p, sparse_convention=conv)
    osk.linear_update(A)
    maps = gpu_maps(_omaps(osk))
    Ad = S.CrsMatrix.from_scipy(A)
    runs = []
    for _ in range(2):
        g = S.SketchySVD(to_gpu_params(p), maps=maps)
        g.linear_update(Ad)
        runs.append(g.state())
    for a, b in zip(*runs):
        assert np.array_equal(a, b)
    X, Y, Z = runs[0]
    assert O.rel_err(X, osk.X) <= TOL and O.rel_err(Y, osk.Y) <= TOL and O.rel_err(Z, osk.Z) <= TOL


def test_csr_bad_column_index_is_rejected(S):
    """A column index >= n is caught by the validation pass (an error, not a
    fault): the pass only reads the col entries rowptr addresses."""
    print("case csr bad column", flush=True)
    m, n = 20, 10
    A = _sparse(m, n, 0.3, 4)
    Ad = S.CrsMatrix.from_scipy(A)
    Ad.col[3] = n + 5
    p = S.AlgParams(matrix_m=m, matrix_n=n, rank=2, window=5, issparse=True)
    g = S.SketchySVD(p)
    with pytest.raises(S.SkemaError) as e:
        g.linear_update(Ad)
    assert "column index" in str(e.value)


def test_sparse_sign_dense_input_cols_convention(S):
    print("case dense x sparse-sign cols", flush=True)
    m, n = 256, 120
    A = np.random.default_rng(5).standard_normal((m, n))
    p = O.AlgParams(matrix_m=m, matrix_n=n, rank=4, window=50, dim_redux="sparse-sign")
    osk = O.SketchySVD(p, sparse_convention="cols")
    osk.linear_update(A)
    Ur, Sr, Vr = osk.low_rank_approx()
    g = S.SketchySVD(to_gpu_params(p), maps=gpu_maps(_omaps(osk)))
    g.linear_update(dev(A))
    X, Y, Z = g.state()
    assert O.rel_err(X, osk.X) <= TOL and O.rel_err(Y, osk.Y) <= TOL and O.rel_err(Z, osk.Z) <= TOL
    U, Sv, V = g.low_rank_approx()
    compare_usv(U, Sv, V, Ur, Sr, Vr)


@pytest.mark.parametrize("transposed", [False, True])
def test_sparse_lmap_rmap(S, transposed):
    print(f"case sparse lmap/rmap transposed={transposed}", flush=True)
    rng = np.random.default_rng(6)
    nrow, ncol = 31, 200
    D = O.sparse_sign_rows(nrow, ncol, 3)
    om = O.SparseSignDimRedux(D, transposed)
    gm = S.SparseSignDimRedux(nrow, ncol, init_transposed=transposed, csr=(D.indptr, D.indices, D.data))
    for tA in ["N", "T"]:
        eff = tA if not transposed else ("T" if tA == "N" else "N")
        for idx in [(0, 0), (17, 140)]:
            wc = idx[1] - idx[0] if idx[0] != idx[1] else ncol
            B = rng.standard_normal((wc if eff == "N" else nrow, 9))
            ref = om.lmap(0.75, B, 0.0, tA, "N", idx)
            got = host(gm.lmap(0.75, dev(B), 0.0, tA, "N", idx))
            assert O.rel_err(got, ref) <= TOL
    for tB in ["N", "T"]:
        A = rng.standard_normal((13, ncol if tB == "T" else nrow))
        ref = om.rmap(1.25, A, 0.0, "N", tB)
        got = host(gm.rmap(1.25, dev(A), 0.0, "N", tB))
        assert O.rel_err(got, ref) <= TOL


# ------------------------------------------------------------------- RBF ---
def test_rbf_window_parity(S):
    print("case rbf window", flush=True)
    X = np.random.default_rng(7).standard_normal((333, 7))
    Xd = dev(X)
    for idx in [(0, 64), (100, 171), (332, 333)]:
        ref = O.gauss_rbf_compute(X[idx[0]:idx[1]], X, 0.3, idx)
        got = host(S.gauss_rbf_window(Xd, 0.3, idx))
        assert np.max(np.abs(got - ref)) <= 1e-15


@pytest.mark.parametrize("window", [4, 25])
def test_kernel_fixture_exact_recovery(S, golden_dir, window):
    print(f"case kernel fixture w={window}", flush=True)
    X = np.loadtxt(f"{golden_dir}/kernel_data.txt")
    sv = np.loadtxt(f"{golden_dir}/kernel_svals.txt")
    p = S.AlgParams(matrix_m=25, matrix_n=25, rank=10, sketch_range=25, sketch_core=51, window=window,
                    kernel_func="gaussrbf", kernel_gamma=1.0, nfeat=2)
    g = S.SketchySVD(p)
    g.linear_update(dev(X))
    U, Sv, V = g.low_rank_approx()
    assert np.max(np.abs(Sv - sv[:10])) <= 1e-13 * sv[0]


def test_rbf_sketch_parity(S):
    print("case rbf sketch", flush=True)
    rng = np.random.default_rng(8)
    n, d = 500, 16
    X = rng.standard_normal((n, d)) + 2.0 * rng.standard_normal((5, d))[rng.integers(0, 5, n)]
    p = O.AlgParams(matrix_m=n, matrix_n=n, rank=6, window=96, kernel_func="gaussrbf", kernel_gamma=1.0 / d)
    osk = O.SketchySVD(p)
    osk.linear_update(X)
    Xr, Yr, Zr = osk.X.copy(), osk.Y.copy(), osk.Z.copy()
    Ur, Sr, Vr = osk.low_rank_approx()
    g = S.SketchySVD(to_gpu_params(p, nfeat=d), maps=gpu_maps(_omaps(osk)))
    g.linear_update(dev(X))
    Xg, Yg, Zg = g.state()
    assert O.rel_err(Xg, Xr) <= TOL and O.rel_err(Yg, Yr) <= TOL and O.rel_err(Zg, Zr) <= TOL
    U, Sv, V = g.low_rank_approx()
    compare_usv(U, Sv, V, Ur, Sr, Vr)


@pytest.mark.parametrize("n,d,w,rank", [(333, 37, 50, 4), (257, 128, 64, 6), (130, 3, 17, 3)])
def test_rbf_fused_odd_shapes(S, n, d, w, rank):
    """Fused window generator at feature counts / windows / point counts that
    are not multiples of the tile sizes."""
    print(f"case rbf fused n={n} d={d} w={w}", flush=True)
    rng = np.random.default_rng(n + d)
    X = rng.standard_normal((n, d))
    p = O.AlgParams(matrix_m=n, matrix_n=n, rank=rank, window=w, kernel_func="gaussrbf", kernel_gamma=0.5 / d,
                    sketch_nu=0.95, sketch_eta=1.05)
    osk = O.SketchySVD(p)
    osk.linear_update(X)
    g = S.SketchySVD(to_gpu_params(p, nfeat=d), maps=gpu_maps(_omaps(osk)))
    g.linear_update(dev(X))
    Xg, Yg, Zg = g.state()
    assert O.rel_err(Xg, osk.X) <= TOL and O.rel_err(Yg, osk.Y) <= TOL and O.rel_err(Zg, osk.Z) <= TOL


@pytest.mark.parametrize("n,d,w,rank", [(2600, 64, 300, 10), (1200, 20, 256, 60)])
def test_rbf_fused_split_points_and_wide_sketches(S, n, d, w, rank):
    """Point counts large enough that the right-hand product splits the
    points over blocks (partial sums reduced afterwards), and k + s above 640 /
    k above 208 so both fused kernels tile their columns over several blocks."""
    print(f"case rbf fused large n={n} d={d} w={w} r={rank}", flush=True)
    rng = np.random.default_rng(n + d)
    X = rng.standard_normal((n, d)) + 2.0 * rng.standard_normal((10, d))[rng.integers(0, 10, n)]
    p = O.AlgParams(matrix_m=n, matrix_n=n, rank=rank, window=w, kernel_func="gaussrbf", kernel_gamma=1.0 / d,
                    sketch_nu=0.9, sketch_eta=1.1)
    osk = O.SketchySVD(p)
    osk.linear_update(X)
    g = S.SketchySVD(to_gpu_params(p, nfeat=d), maps=gpu_maps(_omaps(osk)))
    g.linear_update(dev(X))
    Xg, Yg, Zg = g.state()
    assert O.rel_err(Xg, osk.X) <= TOL and O.rel_err(Yg, osk.Y) <= TOL and O.rel_err(Zg, osk.Z) <= TOL


# BF16 form of the fused generator: BF16 Gram (one product, FP32
# accumulation), FP32 exponent, BF16 sketch MFMAs with FP32 accumulation.
# Rounding K and the maps to BF16 (8 significant bits) and the Gram's
# rounding errors (independent over the d products) leave the sketches within
# a few 1e-3 normwise of the FP64 oracle; 2e-2 is the bound used for the dense
# BF16 form as well.
BF16_TOL = 2e-2


@pytest.mark.parametrize("n,d,w,rank", [(333, 37, 50, 4), (257, 128, 64, 6), (130, 3, 17, 3), (2600, 64, 300, 10),
                                        (1200, 20, 256, 60), (3000, 50, 512, 25), (1000, 100, 300, 40)])
def test_rbf_bf16_form_against_oracle(S, n, d, w, rank):
    print(f"case rbf bf16 n={n} d={d} w={w} r={rank}", flush=True)
    rng = np.random.default_rng(n + d + 1)
    X = rng.standard_normal((n, d)) + 2.0 * rng.standard_normal((10, d))[rng.integers(0, 10, n)]
    p = O.AlgParams(matrix_m=n, matrix_n=n, rank=rank, window=w, kernel_func="gaussrbf", kernel_gamma=1.0 / d,
                    sketch_nu=0.9, sketch_eta=1.1)
    osk = O.SketchySVD(p)
    osk.linear_update(X)
    Ur, Sr, Vr = osk.low_rank_approx()
    g = S.SketchySVD(to_gpu_params(p, nfeat=d, precision="bf16"), maps=gpu_maps(_omaps(osk)))
    g.linear_update(dev(X))
    Xg, Yg, Zg = g.state()
    assert O.rel_err(Xg, osk.X) <= BF16_TOL, O.rel_err(Xg, osk.X)
    assert O.rel_err(Yg, osk.Y) <= BF16_TOL, O.rel_err(Yg, osk.Y)
    assert O.rel_err(Zg, osk.Z) <= 2 * BF16_TOL, O.rel_err(Zg, osk.Z)
    U, Sv, V = g.low_rank_approx()
    # the leading singular values of the approximation agree to the same order
    assert np.max(np.abs(Sv[:rank] - Sr[:rank])) <= BF16_TOL * Sr[0], (Sv[:rank], Sr[:rank])


def test_cli_driver_on_reference_fixtures(golden_dir, tmp_path):
    """The C++ driver (include/skema_mi355x.hpp layer) on the reference's own
    fixture files: MatrixMarket coordinate (dense and --sparse) and the
    Kokkos-Kernels .bin, exact recovery of the stored singular values."""
    import json
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "skema_amd", "skema_sketch")
    if not os.path.exists(exe):
        pytest.skip("driver not built")
    cases = [
        (["--input", f"{golden_dir}/rectangular_sparse.mtx", "--rank", "10", "--range", "14", "--core", "29",
          "--window", "5"], "rectangular_svals.txt", 10),
        (["--input", f"{golden_dir}/rectangular_sparse.mtx", "--sparse", "--rank", "10", "--range", "14",
          "--core", "29", "--window", "6", "--model", "sparse-sign"], "rectangular_svals.txt", 10),
        (["--input", f"{golden_dir}/symmetric_sparse.bin", "--sparse", "--rank", "27", "--range", "27",
          "--core", "55", "--window", "4"], "symmetric_svals.txt", 27),
    ]
    for args, svals, r in cases:
        print("case cli", " ".join(args), flush=True)
        hist = tmp_path / "hist.json"
        subprocess.run([exe, *args, "--history-file", str(hist)], check=True, timeout=120)
        h = json.loads(hist.read_text())
        sv = np.loadtxt(f"{golden_dir}/{svals}")
        assert np.max(np.abs(np.array(h["svals"]) - sv[:r])) <= 1e-12 * sv[0]
        assert set(h["timings"]) >= {"init", "update", "approx"}
        # compute_residuals: exact recovery leaves residuals at rounding level
        assert len(h["rnrms"]) == r and max(h["rnrms"]) <= 1e-10 * sv[0], h["rnrms"]


def test_cli_driver_spd_route(golden_dir, tmp_path):
    """--symmetric (and kernel input unless --force-three-sketch) takes the
    SketchySPD route: the stored eigenvalues of the reference's symmetric and
    kernel fixtures come back as svals, with eigenpair residuals."""
    import json
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "skema_amd", "skema_sketch")
    if not os.path.exists(exe):
        pytest.skip("driver not built")
    X = np.loadtxt(f"{golden_dir}/kernel_data.txt")
    kmtx = tmp_path / "kernel_points.mtx"
    with open(kmtx, "w") as f:
        f.write("%%MatrixMarket matrix array real general\n")
        f.write(f"{X.shape[0]} {X.shape[1]}\n")
        for v in X.T.reshape(-1):
            f.write(f"{v:.17g}\n")
    n = X.shape[0]
    cases = [
        (["--input", f"{golden_dir}/symmetric_sparse.mtx", "--sparse", "--symmetric", "--rank", "27", "--range", "27",
          "--window", "4"], "symmetric_svals.txt", 27, True),
        (["--input", f"{golden_dir}/symmetric_sparse.mtx", "--symmetric", "--rank", "27", "--range", "27",
          "--window", "5"], "symmetric_svals.txt", 27, True),
        (["--input", str(kmtx), "--kernel", "gaussrbf", "--gamma", "1.0", "--rank", str(n), "--range", str(n),
          "--window", "7"], "kernel_svals.txt", n, True),
        (["--input", str(kmtx), "--kernel", "gaussrbf", "--gamma", "1.0", "--force-three-sketch", "--rank", str(n),
          "--range", str(n), "--core", str(2 * n + 1), "--window", "7"], "kernel_svals.txt", n, False),
    ]
    for args, svals, r, spd in cases:
        print("case cli", " ".join(args), flush=True)
        hist = tmp_path / "hist.json"
        subprocess.run([exe, *args, "--history-file", str(hist)], check=True, timeout=120)
        h = json.loads(hist.read_text())
        sv = np.loadtxt(f"{golden_dir}/{svals}")
        assert np.max(np.abs(np.array(h["svals"]) - sv[:r])) <= 1e-12 * sv[0], (h["svals"][:3], sv[:3])
        assert len(h["rnrms"]) == r and max(h["rnrms"]) <= 1e-9 * sv[0], h["rnrms"]
        # both routes write the reference's history keys; the SketchySPD
        # route has its own timing keys (src/Skema_SketchySVD.cpp:776-786)
        assert set(h) >= {"timings", "traces", "svals", "rnrms"}
        assert ("dpotrf" in h["timings"]["approx"]) == spd
