"""GPU checks of the reduced-precision forms of the dense sketch contractions.

Tolerances are normwise relative errors against the FP64 oracle on the same
inputs and maps: single-term BF16 keeps 8 significant bits per operand,
split-BF16 about 16 (then FP32 accumulation), FP32 24, e4m3 FP8 4 with one
power-of-two scale per row of every operand (per row of a map, per row and
per column of a window).
"""
import numpy as np
import pytest

from oracle import skema_oracle as O
from tests.gpu_helpers import dev, gpu_maps, to_gpu_params

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

# About twice the largest normwise error measured on these cases.
TOLS = {"bf16": 5e-3, "bf16x3": 1e-5, "fp32": 2e-6, "fp8": 8e-2}


@pytest.fixture(scope="module")
def S():
    import skema_amd

    skema_amd.load()
    return skema_amd


def _case(m, n, window, seed=0, rank=5):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n)) @ np.diag(0.97 ** np.arange(n)) + 0.05 * rng.standard_normal((m, n))
    p = O.AlgParams(matrix_m=m, matrix_n=n, rank=rank, window=window, sketch_nu=0.9, sketch_eta=1.1)
    osk = O.SketchySVD(p)
    osk.linear_update(A)
    return A, p, osk


def _run(S, A, p, osk, form):
    maps = gpu_maps({"Upsilon": osk.Upsilon, "Omega": osk.Omega, "Phi": osk.Phi, "Psi": osk.Psi})
    try:
        g = S.SketchySVD(to_gpu_params(p, precision=form), maps=maps)
    except S.SkemaError as e:
        if e.code == 7:
            pytest.skip(f"{form} not built: {e}")
        raise
    g.linear_update(dev(A))
    return g


@pytest.mark.parametrize("form", ["bf16", "bf16x3", "fp32", "fp8"])
@pytest.mark.parametrize("m,n,window", [(512, 384, 128), (300, 211, 77)])
def test_lowp_state_against_oracle(S, form, m, n, window):
    print(f"case lowp state {form} {m}x{n} w={window}", flush=True)
    A, p, osk = _case(m, n, window)
    g = _run(S, A, p, osk, form)
    X, Y, Z = g.state()
    tol = TOLS[form]
    ex, ey, ez = O.rel_err(X, osk.X), O.rel_err(Y, osk.Y), O.rel_err(Z, osk.Z)
    print(f"{form} {m}x{n} w={window} rel err X {ex:.3e} Y {ey:.3e} Z {ez:.3e}", flush=True)
    assert ex <= tol and ey <= tol and ez <= 2 * tol, (ex, ey, ez)


@pytest.mark.parametrize("m,n,window,rank", [(384, 200, 128, 10), (1152, 700, 256, 10), (6144, 8192, 2048, 10),
                                             (16514, 8192, 4096, 100)])
def test_fp8_one_pass_over_all_windows(S, m, n, window, rank):
    """nu = 1 and windows of whole 128-row steps: fp8 quantises the whole
    matrix and forms x in one pass (one launch), then runs one right product
    over all m rows and one core product (K = m); the windows only split the
    sums, and their 128-row boundaries are block-scale boundaries too.
    (384, 200, 128): an odd step count, so the pass's padding step lies past
    the last row; (1152, 700, 256): a short last window and a ragged last
    64-column strip; (6144, 8192, 2048): R in 64-column strips (the right
    product on the direct-to-LDS kernel); (16514, 8192, 4096) at rank 100
    (k = 401, s = 803, the largest k the pass takes): one right product of
    65 x 5 tiles of 256 x 256 with no K split, a ragged last row tile and a
    ragged last column tile of [Omega; Psi], and the core product's K split
    chosen to fill the CUs.  The state is not cleared before
    this path (it writes X, Y and Z whole): a first update with another
    matrix must leave no trace.  Every row and column of X and Y is checked."""
    print(f"case fp8 one pass {m}x{n} w={window}", flush=True)
    rng = np.random.default_rng(17)
    A = rng.standard_normal((m, 24)) @ rng.standard_normal((24, n)) + 0.1 * rng.standard_normal((m, n))
    p = O.AlgParams(matrix_m=m, matrix_n=n, rank=rank, window=window, sketch_nu=1.0, sketch_eta=1.1)
    osk = O.SketchySVD(p)
    osk.linear_update(A)
    g = _run(S, 3.0 * rng.standard_normal((m, n)), p, osk, "fp8")
    g.linear_update(dev(A))
    X, Y, Z = g.state()
    tol = TOLS["fp8"]
    ex, ey, ez = O.rel_err(X, osk.X), O.rel_err(Y, osk.Y), O.rel_err(Z, osk.Z)

    def worst(G, R, axis):
        return float((np.linalg.norm(G - R, axis=axis) / np.linalg.norm(R, axis=axis)).max())

    wx = max(worst(X, osk.X, 0), worst(X, osk.X, 1))
    wy = max(worst(Y, osk.Y, 0), worst(Y, osk.Y, 1))
    print(f"fp8 one pass rel err X {ex:.3e} Y {ey:.3e} Z {ez:.3e}; worst row/column X {wx:.3e} Y {wy:.3e}",
          flush=True)
    assert ex <= tol and ey <= tol and ez <= 2 * tol, (ex, ey, ez)
    assert wx <= 2 * tol and wy <= 2 * tol, (wx, wy)


@pytest.mark.parametrize("form", ["bf16", "bf16x3", "fp32"])
@pytest.mark.parametrize("m,n,window,rank", [(1154, 700, 256, 10), (16514, 8192, 4096, 100), (8192, 4096, 2048, 100)])
def test_whole_matrix_update_nu1(S, form, m, n, window, rank):
    """nu = 1 with several windows: bf16 and split-BF16 quantise the whole
    matrix once (their rounding is elementwise, so windows change nothing
    but the FP32 summation order) and run the left, right and core products
    over all rows (fp32 keeps the window loop, checked here the same way):
    (1154, 700, 256) ragged in every dimension;
    (16514, 8192, 4096) at rank 100 (k = 401, s = 803) with the right product
    unsplit over 16514 rows and the core's K split filling the CUs (for
    split-BF16 the right product as one bf16 product over three K segments);
    (8192, 4096, 2048) at rank 100: m a multiple of 64, so split-BF16's left
    product also runs as a segmented bf16 product (K split 2).  A first
    update with another matrix must leave no trace (X, Y and Z are written,
    not accumulated).  Every row and column of X and Y is checked."""
    print(f"case whole-matrix {form} {m}x{n} w={window} rank={rank}", flush=True)
    rng = np.random.default_rng(19)
    A = rng.standard_normal((m, 24)) @ rng.standard_normal((24, n)) + 0.1 * rng.standard_normal((m, n))
    p = O.AlgParams(matrix_m=m, matrix_n=n, rank=rank, window=window, sketch_nu=1.0, sketch_eta=1.1)
    osk = O.SketchySVD(p)
    osk.linear_update(A)
    g = _run(S, 3.0 * rng.standard_normal((m, n)), p, osk, form)
    g.linear_update(dev(A))
    X, Y, Z = g.state()
    tol = TOLS[form]
    ex, ey, ez = O.rel_err(X, osk.X), O.rel_err(Y, osk.Y), O.rel_err(Z, osk.Z)

    def worst(G, R, axis):
        return float((np.linalg.norm(G - R, axis=axis) / np.linalg.norm(R, axis=axis)).max())

    wx = max(worst(X, osk.X, 0), worst(X, osk.X, 1))
    wy = max(worst(Y, osk.Y, 0), worst(Y, osk.Y, 1))
    print(f"{form} whole rel err X {ex:.3e} Y {ey:.3e} Z {ez:.3e}; worst row/column X {wx:.3e} Y {wy:.3e}",
          flush=True)
    assert ex <= tol and ey <= tol and ez <= 2 * tol, (ex, ey, ez)
    assert wx <= 4 * tol and wy <= 4 * tol, (wx, wy)


@pytest.mark.parametrize("form", ["bf16", "bf16x3", "fp32", "fp8"])
def test_lowp_rank_r_error_close_to_fp64(S, form):
    """The finalize is FP64 on every form; the rank-r error of a mixed form
    stays close to the FP64 oracle's on a noisy low-rank matrix."""
    print(f"case lowp rank-r error {form}", flush=True)
    A, p, osk = _case(640, 320, 100, seed=3, rank=8)
    Ur, Sr, Vr = osk.low_rank_approx()
    err_ref = O.rel_fro_error(A, Ur, Sr, Vr)
    g = _run(S, A, p, osk, form)
    U, Sv, V = g.low_rank_approx()
    err = O.rel_fro_error(A, U, Sv, V)
    assert err <= 1.05 * err_ref + 1e-12, (err, err_ref)


def test_lowp_window_offsets_unaligned(S):
    """Windows starting at rows that are not multiples of 8 take the unaligned
    map-load path."""
    print("case lowp unaligned windows", flush=True)
    A, p, osk = _case(203, 130, 29)
    for form in ["bf16x3", "bf16", "fp32", "fp8"]:
        g = _run(S, A, p, osk, form)
        X, Y, Z = g.state()
        ex, ey = O.rel_err(X, osk.X), O.rel_err(Y, osk.Y)
        print(f"{form} unaligned windows rel err X {ex:.3e} Y {ey:.3e}", flush=True)
        assert ex <= TOLS[form] and ey <= TOLS[form], (form, ex, ey)


@pytest.mark.parametrize("form", ["fp8", "bf16"])
def test_lowp_window_products_at_sketch_scale(S, form):
    """k = 401, s = 803 over 4096-row windows (the sizes of the 65536 x 16384
    workload on a narrower matrix): the large-tile NT kernels, their K loop
    over many steps and the split-K partials all take part."""
    print(f"case lowp sketch-scale {form}", flush=True)
    m, n = 8192, 2048
    rng = np.random.default_rng(11)
    A = rng.standard_normal((m, 64)) @ rng.standard_normal((64, n)) + 0.1 * rng.standard_normal((m, n))
    p = O.AlgParams(matrix_m=m, matrix_n=n, rank=100, window=4096, sketch_nu=0.9, sketch_eta=1.1)
    osk = O.SketchySVD(p)
    osk.linear_update(A)
    g = _run(S, A, p, osk, form)
    X, Y, Z = g.state()
    tol = TOLS[form]
    ex, ey, ez = O.rel_err(X, osk.X), O.rel_err(Y, osk.Y), O.rel_err(Z, osk.Z)
    print(f"{form} sketch-scale rel err X {ex:.3e} Y {ey:.3e} Z {ez:.3e}", flush=True)
    assert ex <= tol and ey <= tol and ez <= 2 * tol, (ex, ey, ez)


@pytest.mark.parametrize("form", ["fp8", "bf16"])
def test_lowp_left_product_second_tile_row(S, form):
    """range 480 (between 449 and 512): the fp8 left product's 448-row tile
    covers rows 0..447 and a second tile row holds rows 448..479; 4200 rows
    give one full 4096-row window and a 104-row one."""
    print(f"case lowp {form} range 480 second tile row", flush=True)
    m, n = 4200, 4096
    rng = np.random.default_rng(31)
    A = rng.standard_normal((m, 40)) @ rng.standard_normal((40, n)) + 0.1 * rng.standard_normal((m, n))
    p = O.AlgParams(matrix_m=m, matrix_n=n, rank=100, sketch_range=480, sketch_core=961, window=4096,
                    sketch_nu=0.9, sketch_eta=1.1)
    osk = O.SketchySVD(p)
    osk.linear_update(A)
    g = _run(S, A, p, osk, form)
    X, Y, Z = g.state()
    tol = TOLS[form]
    ex, ey, ez = O.rel_err(X, osk.X), O.rel_err(Y, osk.Y), O.rel_err(Z, osk.Z)
    wx = float((np.linalg.norm(X - osk.X, axis=1) / np.linalg.norm(osk.X, axis=1)).max())
    print(f"{form} range 480 rel err X {ex:.3e} Y {ey:.3e} Z {ez:.3e}; worst row of X {wx:.3e}", flush=True)
    assert ex <= tol and ey <= tol and ez <= 2 * tol, (ex, ey, ez)
    assert wx <= 2 * tol, wx


@pytest.mark.parametrize("form", ["fp8", "bf16", "bf16x3", "fp32"])
def test_lowp_direct_lds_products(S, form):
    """Shapes that take the direct-to-LDS kernel for all three window products:
    right product 4096 x 1204 over K = 8320 (three K splits, the last one
    shorter; the last column tile ragged), left product 401 x 8320 over
    K = 4096 (one 448-row tile for fp8 with column scales, two 256-row tiles for bf16), core 803 x 803
    over K = 4096; the second, 1000-row window takes the other kernels.  Besides the normwise bound, every row and
    column of X and Y is checked on its own, so one wrong tile cannot hide
    in the norm."""
    print(f"case lowp {form} direct-to-LDS products", flush=True)
    m, n = 5096, 8320
    rng = np.random.default_rng(21)
    A = rng.standard_normal((m, 48)) @ rng.standard_normal((48, n)) + 0.1 * rng.standard_normal((m, n))
    p = O.AlgParams(matrix_m=m, matrix_n=n, rank=100, window=4096, sketch_nu=0.9, sketch_eta=1.1)
    osk = O.SketchySVD(p)
    osk.linear_update(A)
    g = _run(S, A, p, osk, form)
    X, Y, Z = g.state()
    tol = TOLS[form]
    ex, ey, ez = O.rel_err(X, osk.X), O.rel_err(Y, osk.Y), O.rel_err(Z, osk.Z)
    print(f"{form} rel err X {ex:.3e} Y {ey:.3e} Z {ez:.3e}", flush=True)
    assert ex <= tol and ey <= tol and ez <= 2 * tol, (ex, ey, ez)

    def worst(G, R, axis):
        return float((np.linalg.norm(G - R, axis=axis) / np.linalg.norm(R, axis=axis)).max())

    wx = max(worst(X, osk.X, 0), worst(X, osk.X, 1))
    wy = max(worst(Y, osk.Y, 0), worst(Y, osk.Y, 1))
    print(f"{form} worst row/column rel err X {wx:.3e} Y {wy:.3e}", flush=True)
    assert wx <= 2 * tol and wy <= 2 * tol, (wx, wy)


@pytest.mark.parametrize("nu", [0.9, 1.0])
@pytest.mark.parametrize("form", ["fp8", "bf16", "bf16x3", "fp32"])
@pytest.mark.parametrize("m,n,window", [(5096, 8320, 4096), (1000, 2100, 256)])
def test_lowp_exact_on_representable_data(S, form, m, n, window, nu):
    """Inputs every reduced form encodes exactly: window entries in {-2, ..., 2}
    times 2^(a + b), a constant over each 32-row block and b over each
    32-column block, and maps with entries in {-2, ..., 2}.  For fp8 each
    block's E8M0 scale then leaves its e4m3 values exact (and the maps' row
    scales theirs); BF16 and FP32 hold these values as they are.  Every
    product and FP32 partial sum is an integer below 2^24, so X and Y of
    every reduced form equal the FP64 oracle's up to the rounding of the FP64
    epilogue; an fp8 scale taken from the wrong block, a dropped scale or a
    misplaced tile is off by a power of two or worse.  The first shape takes
    the direct-to-LDS products, the second (256-row windows) the other
    kernels; with nu = 1 the fp8 form forms x in one pass over all windows.
    The core product of fp8 and bf16 rounds its right operand G to BF16, so
    their Z keeps the BF16 bound; split-BF16 and FP32 carry G exactly, so
    their Z is exact too."""
    print(f"case {form} exact {m}x{n} w={window} nu={nu}", flush=True)
    rng = np.random.default_rng(5)
    a = rng.integers(0, 3, size=-(-m // 32))
    b = rng.integers(0, 3, size=-(-n // 32))
    scale = np.exp2(np.repeat(a, 32)[:m, None] + np.repeat(b, 32)[None, :n])
    A = rng.integers(-2, 3, size=(m, n)).astype(np.float64) * scale
    p = O.AlgParams(matrix_m=m, matrix_n=n, rank=100, window=window, sketch_nu=nu, sketch_eta=1.1)
    maps = O.make_maps(p)
    for mp in maps.values():
        mp.data = rng.integers(-2, 3, size=mp.data.shape).astype(np.float64)
    osk = O.SketchySVD(p, maps=maps)
    osk.linear_update(A)
    g = _run(S, A, p, osk, form)
    X, Y, Z = g.state()
    ex, ey, ez = O.rel_err(X, osk.X), O.rel_err(Y, osk.Y), O.rel_err(Z, osk.Z)
    print(f"{form} exact rel err X {ex:.3e} Y {ey:.3e} Z {ez:.3e}", flush=True)
    assert ex <= 1e-12 and ey <= 1e-12, (ex, ey)
    assert ez <= (1e-12 if form in ("bf16x3", "fp32") else 2 * TOLS["bf16"]), ez
