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

TOLS = {"bf16": 2e-2, "bf16x3": 5e-5, "fp32": 5e-5, "fp8": 1.5e-1}


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
    assert O.rel_err(X, osk.X) <= tol, O.rel_err(X, osk.X)
    assert O.rel_err(Y, osk.Y) <= tol, O.rel_err(Y, osk.Y)
    assert O.rel_err(Z, osk.Z) <= 2 * tol, O.rel_err(Z, osk.Z)


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
        assert O.rel_err(X, osk.X) <= TOLS[form]
        assert O.rel_err(Y, osk.Y) <= TOLS[form]


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
    assert O.rel_err(X, osk.X) <= tol, O.rel_err(X, osk.X)
    assert O.rel_err(Y, osk.Y) <= tol, O.rel_err(Y, osk.Y)
    assert O.rel_err(Z, osk.Z) <= 2 * tol, O.rel_err(Z, osk.Z)


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
