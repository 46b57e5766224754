"""The residual pass entry point that allocates no device memory
(skema_residual_fro_dense_noalloc) against the allocating one
(skema_residual_fro_dense) through the C ABI.

Both run the same windowed kernels; the no-allocation form takes its scratch
from one static 32 MiB device array per GPU, so when the requested window
fits there it must return the same bits, and when windows are cut to fit it
differs only by the order of the FP64 partial sums.  Two calls in a row (the
second on a smaller input) check that nothing of the first call's contents
leaks into the second.
"""
import numpy as np
import pytest

from tests.gpu_helpers import dev

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

SCRATCH_ROW_LIMIT = 4193248  # n + max(r, 1) that one row of the static array holds


@pytest.fixture(scope="module")
def S():
    import skema_amd

    skema_amd.load()
    return skema_amd


def _factors(m, n, r, seed):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((m, n))
    U = rng.standard_normal((m, r))
    V = rng.standard_normal((n, r))
    sv = rng.random(r) + 0.5
    return A, U, sv, V


def _fp64_reference(A, U, sv, V):
    R = A - (U * sv) @ V.T
    return float(np.sum(R * R)), float(np.sum(A * A))


def test_noalloc_matches_allocating_entry_two_calls(S):
    """First call 3000 x 700 (r = 12, windows of 256), second 900 x 300
    (r = 5, windows of 128): both fit the static array, so the sums are
    bit-identical to the allocating entry point's."""
    for m, n, r, w, seed in [(3000, 700, 12, 256, 1), (900, 300, 5, 128, 2)]:
        print(f"case no-alloc residual {m}x{n} r={r} w={w}", flush=True)
        A, U, sv, V = _factors(m, n, r, seed)
        Ad, Ud, Sd, Vd = dev(A), dev(U), dev(sv), dev(V)
        e_ref, a_ref = S.residual_fro(Ad, Ud, Sd, Vd, window=w)
        e, a = S.residual_fro_noalloc(Ad, Ud, Sd, Vd, window=w)
        assert e == e_ref and a == a_ref, (e, e_ref, a, a_ref)
        e64, a64 = _fp64_reference(A, U, sv, V)
        assert abs(e - e64) <= 1e-12 * e64 and abs(a - a64) <= 1e-12 * a64


def test_noalloc_cuts_windows_to_the_static_array(S):
    """n = 40000 columns: a 512-row window needs 164 MB, so the static array
    takes 104 rows at a time; the sums then agree with the allocating entry
    point to FP64 rounding.  A smaller call afterwards is bit-identical again."""
    m, n, r, w = 600, 40000, 20, 512
    print(f"case no-alloc residual cut windows {m}x{n} r={r} w={w}", flush=True)
    A, U, sv, V = _factors(m, n, r, 3)
    Ad, Ud, Sd, Vd = dev(A), dev(U), dev(sv), dev(V)
    e_ref, a_ref = S.residual_fro(Ad, Ud, Sd, Vd, window=w)
    e, a = S.residual_fro_noalloc(Ad, Ud, Sd, Vd, window=w)
    assert abs(e - e_ref) <= 1e-13 * e_ref and abs(a - a_ref) <= 1e-13 * a_ref, (e, e_ref, a, a_ref)
    e64, a64 = _fp64_reference(A, U, sv, V)
    assert abs(e - e64) <= 1e-12 * e64 and abs(a - a64) <= 1e-12 * a64
    del Ad
    print("case no-alloc residual after cut windows 50x60 r=3", flush=True)
    A, U, sv, V = _factors(50, 60, 3, 4)
    args = (dev(A), dev(U), dev(sv), dev(V))
    assert S.residual_fro_noalloc(*args, window=16) == S.residual_fro(*args, window=16)


def test_noalloc_rejects_rows_beyond_the_static_array(S):
    """One row of n + max(r, 1) doubles must fit the static array; a wider
    matrix is refused on the host before any kernel runs, and the next call
    works."""
    n = SCRATCH_ROW_LIMIT  # with r = 1: n + 1 > limit
    print(f"case no-alloc residual rejects 1x{n}", flush=True)
    A = torch.zeros((n, 1), dtype=torch.float64, device="cuda").t()  # 1 x n, column-major (ld = 1)
    U = torch.ones((1, 1), dtype=torch.float64, device="cuda")
    V = torch.zeros((n, 1), dtype=torch.float64, device="cuda")
    sv = torch.ones(1, dtype=torch.float64, device="cuda")
    with pytest.raises(S.SkemaError) as e:
        S.residual_fro_noalloc(A, U, sv, V, window=1)
    assert e.value.code == 2 and "static scratch" in str(e.value)
    # the widest matrix that fits (n + 1 = limit) runs
    A2 = torch.ones((n - 1, 1), dtype=torch.float64, device="cuda").t()
    V2 = torch.zeros((n - 1, 1), dtype=torch.float64, device="cuda")
    e2, a2 = S.residual_fro_noalloc(A2, U, sv, V2, window=4)
    assert e2 == float(n - 1) and a2 == float(n - 1)
