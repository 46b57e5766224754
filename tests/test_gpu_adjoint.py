"""Adjoint of the fused RBF window + Gauss sketch (differentiation form).

Checked against (1) FP64 finite differences through torch.autograd.gradcheck
at n = 512 and (2) a dense PyTorch autograd reference (K materialised, FP64,
on the CPU) at n = 4096 -- both independent of the closed-form adjoint the
kernel implements.
"""
import numpy as np
import pytest

from oracle import skema_oracle as O
from tests.gpu_helpers import dev

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    import skema_amd

    skema_amd.load()
    return skema_amd


def _points(n, d, seed):
    rng = np.random.default_rng(seed)
    mu = 2.0 * rng.standard_normal((10, d))
    return mu[rng.integers(0, 10, n)] + rng.standard_normal((n, d))


def _maps(S, n, k, s, seed=0):
    O_ = O.gauss_map(k, n, seed)
    F_ = O.gauss_map(s, n, seed + 1)
    P_ = O.gauss_map(s, n, seed + 2)
    return (O_, F_, P_), (S.GaussDimRedux(k, n, data=O_), S.GaussDimRedux(s, n, data=F_),
                          S.GaussDimRedux(s, n, data=P_))


def _dense_reference(X, gamma, Om, Ph, Ps):
    Xt = torch.tensor(X, dtype=torch.float64, requires_grad=True)
    gt = torch.tensor(gamma, dtype=torch.float64, requires_grad=True)
    D = torch.cdist(Xt, Xt) ** 2
    K = torch.exp(-gt * D)
    Om_t, Ph_t, Ps_t = (torch.tensor(a) for a in (Om, Ph, Ps))
    Y = K @ Om_t.T
    W = Ph_t @ K @ Ps_t.T
    L = (Y ** 2).sum() + (W ** 2).sum()
    L.backward()
    return L.item(), Xt.grad.numpy(), gt.grad.item()


@pytest.mark.parametrize("n,d,k,s,window", [(300, 7, 21, 43, 64), (512, 16, 33, 67, 100)])
def test_adjoint_matches_dense_autograd(S, n, d, k, s, window):
    print(f"case adjoint dense ref n={n} d={d}", flush=True)
    X = _points(n, d, n)
    gamma = 0.5 / d
    (Om, Ph, Ps), (om, ph, ps) = _maps(S, n, k, s)
    loss, dX, dg = S.sketchy.rbf_sketch_loss_grad(dev(X), gamma, om, ph, ps, window=window)
    L_ref, dX_ref, dg_ref = _dense_reference(X, gamma, Om, Ph, Ps)
    assert abs(loss - L_ref) <= 1e-11 * abs(L_ref)
    assert O.rel_err(dX.cpu().numpy(), dX_ref) <= 1e-10, O.rel_err(dX.cpu().numpy(), dX_ref)
    assert abs(dg - dg_ref) <= 1e-10 * abs(dg_ref)


def test_adjoint_gradcheck_fp64(S):
    """torch.autograd.gradcheck (central finite differences, FP64) through the
    thin autograd wrapper at n = 512, d = 16, k = 21, s = 43: the full
    Jacobian, every one of the 8192 point coordinates and gamma perturbed
    (fast_mode off), the window (128 rows) splitting the forward pass.  The
    loss is ~1e7 and the gradient entries reach ~3e4: with eps = 1e-4 the
    central differences are within 2e-5 of a dense FP64 autograd gradient
    (measured), so atol = 1e-3 (3e-8 of the largest entry), rtol = 1e-5
    keep a ~50x margin while a 0.1% error in any entry fails."""
    from skema_amd.autograd import rbf_sketch_loss

    print("case adjoint gradcheck n=512 d=16 k=21 s=43 (full)", flush=True)
    n, d, k, s = 512, 16, 21, 43
    X = _points(n, d, 1)
    (Om, Ph, Ps), (om, ph, ps) = _maps(S, n, k, s, seed=5)
    Xt = dev(X).requires_grad_(True)
    # gamma enters as 1e-2 g: the loss is far more curved in gamma than in a
    # coordinate, so its step is 1e-6 (eps * 1e-2) while the points' is 1e-4
    gt = torch.tensor(0.5 / d * 1e2, dtype=torch.float64, device="cuda", requires_grad=True)
    ok = torch.autograd.gradcheck(lambda x, g: rbf_sketch_loss(x, g * 1e-2, om, ph, ps, 128), (Xt, gt), eps=1e-4,
                                  atol=1e-3, rtol=1e-5, fast_mode=False)
    assert ok


def test_adjoint_dense_reference_n4096(S):
    """n = 4096 against the dense autograd reference (CPU, FP64)."""
    print("case adjoint n=4096", flush=True)
    n, d, k, s = 4096, 64, 201, 403
    X = _points(n, d, 9)
    gamma = 0.5 / d
    (Om, Ph, Ps), (om, ph, ps) = _maps(S, n, k, s, seed=11)
    loss, dX, dg = S.sketchy.rbf_sketch_loss_grad(dev(X), gamma, om, ph, ps, window=4096)
    L_ref, dX_ref, dg_ref = _dense_reference(X, gamma, Om, Ph, Ps)
    assert abs(loss - L_ref) <= 1e-10 * abs(L_ref)
    assert O.rel_err(dX.cpu().numpy(), dX_ref) <= 1e-9
    assert abs(dg - dg_ref) <= 1e-9 * abs(dg_ref)


# BF16 form: the Gram, S = F R^T and (H X) on BF16 MFMAs with FP32
# accumulation, K and H in FP32.  K carries ~0.3% relative error (BF16
# Gram of the rounded points), S a BF16 rounding of F; dX = -2 gamma
# (rowsum(H) x - H X) subtracts two nearly equal terms inside a cluster, so
# its normwise error is several times that of K.
BF16_LOSS_TOL, BF16_DX_TOL, BF16_DG_TOL = 2e-2, 1e-1, 5e-2


@pytest.mark.parametrize("n,d,k,s,window", [(300, 7, 21, 43, 64), (4096, 64, 201, 403, 4096)])
def test_adjoint_bf16_form_against_dense_reference(S, n, d, k, s, window):
    print(f"case adjoint bf16 n={n} d={d}", flush=True)
    X = _points(n, d, 9)
    gamma = 0.5 / d
    (Om, Ph, Ps), (om, ph, ps) = _maps(S, n, k, s, seed=11)
    loss, dX, dg = S.sketchy.rbf_sketch_loss_grad(dev(X), gamma, om, ph, ps, window=window, precision="bf16")
    L_ref, dX_ref, dg_ref = _dense_reference(X, gamma, Om, Ph, Ps)
    e_l = abs(loss - L_ref) / abs(L_ref)
    e_x = O.rel_err(dX.cpu().numpy(), dX_ref)
    e_g = abs(dg - dg_ref) / abs(dg_ref)
    print(f"bf16 adjoint n={n}: loss {e_l:.3e} dX {e_x:.3e} dgamma {e_g:.3e}", flush=True)
    assert e_l <= BF16_LOSS_TOL, e_l
    assert e_x <= BF16_DX_TOL, e_x
    assert e_g <= BF16_DG_TOL, e_g
