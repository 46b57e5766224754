"""Thin torch.autograd.Function over the hand-written RBF-sketch adjoint
(for gradcheck and for callers that want autograd; all arithmetic runs in
libskema_mi355x)."""
import torch

from .sketchy import rbf_sketch_loss_grad


class RbfSketchLoss(torch.autograd.Function):
    """loss = ||K Omega^T||_F^2 + ||Phi K Psi^T||_F^2, K = exp(-gamma ||x_i - x_j||^2).

    forward evaluates the loss only; backward runs the adjoint (which forms
    the kernel tiles again), so evaluations that need no gradient (finite
    differences) cost one forward pass each."""

    @staticmethod
    def forward(ctx, X, gamma, omega, phi, psi, window):
        Xc = X.detach().t().contiguous().t()
        g = float(gamma)
        loss, _, _ = rbf_sketch_loss_grad(Xc, g, omega, phi, psi, window=window, grad=False)
        ctx.save_for_backward(Xc)
        ctx.args = (g, omega, phi, psi, window)
        return torch.tensor(loss, dtype=torch.float64, device=X.device)

    @staticmethod
    def backward(ctx, gout):
        (Xc,) = ctx.saved_tensors
        g, omega, phi, psi, window = ctx.args
        _, dX, dg = rbf_sketch_loss_grad(Xc, g, omega, phi, psi, window=window, grad=True)
        return gout * dX, gout * torch.tensor(dg, dtype=torch.float64, device=Xc.device), None, None, None, None


def rbf_sketch_loss(X, gamma, omega, phi, psi, window=4096):
    return RbfSketchLoss.apply(X, gamma, omega, phi, psi, window)
