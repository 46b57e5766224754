"""Thin torch.autograd.Function over the hand-written RBF-sketch adjoint
(for gradcheck and for callers that want autograd; all arithmetic runs in
libskema_mi355x)."""
import torch

from .sketchy import rbf_sketch_loss_grad


class RbfSketchLoss(torch.autograd.Function):
    """loss = ||K Omega^T||_F^2 + ||Phi K Psi^T||_F^2, K = exp(-gamma ||x_i - x_j||^2)."""

    @staticmethod
    def forward(ctx, X, gamma, omega, phi, psi, window):
        Xc = X.detach().t().contiguous().t()
        loss, dX, dg = rbf_sketch_loss_grad(Xc, float(gamma), omega, phi, psi, window=window)
        ctx.save_for_backward(dX, torch.tensor(dg, dtype=torch.float64, device=X.device))
        return torch.tensor(loss, dtype=torch.float64, device=X.device)

    @staticmethod
    def backward(ctx, gout):
        dX, dg = ctx.saved_tensors
        return gout * dX, gout * dg, None, None, None, None


def rbf_sketch_loss(X, gamma, omega, phi, psi, window=4096):
    return RbfSketchLoss.apply(X, gamma, omega, phi, psi, window)
