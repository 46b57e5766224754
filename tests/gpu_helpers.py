"""Shared helpers for the GPU parity tests (device buffers, oracle maps)."""
import numpy as np

from oracle import skema_oracle as O

TOL = 1e-12  # relative Frobenius tolerance of the FP64 plain form


def dev(a):
    """numpy (rows, cols) -> column-major float64 CUDA tensor."""
    import torch

    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        return torch.as_tensor(a, device="cuda")
    return torch.as_tensor(np.ascontiguousarray(a.T), device="cuda").t()


def host(t):
    return t.detach().cpu().numpy()


def gpu_maps(omaps, ctx=None):
    """Upload the oracle's materialised maps into the library."""
    import skema_amd as S

    out = {}
    for name, mp in omaps.items():
        r, c = mp.data.shape
        if mp.issparse:
            d = mp.data.tocsr()
            d.sort_indices()
            out[name] = S.SparseSignDimRedux(r, c, init_transposed=mp.init_transposed,
                                             csr=(d.indptr, d.indices, d.data), ctx=ctx)
        else:
            out[name] = S.GaussDimRedux(r, c, init_transposed=mp.init_transposed, data=mp.data, ctx=ctx)
    return out


def to_gpu_params(p, **extra):
    import skema_amd as S

    return S.AlgParams(matrix_m=p.matrix_m, matrix_n=p.matrix_n, rank=p.rank, sketch_range=p.sketch_range,
                       sketch_core=p.sketch_core, window=p.window, seeds=list(p.seeds), sketch_eta=p.sketch_eta,
                       sketch_nu=p.sketch_nu, dim_redux=p.dim_redux, issparse=p.issparse,
                       kernel_func=p.kernel_func, kernel_gamma=p.kernel_gamma, **extra)


def compare_usv(U, S, V, Ur, Sr, Vr, tol=TOL):
    assert O.rel_err(S, Sr) <= tol, O.rel_err(S, Sr)
    Ua, Va = O.align_signs(U, V, Ur, Vr)
    assert O.rel_err(Ua, Ur) <= tol, O.rel_err(Ua, Ur)
    assert O.rel_err(Va, Vr) <= tol, O.rel_err(Va, Vr)
    assert O.rel_err((U * S) @ V.T, (Ur * Sr) @ Vr.T) <= tol
