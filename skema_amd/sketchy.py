"""Host-side mirror of Skema's SketchySVD interface over libskema_mi355x.

Class and method names follow the reference (src/Skema_SketchySVD.hpp:11-64,
src/Skema_DimRedux.hpp:20-308, src/Skema_Window.hpp:15-133) so a driver or a
test reads like the reference's own code.  Device memory comes from PyTorch
(plumbing only); every computation runs in the HIP library.

Layout: dense matrices are torch tensors of shape (rows, cols) in
column-major storage (``t.stride(0) == 1``), like Kokkos::LayoutLeft.  Use
``colmajor(t)`` to obtain such a tensor.
"""
from __future__ import annotations

import ctypes as C
import json
from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple

import numpy as np

from . import _lib as L

try:
    import torch
except ImportError:  # pragma: no cover - torch is required for device buffers
    torch = None


def colmajor(t):
    """Column-major (LayoutLeft) copy of a 2-D tensor, on the same device."""
    return t.t().contiguous().t()


def _ld(t) -> int:
    rows, cols = t.shape
    if t.stride(0) != 1 and cols > 1 and rows > 1:
        raise ValueError("tensor must be column-major (stride(0) == 1); use colmajor()")
    return max(int(t.stride(1)) if cols > 1 else rows, rows, 1)


def _ptr(t) -> int:
    return int(t.data_ptr())


class Context:
    """Device context (replaces Kokkos::initialize, driver/main.cpp:150)."""

    def __init__(self, device: int = 0, stream=None):
        L.load()
        h = C.c_void_p()
        L.call("skema_ctx_create", device, C.byref(h))
        self.h = h
        self.device = device
        if stream is not None:
            L.call("skema_ctx_set_stream", self.h, C.c_void_p(stream))

    def set_stream(self, stream_handle: int) -> None:
        L.call("skema_ctx_set_stream", self.h, C.c_void_p(stream_handle))

    def set_async(self, on: bool) -> None:
        L.call("skema_ctx_set_async", self.h, int(on))

    def synchronize(self) -> None:
        L.call("skema_ctx_synchronize", self.h)

    def set_profiling(self, on: bool, only: Optional[str] = None) -> None:
        """Kernel timing with HIP events; `only` restricts it to one launch
        group name (each event pair costs the stream a few microseconds)."""
        L.call("skema_ctx_set_profiling_only", self.h, only.encode() if only else None)
        L.call("skema_ctx_set_profiling", self.h, int(on))

    def kernel_stats(self, name: str) -> Tuple[int, float]:
        cnt, ms = C.c_int64(), C.c_double()
        L.call("skema_ctx_kernel_stats", self.h, name.encode(), C.byref(cnt), C.byref(ms))
        return cnt.value, ms.value

    def reset_stats(self) -> None:
        L.call("skema_ctx_reset_stats", self.h)

    def close(self) -> None:
        if self.h:
            L.call("skema_ctx_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Optional[Context] = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


# --------------------------------------------------------------------------
# DimRedux maps
# --------------------------------------------------------------------------
class _DimRedux:
    """DimRedux<Derived> (src/Skema_DimRedux.hpp:20-93): a stored (nrow x
    ncol) map, init_transposed meaning it represents the transpose."""

    issparse_: bool = False

    def __init__(self, ctx: Context, handle: C.c_void_p, nrow: int, ncol: int, init_transposed: bool):
        self.ctx, self.h = ctx, handle
        self.nrow, self.ncol = nrow, ncol
        self.init_transposed = init_transposed

    def nrows(self) -> int:
        return self.nrow

    def ncols(self) -> int:
        return self.ncol

    def issparse(self) -> bool:
        return self.issparse_

    def _out(self, rows, cols, like):
        return torch.zeros((cols, rows), dtype=torch.float64, device=like.device).t()

    def lmap(self, alpha: float, B, beta: float = 0.0, transA: str = "N", transB: str = "N",
             idx: Tuple[int, int] = (0, 0), C_out=None):
        """DimRedux::lmap (src/Skema_DimRedux.hpp:60-69): alpha*op(map[:,idx])*op(B)."""
        wc = idx[1] - idx[0] if idx[0] != idx[1] else self.ncol
        eff = transA.upper()
        if self.init_transposed:
            eff = "T" if eff == "N" else "N"
        M = self.nrow if eff == "N" else wc
        N = B.shape[0] if (not self.issparse_ and transB.upper() == "T") else B.shape[1]
        if C_out is None:
            C_out = self._out(M, N, B)
        L.call("skema_map_lmap", self.h, float(alpha), _ptr(B), B.shape[0], B.shape[1], _ld(B), float(beta),
               transA.encode(), transB.encode(), idx[0], idx[1], _ptr(C_out), _ld(C_out))
        return C_out

    def rmap(self, alpha: float, A, beta: float = 0.0, transA: str = "N", transB: str = "T",
             idx: Tuple[int, int] = (0, 0), C_out=None):
        """DimRedux::rmap (src/Skema_DimRedux.hpp:71-80): alpha*op(A)*op(map)."""
        M = A.shape[1] if (not self.issparse_ and transA.upper() == "T") else A.shape[0]
        N = self.nrow if transB.upper() == "T" else self.ncol
        if C_out is None:
            C_out = self._out(M, N, A)
        L.call("skema_map_rmap", self.h, float(alpha), _ptr(A), A.shape[0], A.shape[1], _ld(A), float(beta),
               transA.encode(), transB.encode(), idx[0], idx[1], _ptr(C_out), _ld(C_out))
        return C_out

    def download(self):
        if not self.issparse_:
            out = np.empty((self.ncol, self.nrow), dtype=np.float64)
            L.call("skema_map_download", self.h, out.ctypes.data, None, None)
            return out.T.copy()
        nnz = C.c_int64()
        L.call("skema_map_info", self.h, None, None, C.byref(nnz), None, None)
        rp = np.empty(self.nrow + 1, dtype=np.int64)
        ci = np.empty(max(nnz.value, 1), dtype=np.int32)
        v = np.empty(max(nnz.value, 1), dtype=np.float64)
        L.call("skema_map_download", self.h, v.ctypes.data, rp.ctypes.data, ci.ctypes.data)
        return rp, ci[: nnz.value], v[: nnz.value]

    def close(self):
        if self.h:
            L.call("skema_map_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GaussDimRedux(_DimRedux):
    """GaussDimRedux (src/Skema_DimRedux.hpp:95-162).  Built from a seed on the
    device (counter-based; the reference's U(-a,a) by default, N(0,1) with
    dist="normal"), or from a materialised host array ``data``."""

    issparse_ = False

    def __init__(self, nrow: int, ncol: int, seed: int = 0, init_transposed: bool = False, *,
                 data: Optional[np.ndarray] = None, dist: str = "uniform", ctx: Optional[Context] = None):
        ctx = ctx or default_context()
        h = C.c_void_p()
        if data is not None:
            a = np.asfortranarray(np.asarray(data, dtype=np.float64))
            if a.shape != (nrow, ncol):
                raise ValueError(f"data shape {a.shape} != ({nrow}, {ncol})")
            L.call("skema_map_dense", ctx.h, nrow, ncol, a.ctypes.data, nrow, int(init_transposed), C.byref(h))
        else:
            d = L.DIST_NORMAL if dist == "normal" else L.DIST_UNIFORM
            L.call("skema_map_gauss_seeded", ctx.h, nrow, ncol, C.c_uint64(seed), d, int(init_transposed), C.byref(h))
        super().__init__(ctx, h, nrow, ncol, init_transposed)


class SparseSignDimRedux(_DimRedux):
    """SparseSignDimRedux (src/Skema_DimRedux.hpp:176-308).  Built from a seed
    (convention "rows": the reference's zeta per stored row, with replacement,
    +-1; "cols": zeta distinct rows per column of the map, +-1/sqrt(zeta)), or
    from a host CSR given as ``csr=(rowptr, col, val)``."""

    issparse_ = True

    def __init__(self, nrow: int, ncol: int, seed: int = 0, init_transposed: bool = False, *,
                 csr=None, convention: str = "rows", zeta: int = 0, ctx: Optional[Context] = None):
        ctx = ctx or default_context()
        h = C.c_void_p()
        if csr is not None:
            rp, ci, v = (np.ascontiguousarray(csr[0], dtype=np.int64), np.ascontiguousarray(csr[1], dtype=np.int32),
                         np.ascontiguousarray(csr[2], dtype=np.float64))
            nnz = int(rp[-1] - rp[0])
            L.call("skema_map_csr", ctx.h, nrow, ncol, nnz, rp.ctypes.data, ci.ctypes.data if nnz else None,
                   v.ctypes.data if nnz else None, int(init_transposed), C.byref(h))
        else:
            conv = L.SSIGN_ROWS if convention == "rows" else L.SSIGN_COLS
            L.call("skema_map_sparse_sign_seeded", ctx.h, nrow, ncol, C.c_uint64(seed), conv, zeta,
                   int(init_transposed), C.byref(h))
        super().__init__(ctx, h, nrow, ncol, init_transposed)


# --------------------------------------------------------------------------
# Windows
# --------------------------------------------------------------------------
@dataclass
class CrsMatrix:
    """Device CSR (KokkosSparse CrsMatrix<double,int,...,size_t> analogue,
    src/Skema_Utils.hpp:24-25): rowptr int64 (m+1), col int32, val float64."""

    m: int
    n: int
    rowptr: "torch.Tensor"
    col: "torch.Tensor"
    val: "torch.Tensor"

    @staticmethod
    def from_scipy(A, device="cuda"):
        A = A.tocsr()
        A.sort_indices()
        return CrsMatrix(A.shape[0], A.shape[1], torch.as_tensor(A.indptr.astype(np.int64), device=device),
                         torch.as_tensor(A.indices.astype(np.int32), device=device),
                         torch.as_tensor(A.data.astype(np.float64), device=device))


def gauss_rbf_window(points, gamma: float, idx: Tuple[int, int], ctx: Optional[Context] = None):
    """GaussRBFWindow::get / GaussRBF::compute (src/Skema_Window.hpp:90-107,
    src/Skema_Kernel.cpp:5-81): the materialised K[idx, :] (rows x n)."""
    ctx = ctx or default_context()
    n, d = points.shape
    rows = idx[1] - idx[0]
    K = torch.empty((n, rows), dtype=torch.float64, device=points.device).t()
    L.call("skema_rbf_window", ctx.h, _ptr(points), n, d, _ld(points), float(gamma), idx[0], rows, _ptr(K), _ld(K))
    return K


# --------------------------------------------------------------------------
# SketchySVD
# --------------------------------------------------------------------------
@dataclass
class AlgParams:
    """Sketch fields of AlgParams (src/Skema_AlgParams.hpp:12-90)."""

    matrix_m: int = 0
    matrix_n: int = 0
    rank: int = 1
    sketch_range: int = 1
    sketch_core: int = 1
    window: int = 1
    seeds: Sequence[int] = field(default_factory=lambda: [0, 1, 2, 3])
    sketch_eta: float = 1.0
    sketch_nu: float = 1.0
    dim_redux: str = "gauss"
    issparse: bool = False
    kernel_func: str = "none"
    kernel_gamma: float = 1.0
    nfeat: int = 0
    precision: str = "fp64"

    def sizes(self) -> Tuple[int, int]:
        k = 4 * self.rank + 1 if self.sketch_range < self.rank else self.sketch_range
        s = 2 * k + 1 if self.sketch_core < self.rank else self.sketch_core
        return k, s

    def to_c(self) -> L.skema_params:
        p = L.skema_params()
        L.load().skema_params_default(C.byref(p))
        p.m, p.n, p.rank = self.matrix_m, self.matrix_n, self.rank
        p.range, p.core, p.window = self.sketch_range, self.sketch_core, self.window
        p.eta, p.nu = self.sketch_eta, self.sketch_nu
        p.dim_redux = L.MAP_GAUSS if self.dim_redux == "gauss" else L.MAP_SPARSE_SIGN
        p.issparse = int(self.issparse)
        p.kernel = L.KERNEL_GAUSSRBF if self.kernel_func == "gaussrbf" else L.KERNEL_NONE
        p.precision = L.PRECISIONS[self.precision]
        p.gamma = self.kernel_gamma
        p.nfeat = self.nfeat
        return p


def map_shapes(p: AlgParams):
    """Stored shapes / init_transposed of the four maps, as the reference's
    ctor builds them (src/Skema_SketchySVD.cpp:30-95)."""
    k, s = p.sizes()
    m, n = p.matrix_m, p.matrix_n
    g = p.dim_redux == "gauss"
    tU, tO = p.issparse and g, p.issparse
    return {
        "Upsilon": ((m, k) if tU else (k, m), tU),
        "Omega": ((n, k) if tO else (k, n), tO),
        "Phi": ((m, s) if tU else (s, m), tU),
        "Psi": ((n, s) if tO else (s, n), tO),
    }


def make_maps(p: AlgParams, ctx: Optional[Context] = None, convention: str = "rows", dist: str = "uniform"):
    """Seeded maps (seeds[0..3]) with the reference's orientation; Gauss maps
    U(-a, a) like the reference's fill, N(0, 1) with dist="normal"."""
    ctx = ctx or default_context()
    out = {}
    for name, seed in zip(("Upsilon", "Omega", "Phi", "Psi"), p.seeds):
        (r, c), t = map_shapes(p)[name]
        if p.dim_redux == "gauss":
            out[name] = GaussDimRedux(r, c, seed, t, dist=dist, ctx=ctx)
        else:
            out[name] = SparseSignDimRedux(r, c, seed, t, convention=convention, ctx=ctx)
    return out


class SketchySVD:
    """SketchySVD<M,D> (src/Skema_SketchySVD.hpp:11-64) on one GPU."""

    def __init__(self, params: AlgParams, maps=None, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.params = params
        self.maps = maps if maps is not None else make_maps(params, self.ctx)
        self.rank = params.rank
        self.nrow, self.ncol = params.matrix_m, params.matrix_n
        self.range, self.core = params.sizes()
        h = C.c_void_p()
        cp = params.to_c()
        L.call("skema_sketch_create", self.ctx.h, C.byref(cp), self.maps["Upsilon"].h, self.maps["Omega"].h,
               self.maps["Phi"].h, self.maps["Psi"].h, C.byref(h))
        self.h = h

    # ---- update ---------------------------------------------------------
    def linear_update(self, A) -> None:
        """src/Skema_SketchySVD.cpp:119-246 (resets X, Y, Z first)."""
        if isinstance(A, CrsMatrix):
            L.call("skema_linear_update_csr", self.h, _ptr(A.rowptr), _ptr(A.col), _ptr(A.val))
        elif self.params.kernel_func == "gaussrbf":
            L.call("skema_linear_update_rbf", self.h, _ptr(A), _ld(A))
        else:
            L.call("skema_linear_update_dense", self.h, _ptr(A), _ld(A))

    def update(self, A, idx: Tuple[int, int]) -> None:
        """One window of rows idx: update(H, idx) + axpy (SketchySVD.cpp:176-189).
        A is the whole matrix / point set; the window is taken from it."""
        r0, rows = idx[0], idx[1] - idx[0]
        if isinstance(A, CrsMatrix):
            L.call("skema_update_csr", self.h, _ptr(A.rowptr) + 8 * r0, _ptr(A.col), _ptr(A.val), r0, rows)
        elif self.params.kernel_func == "gaussrbf":
            L.call("skema_update_rbf", self.h, _ptr(A), _ld(A), r0, rows)
        else:
            L.call("skema_update_dense", self.h, _ptr(A) + 8 * r0, _ld(A), r0, rows)

    def reset(self) -> None:
        L.call("skema_sketch_reset", self.h)

    # ---- state ----------------------------------------------------------
    def state(self):
        k, s, m, n = self.range, self.core, self.nrow, self.ncol
        X = np.empty((n, k)); Y = np.empty((k, m)); Z = np.empty((s, s))
        L.call("skema_get_state", self.h, X.ctypes.data, Y.ctypes.data, Z.ctypes.data)
        return X.T.copy(), Y.T.copy(), Z.T.copy()

    def set_state(self, X, Y, Z) -> None:
        X = np.asfortranarray(X, dtype=np.float64)
        Y = np.asfortranarray(Y, dtype=np.float64)
        Z = np.asfortranarray(Z, dtype=np.float64)
        L.call("skema_set_state", self.h, X.ctypes.data, Y.ctypes.data, Z.ctypes.data)

    # ---- finalize -------------------------------------------------------
    def low_rank_approx(self, device: bool = False):
        """src/Skema_SketchySVD.cpp:568-664 -> (U m x r, S r, V n x r)."""
        m, n, r = self.nrow, self.ncol, self.rank
        if device:
            U = torch.empty((r, m), dtype=torch.float64, device="cuda").t()
            V = torch.empty((r, n), dtype=torch.float64, device="cuda").t()
            S = torch.empty(r, dtype=torch.float64, device="cuda")
            L.call("skema_low_rank_approx_device", self.h, _ptr(U), _ld(U), _ptr(S), _ptr(V), _ld(V))
            return U, S, V
        U = np.empty((r, m)); S = np.empty(r); V = np.empty((r, n))
        L.call("skema_low_rank_approx", self.h, U.ctypes.data, S.ctypes.data, V.ctypes.data)
        return U.T.copy(), S, V.T.copy()

    def compute_residuals(self, A, U=None, S=None, V=None) -> np.ndarray:
        """SketchySVD::compute_residuals (src/Skema_SketchySVD.cpp:718-727):
        per-triplet residual norms of (U, S, V), by default the factors of a
        fresh low_rank_approx, against the input A (dense, CrsMatrix, or the
        point set of an RBF sketch)."""
        if U is None:
            U, S, V = self.low_rank_approx(device=True)
        if self.params.kernel_func == "gaussrbf":
            self.rnrms = residual_norms_rbf(A, self.params.kernel_gamma, U, S, V, window=self.params.window,
                                            ctx=self.ctx)
        else:
            self.rnrms = residual_norms(A, U, S, V, ctx=self.ctx)
        return self.rnrms

    def checkpoint(self, name: str) -> np.ndarray:
        rows, cols = C.c_int64(), C.c_int64()
        L.call("skema_get_checkpoint", self.h, name.encode(), None, C.byref(rows), C.byref(cols))
        out = np.empty((cols.value, rows.value))
        L.call("skema_get_checkpoint", self.h, name.encode(), out.ctypes.data, None, None)
        return out.T.copy()

    def timings(self) -> dict:
        buf = C.create_string_buffer(8192)
        L.call("skema_sketch_timings_json", self.h, buf, len(buf))
        return json.loads(buf.value.decode())

    def close(self) -> None:
        if self.h:
            L.call("skema_sketch_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class SketchySPD:
    """SketchySPD<M,D> (src/Skema_SketchySVD.cpp:757-1151) on one GPU: the
    Nystrom sketch Y = A Omega of a symmetric PSD matrix (m == n) and its
    fixed-rank approximation (U, lambda).  Every window is sketched (the
    reference's loop stops after window 0, :795-858); ``shift_removal="eps"``
    removes the eps ||Y||_2 shift, ``"reference"`` subtracts ||Y||_2 as the
    reference does (:1141-1145)."""

    SHIFTS = {"eps": 0, "reference": 1}

    def __init__(self, params: AlgParams, omega: Optional[_DimRedux] = None, ctx: Optional[Context] = None):
        self.ctx = ctx or default_context()
        self.params = params
        self.rank = params.rank
        self.n = params.matrix_n
        self.range = 4 * params.rank + 1 if params.sketch_range < params.rank else params.sketch_range
        if omega is None:  # DimReduxT(ncol, range, seeds[0]) (:768-773)
            if params.dim_redux == "gauss":
                omega = GaussDimRedux(self.n, self.range, seed=params.seeds[0], ctx=self.ctx)
            else:
                omega = SparseSignDimRedux(self.n, self.range, seed=params.seeds[0], ctx=self.ctx)
        self.omega = omega
        h = C.c_void_p()
        cp = params.to_c()
        L.call("skema_spd_create", self.ctx.h, C.byref(cp), omega.h, C.byref(h))
        self.h = h

    def linear_update(self, A) -> None:
        """src/Skema_SketchySVD.cpp:790-880 (zeroes Y first)."""
        if isinstance(A, CrsMatrix):
            L.call("skema_spd_linear_update_csr", self.h, _ptr(A.rowptr), _ptr(A.col), _ptr(A.val))
        elif self.params.kernel_func == "gaussrbf":
            L.call("skema_spd_linear_update_rbf", self.h, _ptr(A), _ld(A))
        else:
            L.call("skema_spd_linear_update_dense", self.h, _ptr(A), _ld(A))

    def update(self, A, idx: Tuple[int, int]) -> None:
        r0, rows = idx[0], idx[1] - idx[0]
        if isinstance(A, CrsMatrix):
            L.call("skema_spd_update_csr", self.h, _ptr(A.rowptr) + 8 * r0, _ptr(A.col), _ptr(A.val), r0, rows)
        elif self.params.kernel_func == "gaussrbf":
            L.call("skema_spd_update_rbf", self.h, _ptr(A), _ld(A), r0, rows)
        else:
            L.call("skema_spd_update_dense", self.h, _ptr(A) + 8 * r0, _ld(A), r0, rows)

    def state(self) -> np.ndarray:
        Y = np.empty((self.range, self.params.matrix_m))
        L.call("skema_spd_get_state", self.h, Y.ctypes.data)
        return Y.T.copy()

    def low_rank_approx(self, device: bool = False, shift_removal: str = "eps"):
        """src/Skema_SketchySVD.cpp:934-1151 -> (U n x r, lambda r)."""
        n, r = self.n, self.rank
        sh = self.SHIFTS[shift_removal]
        if device:
            U = torch.empty((r, n), dtype=torch.float64, device="cuda").t()
            S = torch.empty(r, dtype=torch.float64, device="cuda")
            L.call("skema_spd_low_rank_approx_device", self.h, sh, _ptr(U), _ld(U), _ptr(S))
            return U, S
        U = np.empty((r, n)); S = np.empty(r)
        L.call("skema_spd_low_rank_approx", self.h, sh, U.ctypes.data, S.ctypes.data)
        return U.T.copy(), S

    def compute_residuals(self, A, U=None, S=None) -> np.ndarray:
        """Eigenpair residual norms (residuals(A, V, S, ...),
        src/Skema_Residuals.hpp:123-158) of a dense or CSR A or an RBF point set."""
        if U is None:
            U, S = self.low_rank_approx(device=True)
        gamma = self.params.kernel_gamma if self.params.kernel_func == "gaussrbf" else None
        self.rnrms = residual_norms_sym(A, U, S, window=self.params.window, gamma=gamma, ctx=self.ctx)
        return self.rnrms

    def close(self) -> None:
        if self.h:
            L.call("skema_spd_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def residual_fro(A, U, S, V, window: int = 4096, ctx: Optional[Context] = None) -> Tuple[float, float]:
    """(||A - U diag(S) V^T||_F^2, ||A||_F^2) streamed over row windows (device
    tensors, column-major)."""
    ctx = ctx or default_context()
    m, n = A.shape
    r = U.shape[1]
    e2, a2 = C.c_double(), C.c_double()
    L.call("skema_residual_fro_dense", ctx.h, m, n, r, _ptr(A), _ld(A), _ptr(U), _ld(U), _ptr(S), _ptr(V), _ld(V),
           window, C.byref(e2), C.byref(a2))
    return e2.value, a2.value


def residual_fro_noalloc(A, U, S, V, window: int = 4096, ctx: Optional[Context] = None) -> Tuple[float, float]:
    """residual_fro through the entry point that allocates no device memory
    (its scratch is the library's static per-GPU array; calls on one GPU must
    not overlap)."""
    ctx = ctx or default_context()
    m, n = A.shape
    r = U.shape[1]
    e2, a2 = C.c_double(), C.c_double()
    L.call("skema_residual_fro_dense_noalloc", ctx.h, m, n, r, _ptr(A), _ld(A), _ptr(U), _ld(U), _ptr(S), _ptr(V),
           _ld(V), window, C.byref(e2), C.byref(a2))
    return e2.value, a2.value


def residual_norms(A, U, S, V, ctx: Optional[Context] = None) -> np.ndarray:
    """Per-triplet residual norms ||[A^T u_i - s_i v_i ; A v_i - s_i u_i]||_2
    (residuals, src/Skema_Residuals.hpp:160-217) for a dense column-major
    device A or a CrsMatrix; U, S, V device tensors (m x r, r, n x r)."""
    ctx = ctx or default_context()
    r = U.shape[1]
    out = np.empty(r)
    if isinstance(A, CrsMatrix):
        m, n = A.m, A.n
        L.call("skema_residuals_csr", ctx.h, m, n, r, _ptr(A.rowptr), _ptr(A.col), _ptr(A.val), _ptr(U), _ld(U),
               _ptr(S), _ptr(V), _ld(V), out.ctypes.data)
    else:
        m, n = A.shape
        L.call("skema_residuals_dense", ctx.h, m, n, r, _ptr(A), _ld(A), _ptr(U), _ld(U), _ptr(S), _ptr(V), _ld(V),
               out.ctypes.data)
    return out


def residual_norms_rbf(points, gamma: float, U, S, V, window: int = 4096,
                       ctx: Optional[Context] = None) -> np.ndarray:
    """The same for A = the Gaussian RBF matrix of the device point set,
    generated window by window (the GaussRBFWindow path of map_A_by_window,
    src/Skema_Residuals.hpp:60-121)."""
    ctx = ctx or default_context()
    n, d = points.shape
    r = U.shape[1]
    out = np.empty(r)
    L.call("skema_residuals_rbf", ctx.h, _ptr(points), n, d, _ld(points), float(gamma), r, _ptr(U), _ld(U), _ptr(S),
           _ptr(V), _ld(V), window, out.ctypes.data)
    return out


def residual_norms_sym(A, V, S, window: int = 4096, gamma: Optional[float] = None,
                       ctx: Optional[Context] = None) -> np.ndarray:
    """Eigenpair residuals ||A v_i - s_i v_i||_2 (src/Skema_Residuals.hpp:123-158):
    A a dense symmetric device matrix, a CrsMatrix, or (gamma given) the RBF
    matrix of the device point set A."""
    ctx = ctx or default_context()
    r = V.shape[1]
    out = np.empty(r)
    if isinstance(A, CrsMatrix):
        L.call("skema_residuals_sym_csr", ctx.h, A.m, r, _ptr(A.rowptr), _ptr(A.col), _ptr(A.val), _ptr(V), _ld(V),
               _ptr(S), out.ctypes.data)
    elif gamma is None:
        n = A.shape[0]
        L.call("skema_residuals_sym_dense", ctx.h, n, r, _ptr(A), _ld(A), _ptr(V), _ld(V), _ptr(S), out.ctypes.data)
    else:
        n, d = A.shape
        L.call("skema_residuals_sym_rbf", ctx.h, _ptr(A), n, d, _ld(A), float(gamma), r, _ptr(V), _ld(V), _ptr(S),
               window, out.ctypes.data)
    return out


def rbf_sketch_loss_grad(points, gamma: float, omega: _DimRedux, phi: _DimRedux, psi: _DimRedux,
                         window: int = 4096, precision: str = "fp64", grad: bool = True,
                         ctx: Optional[Context] = None):
    """L = ||K Omega^T||_F^2 + ||Phi K Psi^T||_F^2 for the Gaussian RBF kernel
    K of the n x d device point set, with dL/dX (n x d, column-major device
    tensor) and dL/dgamma, kernel tiles recomputed in the backward pass."""
    ctx = ctx or default_context()
    n, d = points.shape
    loss, dg = C.c_double(), C.c_double()
    dX = torch.empty((d, n), dtype=torch.float64, device=points.device).t() if grad else None
    L.call("skema_rbf_sketch_loss_grad", ctx.h, _ptr(points), n, d, _ld(points), float(gamma), omega.h, phi.h, psi.h,
           window, L.PRECISIONS[precision], C.byref(loss), _ptr(dX) if grad else None, _ld(dX) if grad else n,
           C.byref(dg))
    return loss.value, dX, dg.value
