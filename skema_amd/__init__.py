"""skema_amd -- MI355X-native streaming SketchySVD (update + fixed-rank finalize).

The compute lives in ``libskema_mi355x.so`` (hand-written gfx950 HIP kernels
behind the C ABI in ``include/skema_mi355x.h``); this package is the
reference-shaped host layer over it.
"""
from ._lib import SkemaError, load  # noqa: F401
from .sketchy import (  # noqa: F401
    AlgParams,
    Context,
    CrsMatrix,
    GaussDimRedux,
    SketchySPD,
    SketchySVD,
    SparseSignDimRedux,
    colmajor,
    default_context,
    gauss_rbf_window,
    make_maps,
    map_shapes,
    rbf_sketch_loss_grad,
    residual_fro,
    residual_fro_noalloc,
    residual_norms,
    residual_norms_rbf,
    residual_norms_sym,
)

__version__ = "1.0.0"
