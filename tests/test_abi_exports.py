"""The C-ABI library loads and exports every symbol include/skema_mi355x.h
declares (no GPU call is made)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "skema_mi355x.h")
LIB = os.path.join(ROOT, "skema_amd", "libskema_mi355x.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(skema_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_surface():
    names = declared_functions()
    for must in ["skema_ctx_create", "skema_map_lmap", "skema_map_rmap", "skema_sketch_create",
                 "skema_update_dense", "skema_update_csr", "skema_update_rbf", "skema_linear_update_dense",
                 "skema_low_rank_approx", "skema_get_state", "skema_rbf_sketch_loss_grad"]:
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_python_binding_covers_the_header():
    from skema_amd import _lib

    assert set(_lib.SIGNATURES) == set(declared_functions())
    lib = _lib.load()
    assert lib.skema_version() >= 10000
    p = _lib.skema_params()
    lib.skema_params_default(ctypes.byref(p))
    assert (p.rank, p.window, p.eta, p.nu, p.gamma) == (1, 1, 1.0, 1.0, 1.0)
