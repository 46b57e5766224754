"""The command-line driver's sketch-route flags on the GPU: --normalize-matrix,
--sketch-compute-svals-iters (history traces), --debug-file dumps,
--print-level banners, the SketchySPD history timings and the reference's
U(-a, a) Gauss map law as the default.

Exact recovery (range = min(m, n)) on the reference's fixtures makes the
expected singular values known exactly (tests/golden/*_svals.txt)."""
import json
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "skema_amd", "skema_sketch")

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not os.path.exists(EXE), reason="driver not built")]


def _run(args, cwd, timeout=120):
    print("case cli", " ".join(args), flush=True)
    r = subprocess.run([EXE, *args], capture_output=True, text=True, timeout=timeout, cwd=cwd)
    assert r.returncode == 0, r.stderr
    return r


RECT = ["--rank", "10", "--range", "14", "--core", "29", "--window", "5"]


def test_normalize_matrix_scales_the_input(golden_dir, tmp_path):
    """--normalize-matrix x divides A by x (driver/main.cpp:110-114,
    :139-141): dense and CSR input give the stored singular values / x."""
    sv = np.loadtxt(f"{golden_dir}/rectangular_svals.txt")[:10]
    for extra in ([], ["--sparse"]):
        hist = tmp_path / "h.json"
        _run(["--input", f"{golden_dir}/rectangular_sparse.mtx", *extra, *RECT, "--normalize-matrix", "4",
              "--history-file", str(hist)], tmp_path)
        h = json.loads(hist.read_text())
        assert np.max(np.abs(np.array(h["svals"]) - sv / 4.0)) <= 1e-12 * sv[0], (extra, h["svals"][:3])


def test_svals_traces_per_window(golden_dir, tmp_path):
    """--sketch-compute-svals-iters: one traces entry per window ("0", "1",
    ...) with the rank-r singular values of the sketch so far; the last one
    is the final answer (src/Skema_SketchySVD.cpp:208-216)."""
    sv = np.loadtxt(f"{golden_dir}/rectangular_svals.txt")[:10]
    hist = tmp_path / "h.json"
    _run(["--input", f"{golden_dir}/rectangular_sparse.mtx", *RECT, "--sketch-compute-svals-iters",
          "--history-file", str(hist)], tmp_path)
    h = json.loads(hist.read_text())
    windows = math.ceil(32 / 5)
    assert sorted(h["traces"], key=int) == [str(i) for i in range(windows)]
    for key, t in h["traces"].items():
        assert len(t["svals"]) == 10
    last = np.array(h["traces"][str(windows - 1)]["svals"])
    assert np.max(np.abs(last - np.array(h["svals"]))) <= 1e-13 * sv[0]
    assert np.max(np.abs(last - sv)) <= 1e-12 * sv[0]
    # a partial sketch differs from the final one
    assert np.max(np.abs(np.array(h["traces"]["0"]["svals"]) - sv)) > 1e-6 * sv[0]
    # without the flag the history still has the (empty) traces key
    _run(["--input", f"{golden_dir}/rectangular_sparse.mtx", *RECT, "--history-file", str(hist)], tmp_path)
    assert json.loads(hist.read_text())["traces"] == {}


def test_debug_file_dumps(golden_dir, tmp_path):
    """--debug-file dbg.txt writes dbg_<name>.txt in the working directory:
    the maps, X/Y/Z after every window and at the end, the finalize
    intermediates and U, S, V, as "%.16f" text with one matrix row per line.
    The per-window path matches the whole-matrix path."""
    A = np.loadtxt(f"{golden_dir}/rectangular_dense.txt")
    sv = np.loadtxt(f"{golden_dir}/rectangular_svals.txt")[:10]
    hist = tmp_path / "h.json"
    _run(["--input", f"{golden_dir}/rectangular_sparse.mtx", *RECT, "--debug-file", "/somewhere/dbg.txt",
          "--history-file", str(hist)], tmp_path)
    m, n, k, s = 32, 14, 14, 29
    shapes = {"Upsilon.txt": (k, m), "Omega.txt": (k, n), "Phi.txt": (s, m), "Psi.txt": (s, n),
              "X.txt": (k, n), "Y.txt": (m, k), "Z.txt": (s, s), "X.txt.0": (k, n), "X.txt.6": (k, n), "Y.txt.6": (m, k),
              "P.txt": (n, k), "Q.txt": (m, k), "PhiQ.txt": (s, k), "PsiP.txt": (s, k), "T1.txt": (k, k),
              "T2.txt": (k, k), "U.txt": (m, 10), "V.txt": (n, 10)}
    mats = {}
    for name, shape in shapes.items():
        f = tmp_path / f"dbg_{name}"
        assert f.exists(), name
        mats[name] = np.loadtxt(f, ndmin=2)
        assert mats[name].shape == shape, (name, mats[name].shape)
    S = np.loadtxt(tmp_path / "dbg_S.txt")
    assert np.max(np.abs(S - sv)) <= 1e-12 * sv[0]
    # the dumped maps and the final state agree: X = Upsilon A (16 fixed decimals)
    X = mats["Upsilon.txt"] @ A
    assert np.max(np.abs(mats["X.txt"] - X)) <= 1e-9 * np.max(np.abs(X))
    assert np.max(np.abs(mats["X.txt.6"] - mats["X.txt"])) == 0.0
    assert np.max(np.abs(mats["Y.txt"] - A @ mats["Omega.txt"].T)) <= 1e-9 * np.max(np.abs(mats["Y.txt"]))
    U, V = mats["U.txt"], mats["V.txt"]
    assert np.linalg.norm((U * S) @ V.T - A) <= 1e-9 * np.linalg.norm(A)
    # the sparse-sign maps are written as MatrixMarket coordinate files
    _run(["--input", f"{golden_dir}/rectangular_sparse.mtx", "--sparse", *RECT, "--model", "sparse-sign",
          "--debug-file", "sp.txt", "--history-file", str(hist)], tmp_path)
    from scipy.io import mmread

    Ups = mmread(str(tmp_path / "sp_Upsilon.mtx")).tocoo()
    assert Ups.shape == (k, m) and Ups.nnz == 8 * k and set(np.unique(Ups.data)) <= {-1.0, 1.0}


def test_print_level_and_default_map_law(golden_dir, tmp_path):
    """--print-level 1 prints the options banner and the finalize phases.
    The seeded Gauss maps are U(-a, a), a = sqrt(2 ln(rows * cols)), like the
    reference's fill (src/Skema_DimRedux.hpp:112-115); --map-dist normal
    selects N(0, 1)."""
    r = _run(["--input", f"{golden_dir}/rectangular_sparse.mtx", *RECT, "--print-level", "1",
              "--debug-file", "u.txt"], tmp_path)
    assert "==================== Skema ====================" in r.stdout
    assert "SketchSVD::initial_approx: [P,~] = qr(X^T,0)" in r.stdout
    U = np.loadtxt(tmp_path / "u_Upsilon.txt")
    a = math.sqrt(2.0 * math.log(14 * 32))
    assert np.max(np.abs(U)) <= a and np.max(np.abs(U)) > 0.9 * a
    assert abs(np.mean(np.abs(U)) - a / 2) < 0.1 * a  # E|U(-a,a)| = a/2
    _run(["--input", f"{golden_dir}/rectangular_sparse.mtx", *RECT, "--map-dist", "normal", "--debug-file",
          "g.txt"], tmp_path)
    G = np.loadtxt(tmp_path / "g_Upsilon.txt")
    assert abs(np.mean(np.abs(G)) - math.sqrt(2 / math.pi)) < 0.1  # E|N(0,1)| = 0.798


def test_spd_history_has_timings(golden_dir, tmp_path):
    """The SketchySPD route's history carries the reference's SketchySPD
    timing keys (src/Skema_SketchySVD.cpp:776-786)."""
    hist = tmp_path / "h.json"
    _run(["--input", f"{golden_dir}/symmetric_sparse.mtx", "--symmetric", "--rank", "27", "--range", "27",
          "--window", "5", "--history-file", str(hist)], tmp_path)
    h = json.loads(hist.read_text())
    t = h["timings"]
    assert set(t["init"]) >= {"omega"}
    assert set(t["update"]) >= {"omega", "window", "daxpy"}
    assert set(t["approx"]) >= {"norm2", "daxpy", "omega", "update", "dpotrf", "dgels", "dgesvd"}
    assert t["update"]["omega"] > 0.0 and t["approx"]["dgesvd"] > 0.0
    sv = np.loadtxt(f"{golden_dir}/symmetric_svals.txt")
    assert np.max(np.abs(np.array(h["svals"]) - sv)) <= 1e-12 * sv[0]
