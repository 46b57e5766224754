#!/usr/bin/env python3
"""rows/s of the streaming SketchySVD window update on one MI355X.

Workload (default): dense A = U_p diag(sigma) V_p^T + E, 65536 x 16384,
sigma_i = 0.95^i (p = 720 terms, below 1e-16*sigma_0 after that), U_p, V_p
Haar (QR of N(0,1) with the R-diagonal sign fix, seeds 6, 7), E ~ N(0, 1e-4)
read as a variance (seed 8); rank 100, range 401, core 803, windows of 4096
rows; Gaussian N(0,1) maps from seeds 0..3.  A is resident in HBM before the
timed region.  One step = linear_update over all 65536 rows.

Forms timed: the plain FP64 path and every reduced-precision form the library
supports; `value` is the fastest form whose rank-r relative Frobenius error is
within 1.05x of the FP64 path's (the FP64 path itself if none is).  The
roofline object covers the dominant kernel of that form, timed with HIP
events recorded by the library on the launch stream during the timed steps.
The CPU baseline is the oracle (numpy + LAPACK/BLAS on all host threads) on a
bounded sample of the same windows, same maps.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "rows/s of SketchySVD window update, mixed vs FP64 plain at equal rel. Frobenius err"

WORKLOADS = {
    "dense-65536x16384": dict(m=65536, n=16384, rank=100, k=401, s=803, window=4096, p=720, decay=0.95,
                              noise_var=1e-4, seeds=(6, 7, 8)),
    # The same matrix with the noise level read as a standard deviation
    # (variance 1e-8): signal and noise energies are comparable, so the
    # reduced forms' rounding shows in the rank-r error.
    "dense-65536x16384-lownoise": dict(m=65536, n=16384, rank=100, k=401, s=803, window=4096, p=720, decay=0.95,
                                       noise_var=1e-8, seeds=(6, 7, 8)),
    "dense-2000x1000": dict(m=2000, n=1000, rank=10, k=41, s=83, window=100, p=1000, decay=0.8,
                            noise_var=1e-6, seeds=(1, 2, 3)),
    # CSR A with density 1e-3 (uniform positions, duplicates merged, N(0,1)
    # values, seed 4); sparse-sign maps, zeta = 8 distinct rows per column.
    "sparse-200000x20000": dict(m=200000, n=20000, rank=25, k=101, s=203, window=10000, density=1e-3, seed=4,
                                zeta=8),
    # Gaussian-RBF kernel of n = 1e5 points in d = 128 (mixture of 10
    # Gaussians, means N(0, 4I), covariance I, FP32 draws widened to FP64,
    # seed 5), gamma = 1/d.
    "rbf-100000x128": dict(m=100000, n=100000, d=128, rank=50, k=201, s=403, window=2048, clusters=10, seed=5),
    # Loss L = ||K Omega^T||^2 + ||Phi K Psi^T||^2 of the RBF kernel of 32768
    # points in d = 64 (same mixture, seed 9), gamma = 0.5/d, k = 201,
    # s = 403, forward windows of 4096 rows; one step = loss + dL/dX + dL/dgamma.
    "rbf-adjoint-32768x64": dict(m=32768, n=32768, d=64, k=201, s=403, window=4096, clusters=10, seed=9),
}

PEAKS_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16": 2500.0, "bf16x3": 2500.0, "fp8": 5000.0}
# MFMA operations issued per algorithmic flop (split-BF16 issues 3 products).
MFMA_PER_FLOP = {"fp64": 1, "fp32": 1, "bf16": 1, "bf16x3": 3, "fp8": 1}
DTYPE = {"fp64": "f64", "fp32": "f32", "bf16": "bf16", "bf16x3": "bf16x3", "fp8": "fp8e4m3"}


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "pmc_traffic_dense-65536x16384.txt")


def committed_traffic(form):
    """HBM bytes per launch of the dominant kernel for a form, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary (each scaled by the
    copy-kernel calibration stated there); None when no summary covers it."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                parts = line.split()
                if len(parts) >= 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == form:
                    return float(parts[2])
    except OSError:
        return None
    return None


def haar(rows, cols, seed):
    G = np.random.default_rng(seed).standard_normal((rows, cols))
    Q, R = np.linalg.qr(G)
    return Q * np.sign(np.diag(R))


def make_dense(cfg, torch):
    """A (m x n) column-major on cuda, built as (V (U sigma)^T + E^T)^T."""
    m, n, p = cfg["m"], cfg["n"], cfg["p"]
    su, sv, se = cfg["seeds"]
    sig = cfg["decay"] ** np.arange(p)
    U = torch.as_tensor(haar(m, p, su) * sig, device="cuda")
    V = torch.as_tensor(haar(n, p, sv), device="cuda")
    g = torch.Generator(device="cuda")
    g.manual_seed(se)
    At = torch.randn((n, m), dtype=torch.float64, device="cuda", generator=g)
    At.mul_(math.sqrt(cfg["noise_var"]))
    At.addmm_(V, U.t())
    del U, V
    torch.cuda.synchronize()
    return At.t()


def optimal_tail(A, r, torch, iters=8, over=100):
    """sqrt(sum_{i>r} sigma_i^2) / ||A||_F by block subspace iteration (a
    measurement aid, not the sketch path): the products with A run on the
    device (torch FP64), the orthonormalisations and the final singular
    values on the host (numpy LAPACK); ||A||^2 minus the top-r Ritz values
    squared of an (r + over)-dimensional subspace refined `iters` times."""
    n = A.shape[1]

    def orth(M):
        Q, _ = np.linalg.qr(M.cpu().numpy())
        return torch.as_tensor(Q, device="cuda")

    G = np.random.default_rng(12345).standard_normal((n, r + over))
    Q = orth(A @ torch.as_tensor(G, device="cuda"))
    for _ in range(iters):
        Z = orth(A.t() @ Q)
        Q = orth(A @ Z)
    sv = np.linalg.svd((Q.t() @ A).cpu().numpy(), compute_uv=False)
    a2 = float((A * A).sum())
    top = float((sv[:r] ** 2).sum())
    return math.sqrt(max(a2 - top, 0.0) / a2)


def host_cpu():
    """CPU model and the cores this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = os.cpu_count()
    return {"cpu_model": model, "host_cores_available": avail, "host_cpus_online": os.cpu_count()}


def one_pass_fp8(cfg, form):
    """Whether the fp8 update runs as one pass over the whole matrix (nu = 1,
    windows of whole 128-row steps, k <= 448): quantisation and left product
    in one launch, then one right and one core product over all m rows."""
    w, m = cfg["window"], cfg["m"]
    return form == "fp8" and w % 128 == 0 and w < m and cfg["k"] <= 448 and m % 2 == 0


def right_product(cfg, form):
    """(name, algorithmic flops per launch) of [y | G] = A_w [Omega; Psi]^T,
    2 w n (k + s) flops per window (2 m n (k + s) in the one-pass form)."""
    rows = cfg["m"] if one_pass_fp8(cfg, form) else cfg["window"]
    return "sketch_right", 2.0 * rows * cfg["n"] * (cfg["k"] + cfg["s"])


def dominant_kernel(cfg, form):
    """(name, algorithmic work per launch, bound) of the launch group with
    most time in a step.  fp8 with the one-pass update: the pass
    (sketch_left, one launch per update), HBM-bound: A read once (8 B an
    element), R and its scales written (1 + 1/32 B an element), x written
    (8 k n B) and Upsilon read (k m B).  Otherwise the right product, MFMA-bound."""
    if one_pass_fp8(cfg, form):
        m, n, k = cfg["m"], cfg["n"], cfg["k"]
        return "sketch_left", m * n * (8.0 + 1.0 + 1.0 / 32.0) + 8.0 * k * n + 1.0 * k * m, "hbm"
    name, flops = right_product(cfg, form)
    return name, flops, "mfma"


def flops_per_row(cfg):
    n, k, s = cfg["n"], cfg["k"], cfg["s"]
    return 2.0 * n * (2 * k + s) + 2.0 * s * s


def run_form(S, ctx, cfg, A, maps, form, steps, warmup, torch, error=True, ref_state=None, keep_state=False):
    """Times `steps` linear_update calls of one form (after `warmup`), then
    (error=True) the finalize and the streamed rank-r error.  ref_state: the
    FP64 form's (X, Y, Z), against which this form's sketches are compared."""
    p = S.AlgParams(matrix_m=cfg["m"], matrix_n=cfg["n"], rank=cfg["rank"], sketch_range=cfg["k"],
                    sketch_core=cfg["s"], window=cfg["window"], precision=form)
    sk = S.SketchySVD(p, maps=maps, ctx=ctx)
    for _ in range(warmup):
        sk.linear_update(A)
    torch.cuda.synchronize()
    # Timed region: HIP events only around the dominant kernel's launch group
    # (one pair per window; timing every group adds ~10 us of stream stalls
    # per event pair).
    name, _, _ = dominant_kernel(cfg, form)
    ctx.reset_stats()
    ctx.set_profiling(True, only=name)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        sk.linear_update(A)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.set_profiling(False)
    cnt, tot_ms = ctx.kernel_stats(name)
    ms_per_step = (t1 - t0) * 1e3 / steps
    # Per-group breakdown from one more, fully profiled step (not timed).
    ctx.reset_stats()
    ctx.set_profiling(True)
    sk.linear_update(A)
    torch.cuda.synchronize()
    ctx.set_profiling(False)
    stats = {k: ctx.kernel_stats(k) for k in ("window", "sketch_left", "sketch_right", "sketch_core")}
    out = dict(form=form, ms_per_step=ms_per_step, rows_per_s=cfg["m"] / (ms_per_step * 1e-3),
               kernel_avg_ms=(tot_ms / cnt if cnt else None), kernel_launches=cnt, rel_fro_err=None,
               finalize_ms=None, state_err_vs_fp64=None,
               kernel_ms={k: (v[1] / v[0] if v[0] else None) for k, v in stats.items()})
    if error:
        X, Y, Z = sk.state()
        if ref_state is not None:
            out["state_err_vs_fp64"] = {nm: float(np.linalg.norm(a - b) / np.linalg.norm(b))
                                        for nm, a, b in zip("XYZ", (X, Y, Z), ref_state)}
        if keep_state:
            out["state"] = (X, Y, Z)
        # finalize (QR of Y and X^T, core solve, k x k SVD; FP64 on every form)
        sk.low_rank_approx(device=True)
        torch.cuda.synchronize()
        before = dict(sk.timings()["timings"].get("approx", {}))
        t0 = time.perf_counter()
        U, Sv, V = sk.low_rank_approx(device=True)
        torch.cuda.synchronize()
        out["finalize_ms"] = (time.perf_counter() - t0) * 1e3
        after = sk.timings()["timings"].get("approx", {})
        out["finalize_phases_ms"] = {key: (after[key] - before.get(key, 0.0)) * 1e3 for key in after}
        # rank-r quality: ||A - U S V^T||_F / ||A||_F by the streamed residual pass
        e2, a2 = S.residual_fro(A, U, Sv, V, window=cfg["window"], ctx=ctx)
        out["rel_fro_err"] = math.sqrt(e2 / a2)
    sk.close()
    return out


def h2d_inclusive(S, ctx, cfg, A, maps, form, torch, reps=2):
    """rows/s when A starts in pinned host memory: one host-to-device copy of
    the whole matrix followed by linear_update (sequential, no overlap)."""
    m, n = cfg["m"], cfg["n"]
    Ah = torch.empty((n, m), dtype=torch.float64, pin_memory=True)
    Ah.copy_(A.t())
    Ad = torch.empty((n, m), dtype=torch.float64, device="cuda")
    p = S.AlgParams(matrix_m=m, matrix_n=n, rank=cfg["rank"], sketch_range=cfg["k"], sketch_core=cfg["s"],
                    window=cfg["window"], precision=form)
    sk = S.SketchySVD(p, maps=maps, ctx=ctx)
    sk.linear_update(Ad.t())
    torch.cuda.synchronize()
    t_copy = t_all = 0.0
    for _ in range(reps):
        t0 = time.perf_counter()
        Ad.copy_(Ah, non_blocking=True)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        sk.linear_update(Ad.t())
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        t_copy += t1 - t0
        t_all += t2 - t0
    sk.close()
    del Ad, Ah
    return {"form": form, "rows_per_s": m * reps / t_all, "ms_per_step": t_all * 1e3 / reps,
            "h2d_ms": t_copy * 1e3 / reps, "h2d_gbs": 8.0 * m * n * reps / t_copy / 1e9,
            "note": "pinned host A copied whole, then linear_update; no overlap"}


def forms_block(results, base, tail):
    return {f: {"rows_per_s": r["rows_per_s"], "ms_per_step": r["ms_per_step"], "finalize_ms": r["finalize_ms"],
                "finalize_phases_ms": r.get("finalize_phases_ms"),
                "rel_fro_err": r["rel_fro_err"],
                "err_vs_fp64": (r["rel_fro_err"] / base["rel_fro_err"])
                if (r["rel_fro_err"] and base.get("rel_fro_err")) else None,
                "err_vs_optimal": (r["rel_fro_err"] / tail) if (tail and r["rel_fro_err"]) else None,
                "state_err_vs_fp64": r["state_err_vs_fp64"],
                "speedup_vs_fp64": r["rows_per_s"] / base["rows_per_s"], "kernel_ms": r["kernel_ms"]}
            for f, r in results.items()}


def time_forms(S, ctx, cfg, A, maps, forms, args, torch):
    results = {}
    ref_state = None
    for form in forms:
        try:
            print(f"[bench] {cfg['label']}: timing form {form}", file=sys.stderr, flush=True)
            r = run_form(S, ctx, cfg, A, maps, form, args.steps, args.warmup, torch, error=not args.no_error,
                         ref_state=ref_state, keep_state=(form == "fp64"))
            if form == "fp64" and "state" in r:
                ref_state = r.pop("state")
            results[form] = r
            print(f"[bench] {form}: {r['ms_per_step']:.2f} ms/step, err {r['rel_fro_err']}, "
                  f"finalize {r['finalize_ms']} ms, state err {r['state_err_vs_fp64']}", file=sys.stderr, flush=True)
        except S.SkemaError as e:
            if e.code != 7:
                raise
            print(f"[bench] form {form} not available: {e}", file=sys.stderr, flush=True)
    return results


def cpu_baseline(S, cfg, A, maps, budget_s):
    """The oracle's window update on host cores over consecutive windows of
    the same A and maps until budget_s of CPU time has been spent."""
    from oracle import skema_oracle as O  # the checker, timed as the baseline

    try:
        from threadpoolctl import threadpool_info

        threads = max([i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"] or [1])
    except Exception:
        threads = os.cpu_count() or 1
    op = O.AlgParams(matrix_m=cfg["m"], matrix_n=cfg["n"], rank=cfg["rank"], sketch_range=cfg["k"],
                     sketch_core=cfg["s"], window=cfg["window"])
    omaps = {name: O.GaussDimRedux(maps[name].download(), False) for name in ("Upsilon", "Omega", "Phi", "Psi")}
    osk = O.SketchySVD(op, maps=omaps)
    osk.init_state()
    w, m = cfg["window"], cfg["m"]
    spent, rows = 0.0, 0
    for r0 in range(0, m, w):
        r1 = min(m, r0 + w)
        H = A[r0:r1].cpu().numpy()
        t0 = time.perf_counter()
        osk.update_window(H, (r0, r1))
        spent += time.perf_counter() - t0
        rows += r1 - r0
        if spent >= budget_s:
            break
    return dict(value=rows / spent, unit="rows/s", cores=int(threads), kind="port",
                sample=f"oracle (numpy/OpenBLAS, {threads} threads) window update over the first {rows} of "
                       f"{m} rows ({rows // w} windows of {w}) of {cfg.get('label', 'the matrix')}, same maps; "
                       f"{spent:.1f} s", **host_cpu())


def cpu_baseline_windows(op, H_of, m, w, budget_s, what, sparse_convention="rows"):
    """The oracle's window update on host cores, window after window of the
    workload (H_of(r0, r1) builds the window the oracle's Window::get would),
    until budget_s of CPU time has been spent; rows/s of the windows done."""
    from oracle import skema_oracle as O  # the checker, timed as the baseline

    try:
        from threadpoolctl import threadpool_info

        threads = max([i.get("num_threads", 1) for i in threadpool_info() if i.get("user_api") == "blas"] or [1])
    except Exception:
        threads = os.cpu_count() or 1
    osk = O.SketchySVD(op, sparse_convention=sparse_convention)
    osk.init_state()
    spent, rows = 0.0, 0
    for r0 in range(0, m, w):
        r1 = min(m, r0 + w)
        t0 = time.perf_counter()
        osk.update_window(H_of(r0, r1), (r0, r1))
        spent += time.perf_counter() - t0
        rows = r1
        if spent >= budget_s:
            break
    return dict(value=rows / spent, unit="rows/s", cores=threads, kind="port",
                sample=f"oracle (numpy/scipy/OpenBLAS, {threads} threads) {what} over the first {rows} of {m} rows "
                       f"(windows of {w}, the oracle's own maps of the same kind and sizes); {spent:.1f} s",
                **host_cpu())


def make_sparse(cfg):
    import scipy.sparse as sp

    m, n = cfg["m"], cfg["n"]
    rng = np.random.default_rng(cfg["seed"])
    nnz = int(round(cfg["density"] * m * n))
    lin = np.unique(rng.integers(0, m, nnz).astype(np.int64) * n + rng.integers(0, n, nnz))
    rows, cols = lin // n, lin % n
    vals = rng.standard_normal(len(lin))
    A = sp.csr_matrix((vals, (rows, cols)), shape=(m, n))
    A.sort_indices()
    return A


def make_points(cfg):
    rng = np.random.default_rng(cfg["seed"])
    n, d = cfg["m"], cfg["d"]
    mu = (2.0 * rng.standard_normal((cfg["clusters"], d))).astype(np.float32)
    c = rng.integers(0, cfg["clusters"], n)
    X = (mu[c] + rng.standard_normal((n, d)).astype(np.float32)).astype(np.float64)
    return X


def bench_sparse(args, S, torch):
    """CSR linear_update with sparse-sign maps (one pass over the whole CSR):
    bandwidth-bound gathers and scatters."""
    cfg = WORKLOADS[args.workload]
    A = make_sparse(cfg)
    Ad = S.CrsMatrix.from_scipy(A)
    ctx = S.Context(0)
    p = S.AlgParams(matrix_m=cfg["m"], matrix_n=cfg["n"], rank=cfg["rank"], sketch_range=cfg["k"],
                    sketch_core=cfg["s"], window=cfg["window"], dim_redux="sparse-sign", issparse=True)
    maps = S.make_maps(p, ctx, convention="cols")
    sk = S.SketchySVD(p, maps=maps, ctx=ctx)
    for _ in range(args.warmup):
        sk.linear_update(Ad)
    torch.cuda.synchronize()
    ctx.reset_stats()
    ctx.set_profiling(True, only="window")  # the dominant group only inside the timed region
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sk.linear_update(Ad)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    ctx.set_profiling(False)
    win_stats = ctx.kernel_stats("window")
    ctx.reset_stats()  # per-group breakdown from one more, fully profiled step
    ctx.set_profiling(True)
    sk.linear_update(Ad)
    torch.cuda.synchronize()
    ctx.set_profiling(False)
    stats = {kname: ctx.kernel_stats(kname) for kname in ("window", "sketch_left", "sketch_right", "sketch_core")}
    stats["window"] = win_stats
    m, n, k, s, w, z = cfg["m"], cfg["n"], cfg["k"], cfg["s"], cfg["window"], cfg["zeta"]
    nnz = A.nnz
    # Algorithmic HBM bytes of each stage of the whole-matrix pass (one
    # group of launches per step): CSR / CSC entries 12 B (int32 index +
    # double), row pointers 8 B, map-transpose entries 12 B, dense outputs 8 B
    # per element (Y read and written: nu != 0).
    bytes_per = {
        "window": 24 * nnz + 8 * (m + 1) + 8 * (n + 1),
        "sketch_left": 12 * nnz + 8 * (n + 1) + 2 * (12 * z * m + 8 * (m + 1)) + 8 * (k + s) * n,
        "sketch_right": 12 * nnz + 8 * (m + 1) + 12 * z * n + 8 * (n + 1) + 16 * m * k,
        "sketch_core": 8 * s * n + 12 * z * n + 8 * (s + 1) + 8 * s * s,
    }
    dom = max((kn for kn in bytes_per if stats[kn][0]), key=lambda kname: stats[kname][1])
    avg_ms = stats[dom][1] / stats[dom][0]
    achieved = bytes_per[dom] / (avg_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC, "value": m / dt, "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded generators; no dataset)",
        "config": {"workload": f"{args.workload}: CSR {m}x{n}, nnz={A.nnz}, r={cfg['rank']} k={k} s={s} "
                               f"window={w}, sparse-sign maps zeta={z} per column", "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                     "frac": achieved / 8000.0, "traffic": None, "bytes_per_launch": bytes_per[dom],
                     "avg_launch_ms": avg_ms},
        "step_bytes": sum(bytes_per.values()),
        "step_hbm_gbs": sum(bytes_per.values()) / dt / 1e9,
        "kernel_ms": {kname: (v[1] / v[0] if v[0] else None) for kname, v in stats.items()},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle)", file=sys.stderr, flush=True)
        from oracle import skema_oracle as O  # the checker, timed as the baseline

        op = O.AlgParams(matrix_m=m, matrix_n=n, rank=cfg["rank"], sketch_range=k, sketch_core=s, window=w,
                         dim_redux="sparse-sign", issparse=True)
        out["cpu_baseline"] = cpu_baseline_windows(op, lambda r0, r1: A[r0:r1], m, w, args.cpu_budget,
                                                   "CSR window update (scipy sparse x sparse-sign maps)",
                                                   sparse_convention="cols")
    print(json.dumps(out))


def rbf_residual(S, Xd, gamma, U, Sv, V, window, ctx):
    """||K - U diag(S) V^T||_F / ||K||_F with K streamed in materialised
    windows (the parity generator), one window at a time."""
    n = Xd.shape[0]
    e2 = a2 = 0.0
    for r0 in range(0, n, window):
        r1 = min(n, r0 + window)
        Kw = S.gauss_rbf_window(Xd, gamma, (r0, r1), ctx=ctx)
        e, a = S.residual_fro(Kw, U[r0:r1], Sv, V, window=window, ctx=ctx)
        e2 += e
        a2 += a
        del Kw
    return math.sqrt(e2 / a2)


def bench_rbf(args, S, torch):
    """Fused RBF window + Gauss sketch (K never written to HBM), FP64 and
    BF16 forms; value = the faster form whose rank-r error is within 1.05x of
    the FP64 form's."""
    cfg = WORKLOADS[args.workload]
    X = make_points(cfg)
    Xd = S.colmajor(torch.as_tensor(X, device="cuda"))
    ctx = S.Context(0)
    n, d, k, s, w = cfg["m"], cfg["d"], cfg["k"], cfg["s"], cfg["window"]
    gamma = 1.0 / d
    p64 = S.AlgParams(matrix_m=n, matrix_n=n, rank=cfg["rank"], sketch_range=k, sketch_core=s, window=w,
                      kernel_func="gaussrbf", kernel_gamma=gamma, nfeat=d)
    maps = S.make_maps(p64, ctx, dist="normal")
    forms = [f for f in args.forms.split(",") if f in ("fp64", "bf16")] or ["fp64"]
    if "fp64" not in forms:
        forms.insert(0, "fp64")
    res = {}
    for form in forms:
        print(f"[bench] rbf form {form}", file=sys.stderr, flush=True)
        p = S.AlgParams(matrix_m=n, matrix_n=n, rank=cfg["rank"], sketch_range=k, sketch_core=s, window=w,
                        kernel_func="gaussrbf", kernel_gamma=gamma, nfeat=d, precision=form)
        sk = S.SketchySVD(p, maps=maps, ctx=ctx)
        for _ in range(args.warmup):
            sk.linear_update(Xd)
        torch.cuda.synchronize()
        ctx.reset_stats()
        ctx.set_profiling(True, only="rbf_right")  # the dominant group only inside the timed region
        t0 = time.perf_counter()
        for _ in range(args.steps):
            sk.linear_update(Xd)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        ctx.set_profiling(False)
        right_stats = ctx.kernel_stats("rbf_right")
        ctx.reset_stats()  # per-group breakdown from one more, fully profiled step
        ctx.set_profiling(True)
        sk.linear_update(Xd)
        torch.cuda.synchronize()
        ctx.set_profiling(False)
        stats = {kname: ctx.kernel_stats(kname) for kname in ("window", "rbf_left", "rbf_right", "sketch_core")}
        stats["rbf_right"] = right_stats
        sk.low_rank_approx(device=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        U, Sv, V = sk.low_rank_approx(device=True)
        torch.cuda.synchronize()
        fin_ms = (time.perf_counter() - t0) * 1e3
        err = rbf_residual(S, Xd, gamma, U, Sv, V, w, ctx)
        sk.close()
        res[form] = dict(form=form, ms=dt * 1e3, rows_per_s=n / dt, err=err, finalize_ms=fin_ms,
                         kernel_ms={kn: (v[1] / v[0] if v[0] else None) for kn, v in stats.items()},
                         right_avg_ms=stats["rbf_right"][1] / stats["rbf_right"][0],
                         svals_head=[float(x) for x in Sv[:5].cpu()])
        print(f"[bench] rbf {form}: {dt * 1e3:.1f} ms/step, err {err:.6e}", file=sys.stderr, flush=True)
    base = res["fp64"]
    best = base
    for f, r in res.items():
        if f != "fp64" and r["err"] <= 1.05 * base["err"] and r["rows_per_s"] > best["rows_per_s"]:
            best = r
    nwin = math.ceil(n / w)
    avg_rows = n / nwin
    # rbf_right per launch: sketch 2 w n (k+s) + Gram 2 w n d (algorithmic, one Gram)
    flops_alg = 2.0 * avg_rows * n * (k + s) + 2.0 * avg_rows * n * d
    achieved = flops_alg / (best["right_avg_ms"] * 1e-3) / 1e12
    peak = PEAKS_TFLOPS[best["form"]]
    out = {
        "metric": METRIC, "value": best["rows_per_s"], "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": best["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": DTYPE[best["form"]], "data": "synthetic (seeded generators; no dataset)",
        "config": {"workload": f"{args.workload}: Gaussian RBF of {n} points in d={d} (10-cluster mixture), "
                               f"gamma=1/d, r={cfg['rank']} k={k} s={s} window={w}, fused generator",
                   "form": best["form"], "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "kernel": "rbf_right", "achieved": achieved, "peak": peak,
                     "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                     "flops_per_launch": flops_alg, "avg_launch_ms": best["right_avg_ms"]},
        "finalize_ms": best["finalize_ms"],
        "forms": {f: {"rows_per_s": r["rows_per_s"], "ms_per_step": r["ms"], "rel_fro_err": r["err"],
                      "finalize_ms": r["finalize_ms"],
                      "err_vs_fp64": r["err"] / base["err"], "speedup_vs_fp64": r["rows_per_s"] / base["rows_per_s"],
                      "kernel_ms": r["kernel_ms"], "svals_head": r["svals_head"]} for f, r in res.items()},
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle)", file=sys.stderr, flush=True)
        from oracle import skema_oracle as O  # the checker, timed as the baseline

        op = O.AlgParams(matrix_m=n, matrix_n=n, rank=cfg["rank"], sketch_range=k, sketch_core=s, window=w,
                         kernel_func="gaussrbf", kernel_gamma=gamma)
        out["cpu_baseline"] = cpu_baseline_windows(
            op, lambda r0, r1: O.gauss_rbf_compute(X[r0:r1], X, gamma, (r0, r1)), n, 64, args.cpu_budget,
            "RBF window (K rows by direct differences, as the reference's GaussRBFWindow) and update, "
            "64-row windows")
    print(json.dumps(out))


def bench_adjoint(args, S, torch):
    """Forward loss and the hand-written adjoint (dL/dX, dL/dgamma) in the FP64
    form and the BF16 form (BF16 Gram, sketch and tile products, FP32
    accumulation); the BF16 gradient's error is stated against FP64's."""
    cfg = WORKLOADS[args.workload]
    X = make_points(cfg)
    Xd = S.colmajor(torch.as_tensor(X, device="cuda"))
    ctx = S.Context(0)
    n, d, k, s = cfg["m"], cfg["d"], cfg["k"], cfg["s"]
    gamma = 0.5 / d
    om = S.GaussDimRedux(k, n, seed=0, dist="normal", ctx=ctx)
    ph = S.GaussDimRedux(s, n, seed=1, dist="normal", ctx=ctx)
    ps = S.GaussDimRedux(s, n, seed=2, dist="normal", ctx=ctx)
    forms = [f for f in args.forms.split(",") if f in ("fp64", "bf16")] or ["fp64", "bf16"]
    h = k + s
    # tile pass: (F_I R_J^T) with inner 2h, the Gram over d, and H X_J over d
    flops = float(n) * n * (4.0 * h + 4.0 * d)
    res = {}
    for form in forms:
        def run(grad, steps):
            out = None
            for _ in range(steps):
                out = S.rbf_sketch_loss_grad(Xd, gamma, om, ph, ps, window=cfg["window"], grad=grad, ctx=ctx,
                                             precision=form)
            return out

        print(f"[bench] adjoint form {form}", file=sys.stderr, flush=True)
        run(True, args.warmup)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        run(False, args.steps)
        torch.cuda.synchronize()
        t_fwd = (time.perf_counter() - t0) / args.steps
        ctx.reset_stats()
        ctx.set_profiling(True)
        t0 = time.perf_counter()
        loss, dX, dg = run(True, args.steps)
        torch.cuda.synchronize()
        t_all = (time.perf_counter() - t0) / args.steps
        ctx.set_profiling(False)
        cnt, tot = ctx.kernel_stats("rbf_adjoint")
        fc, ft = ctx.kernel_stats("rbf_forward")
        adj_ms = tot / cnt
        peak = PEAKS_TFLOPS[form]
        achieved = flops / (adj_ms * 1e-3) / 1e12
        res[form] = dict(ms=t_all * 1e3, forward_ms=t_fwd * 1e3, forward_kernel_ms=(ft / fc if fc else None),
                         adjoint_ms=adj_ms, achieved=achieved, peak=peak, loss=loss, dgamma=dg,
                         dX=dX.clone())
        print(f"[bench] {form}: {t_all * 1e3:.2f} ms/step (tile pass {adj_ms:.2f} ms, {achieved:.1f} TFLOP/s)",
              file=sys.stderr, flush=True)
    best = min(res, key=lambda f: res[f]["ms"])
    b = res[best]
    forms_out = {}
    for f, r in res.items():
        e = {"ms_per_step": r["ms"], "forward_ms": r["forward_ms"], "forward_kernels_ms": r["forward_kernel_ms"],
             "adjoint_tile_pass_ms": r["adjoint_ms"], "achieved_tflops": r["achieved"], "loss": r["loss"],
             "dgamma": r["dgamma"]}
        if "fp64" in res and f != "fp64":
            ref = res["fp64"]
            e["speedup_vs_fp64"] = ref["ms"] / r["ms"]
            e["loss_rel_err_vs_fp64"] = abs(r["loss"] - ref["loss"]) / abs(ref["loss"])
            e["dX_rel_err_vs_fp64"] = float(torch.linalg.norm(r["dX"] - ref["dX"]) / torch.linalg.norm(ref["dX"]))
            e["dgamma_rel_err_vs_fp64"] = abs(r["dgamma"] - ref["dgamma"]) / abs(ref["dgamma"])
        forms_out[f] = e
    out = {
        "metric": METRIC, "value": n / (b["ms"] * 1e-3), "unit": "rows/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": b["ms"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": DTYPE[best], "data": "synthetic (seeded generators; no dataset)",
        "config": {"workload": f"{args.workload}: loss + adjoint of the fused RBF sketch, {n} points in d={d}, "
                               f"gamma=0.5/d, k={k} s={s}, forward windows {cfg['window']}",
                   "form": best, "parallelism": "single GPU"},
        "roofline": {"bound": "mfma", "kernel": "rbf_adjoint", "achieved": b["achieved"], "peak": b["peak"],
                     "unit": "TFLOP/s", "frac": b["achieved"] / b["peak"], "traffic": None,
                     "flops_per_launch": flops, "avg_launch_ms": b["adjoint_ms"]},
        "forms": forms_out, "cpu_baseline": None,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="dense-65536x16384", choices=sorted(WORKLOADS))
    ap.add_argument("--forms", default="fp64,bf16x3,fp8,bf16,fp32")
    ap.add_argument("--cpu-budget", type=float, default=12.0, help="seconds of CPU baseline work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a known-size kernel that "
                         "calibrates the FETCH_SIZE / WRITE_SIZE counters of a rocprofv3 --pmc run)")
    ap.add_argument("--no-lownoise", action="store_true",
                    help="skip the low-noise companion matrix of the default dense workload")
    ap.add_argument("--no-error", action="store_true",
                    help="time the update only: no finalize, residual or optimal-tail passes (counter runs); "
                         "the forms are then ranked by speed alone")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the sketch path is built and measured on a single MI355X; "
              "multi-GPU sharding is a design note only (row-window sharding + one all-reduce of X and Z).",
              file=sys.stderr)
        sys.exit(2)

    import torch

    import skema_amd as S

    if args.copy_probe_mb:
        nb = args.copy_probe_mb << 20
        src = torch.ones(nb // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"[bench] copy probe: {nb} bytes read, {nb} bytes written", file=sys.stderr, flush=True)
        del src, dst
    if args.workload.startswith("sparse-"):
        return bench_sparse(args, S, torch)
    if args.workload.startswith("rbf-adjoint"):
        return bench_adjoint(args, S, torch)
    if args.workload.startswith("rbf-"):
        return bench_rbf(args, S, torch)
    return bench_dense(args, S, torch)


def bench_dense(args, S, torch):
    """Every form on the dense workload and, for the default workload, on its
    low-noise companion (noise standard deviation 1e-4 instead of variance
    1e-4).  `value` is the fastest form whose rank-r error is within 1.05x of
    the FP64 form's on every matrix timed."""
    cfg = dict(WORKLOADS[args.workload], label=args.workload)
    forms = args.forms.split(",")
    if not args.no_error and "fp64" not in forms:
        forms.insert(0, "fp64")
    print(f"[bench] building {args.workload} on cuda:0", file=sys.stderr, flush=True)
    A = make_dense(cfg, torch)
    ctx = S.Context(0)
    p64 = S.AlgParams(matrix_m=cfg["m"], matrix_n=cfg["n"], rank=cfg["rank"], sketch_range=cfg["k"],
                      sketch_core=cfg["s"], window=cfg["window"])
    maps = S.make_maps(p64, ctx, dist="normal")
    results = time_forms(S, ctx, cfg, A, maps, forms, args, torch)
    tail = None
    if not args.no_error:
        print("[bench] optimal rank-r tail (subspace iteration)", file=sys.stderr, flush=True)
        tail = optimal_tail(A, cfg["rank"], torch)
    # the low-noise companion matrix (same spectrum and seeds)
    low = None
    low_name = args.workload + "-lownoise"
    if not args.no_error and not args.no_lownoise and low_name in WORKLOADS:
        cfg_low = dict(WORKLOADS[low_name], label=low_name)
        print(f"[bench] building {low_name} on cuda:0", file=sys.stderr, flush=True)
        A_keep = A
        A = None
        del A_keep
        torch.cuda.empty_cache()
        A_low = make_dense(cfg_low, torch)
        res_low = time_forms(S, ctx, cfg_low, A_low, maps, forms, args, torch)
        print("[bench] optimal rank-r tail of the low-noise matrix", file=sys.stderr, flush=True)
        tail_low = optimal_tail(A_low, cfg_low["rank"], torch)
        low = dict(cfg=cfg_low, results=res_low, tail=tail_low)
        A = A_low
    # headline: fastest form within 1.05x of FP64's rank-r error on every matrix timed
    if "fp64" in results and not args.no_error:
        base = results["fp64"]
        ok = [f for f, r in results.items() if r["rel_fro_err"] <= 1.05 * base["rel_fro_err"]
              and (low is None or (f in low["results"] and
                                   low["results"][f]["rel_fro_err"] <= 1.05 * low["results"]["fp64"]["rel_fro_err"]))]
        best = max((results[f] for f in ok), key=lambda r: r["rows_per_s"])
    else:  # no FP64 reference run (profiling a single form): report the fastest form
        best = max(results.values(), key=lambda r: r["rows_per_s"])
        base = best
    form = best["form"]
    name, work, bound = dominant_kernel(cfg, form)
    if bound == "hbm":
        achieved = work / (best["kernel_avg_ms"] * 1e-3) / 1e9 if best["kernel_avg_ms"] else None
        roof = {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
                "frac": (achieved / 8000.0) if achieved else None, "traffic": committed_traffic(form + "_pass"),
                "bytes_per_launch": work, "avg_launch_ms": best["kernel_avg_ms"]}
    else:
        mfma_flops = work * MFMA_PER_FLOP[form]
        achieved = mfma_flops / (best["kernel_avg_ms"] * 1e-3) / 1e12 if best["kernel_avg_ms"] else None
        roof = {"bound": "mfma", "kernel": name, "achieved": achieved, "peak": PEAKS_TFLOPS[form], "unit": "TFLOP/s",
                "frac": (achieved / PEAKS_TFLOPS[form]) if achieved else None, "traffic": committed_traffic(form),
                "flops_per_launch": mfma_flops, "avg_launch_ms": best["kernel_avg_ms"]}
    # the right product's own figure (from the profiled breakdown step)
    rname, rflops = right_product(cfg, form)
    rms = best["kernel_ms"].get(rname)
    right_roof = None
    if rms:
        rach = rflops * MFMA_PER_FLOP[form] / (rms * 1e-3) / 1e12
        right_roof = {"bound": "mfma", "kernel": rname, "achieved": rach, "peak": PEAKS_TFLOPS[form],
                      "unit": "TFLOP/s", "frac": rach / PEAKS_TFLOPS[form], "traffic": committed_traffic(form),
                      "avg_launch_ms": rms, "note": "from the fully profiled breakdown step"}
    h2d = None
    if not args.no_error:
        print(f"[bench] host-buffer rate of {form}", file=sys.stderr, flush=True)
        h2d = h2d_inclusive(S, ctx, cfg, A, maps, form, torch)
    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (oracle)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(S, cfg, A, maps, args.cpu_budget)
    out = {
        "metric": METRIC,
        "value": best["rows_per_s"],
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": best["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[form],
        "data": "synthetic (seeded generators; no dataset)",
        "config": {
            "workload": f"{args.workload}: dense A = U diag({cfg['decay']}^i) V^T + N(0, {cfg['noise_var']:g}) noise "
                        f"(variance), r={cfg['rank']} k={cfg['k']} s={cfg['s']} window={cfg['window']}, "
                        f"Gaussian maps seeds 0-3, A resident in HBM",
            "m": cfg["m"], "n": cfg["n"], "rank": cfg["rank"], "range": cfg["k"], "core": cfg["s"],
            "window": cfg["window"], "form": form, "parallelism": "single GPU",
            "selection": "fastest form whose rank-r error is within 1.05x of the FP64 form's on this matrix"
                         + (" and on the low-noise matrix" if low else ""),
        },
        "roofline": roof,
        "roofline_right_product": right_roof,
        "cpu_baseline": cpu,
        "finalize_ms": best["finalize_ms"],
        "h2d_inclusive": h2d,
        "forms": forms_block(results, base, tail),
        "flops_per_row": flops_per_row(cfg),
        "optimal_rank_r_err": tail,
    }
    if low is not None:
        out["lownoise"] = {
            "workload": f"{low_name}: as above with N(0, {low['cfg']['noise_var']:g}) noise (variance)",
            "optimal_rank_r_err": low["tail"],
            "forms": forms_block(low["results"], low["results"]["fp64"], low["tail"]),
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
