"""Argument handling of the command-line driver (skema_amd/skema_sketch).

Every case here fails (or prints the usage) while the arguments are parsed,
before the driver touches a GPU, so these run on the CPU.  The flag table is
the reference's (src/Skema_AlgParams.cpp:242-330, driver/main.cpp:149-224).
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "skema_amd", "skema_sketch")
FIXTURE = os.path.join(ROOT, "tests", "golden", "rectangular_sparse.mtx")

pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="driver not built")


def _run(*args):
    return subprocess.run([EXE, *args], capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("args,needle", [
    (["--input", FIXTURE, "--bogus", "1"], "unknown argument --bogus"),
    # a value-less flag must not swallow the next token: --bogus is seen
    (["--input", FIXTURE, "--force-three-sketch-off", "--bogus"], "unknown argument --bogus"),
    (["--input", FIXTURE, "--debug", "--bogus"], "unknown argument --bogus"),
    (["--input", FIXTURE, "--sketch-compute-svals-iters", "--bogus"], "unknown argument --bogus"),
    (["--input", FIXTURE, "--rr-off", "--bogus"], "unknown argument --bogus"),
    (["--input", FIXTURE, "--isvd-initial-guess", "--bogus"], "unknown argument --bogus"),
    (["--input", FIXTURE, "--normalize-matrix", "-2"], "Bad input: --normalize-matrix"),
    (["--input", FIXTURE, "--normalize-matrix", "two"], "Bad input: --normalize-matrix"),
    (["--input", FIXTURE, "--rank", "0"], "Bad input: --rank"),
    (["--input", FIXTURE, "--window", "3x"], "Bad input: --window"),
    (["--input", FIXTURE, "--print-level", "7"], "--print-level"),
    (["--input", FIXTURE, "--solver", "isvd"], "only --solver sketch"),
    (["--input", FIXTURE, "--rr"], "--rr"),
    (["--input", FIXTURE, "--model", "fourier"], "Bad input: --model fourier"),
    (["--input", FIXTURE, "--kernel", "laplace"], "Bad input: --kernel laplace"),
    (["--input", FIXTURE, "--map-dist", "cauchy"], "Bad input: --map-dist"),
    (["--input", FIXTURE, "--precision", "fp4"], "unknown --precision"),
    (["--input", FIXTURE, "--rank"], "missing value for --rank"),
    (["--rank", "3"], "must provide --input"),
    # the reference also takes flags without their leading "--"
    (["input", FIXTURE, "bogus"], "unknown argument bogus"),
])
def test_rejected(args, needle):
    r = _run(*args)
    assert r.returncode != 0
    assert needle in r.stderr, r.stderr


def test_help_prints_usage():
    r = _run("--help")
    assert r.returncode == 0 and "Usage: skema_sketch" in r.stdout


def test_other_solver_flags_are_parsed_not_swallowed():
    """iSVD / PRIMME flags are accepted (with a note that they are ignored)
    and keep their value arity: the token after a value flag is its value,
    the token after a value-less one is parsed as a flag."""
    r = _run("--input", FIXTURE, "--isvd-num-samples", "5", "--primme_locking", "--bogus")
    assert r.returncode != 0 and "unknown argument --bogus" in r.stderr
    assert "--isvd-num-samples applies to the isvd / primme solvers; ignored" in r.stderr
