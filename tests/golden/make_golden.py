"""Writes tests/golden/toy_sketch.npz: oracle X, Y, Z, U, S, V (and the
materialised maps) for a 64 x 48 toy, N(0,1) Gauss maps from the counter stream,
windows of 10 rows, rank 4, nu = 0.9, eta = 1.1.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import skema_oracle as O  # noqa: E402


def build():
    m, n = 64, 48
    rng = np.random.default_rng(2024)
    A = rng.standard_normal((m, n)) @ np.diag(0.8 ** np.arange(n)) + 1e-3 * rng.standard_normal((m, n))
    p = O.AlgParams(matrix_m=m, matrix_n=n, rank=4, window=10, sketch_nu=0.9, sketch_eta=1.1)
    sk = O.SketchySVD(p, dist="normal")
    sk.linear_update(A)
    X, Y, Z = sk.X.copy(), sk.Y.copy(), sk.Z.copy()
    U, S, V = sk.low_rank_approx()
    return dict(A=A, X=X, Y=Y, Z=Z, U=U, S=S, V=V, Upsilon=sk.Upsilon.data, Omega=sk.Omega.data,
                Phi=sk.Phi.data, Psi=sk.Psi.data, params=np.array([m, n, 4, 10, 0.9, 1.1]))


if __name__ == "__main__":
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "toy_sketch.npz")
    np.savez(out, **build())
    print("wrote", out)
