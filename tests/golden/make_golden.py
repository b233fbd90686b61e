"""Writes tests/golden/golden.npz: data fixtures that pin the generator and the QZ eigenvalues.

The reference ships no expected outputs (its tests only check residuals, testing/EXRAND2.f:
404-407), so the eigenvalue fixtures are LAPACK DGGES results (scipy.linalg, bundled LAPACK) on
seeded N(0,1) pencils from the SplitMix64/Box-Muller generator.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np
import scipy.linalg as sl

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_ffi as O  # noqa: E402
import qzcheck as Q  # noqa: E402


def main():
    out = {}
    out["normal_seed1_16"] = O.normal(1, (16,))
    out["uniform_seed7_16"] = O.uniform(7, (16,))
    for n, seed in ((16, 3), (48, 5)):
        A, B = Q.pencil_randn(n, seed)
        a, b = sl.eigvals(A, B, homogeneous_eigvals=True)
        out["seed_%d" % n] = np.array(seed)
        out["A_%d_corner" % n] = A[:2, :2].copy()
        out["lapack_alpha_%d" % n] = a
        out["lapack_beta_%d" % n] = b
    np.savez(os.path.join(HERE, "golden.npz"), **out)
    print("wrote", os.path.join(HERE, "golden.npz"))


if __name__ == "__main__":
    main()
