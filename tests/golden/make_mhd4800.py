"""Fixtures for the MHD4800 real-pencil test (examples/EXRAND3.f inputs of NLAFET/GEVP-PDHGEQZ).

  mhd4800a.mtx.gz, mhd4800b.mtx.gz  the reference's examples/sparseA.mtx and sparseB.mtx (Matrix Market
                                    data files of its example driver), gzip-compressed;
  mhd4800_lapack.npz                LAPACK's generalized eigenvalues of (A, B) (DGGEV via scipy, with left
                                    and right eigenvectors) as homogeneous pairs (alpha, beta) and their
                                    condition numbers kappa_j = ||x|| ||y|| / |(y^H A x, y^H B x)|, so that the
                                    GPU test can compare without running LAPACK on a 4800 pencil.
Run in a checkout next to the reference tree: python tests/golden/make_mhd4800.py /path/to/reference
"""
import gzip
import os
import shutil
import sys

import numpy as np
import scipy.io as sio
import scipy.linalg as sl

HERE = os.path.dirname(os.path.abspath(__file__))


def main(ref):
    for src, dst in (("sparseA.mtx", "mhd4800a.mtx.gz"), ("sparseB.mtx", "mhd4800b.mtx.gz")):
        with open(os.path.join(ref, "examples", src), "rb") as f, gzip.open(os.path.join(HERE, dst), "wb", 9) as g:
            shutil.copyfileobj(f, g)
    A = sio.mmread(os.path.join(ref, "examples", "sparseA.mtx")).toarray()
    B = sio.mmread(os.path.join(ref, "examples", "sparseB.mtx")).toarray()
    w, vl, vr = sl.eig(A, B, left=True, right=True, homogeneous_eigvals=True)
    a, b = w
    # y_j^H A x_j and y_j^H B x_j for all j at once (real A, B times the complex eigenvector blocks)
    ya = np.einsum("ij,ij->j", vl.conj(), A @ vr.real + 1j * (A @ vr.imag))
    yb = np.einsum("ij,ij->j", vl.conj(), B @ vr.real + 1j * (B @ vr.imag))
    den = np.maximum(np.sqrt(np.abs(ya) ** 2 + np.abs(yb) ** 2), 1e-300)
    cond = np.linalg.norm(vr, axis=0) * np.linalg.norm(vl, axis=0) / den
    np.savez_compressed(os.path.join(HERE, "mhd4800_lapack.npz"), alpha=a, beta=b, cond=cond,
                        fro_a=np.linalg.norm(A), fro_b=np.linalg.norm(B))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
