"""MHD4800: the reference's example pencil (examples/sparseA.mtx, sparseB.mtx, driven by
examples/EXRAND3.f:300-406) read with the library's Matrix Market reader, taken through the whole
pipeline on the GPU (QR of B, HT reduction, multishift QZ with the PIPARMQ defaults), and compared
with LAPACK's generalized eigenvalues of the same pencil (tests/golden/mhd4800_lapack.npz, written by
tests/golden/make_mhd4800.py).

The comparison is condition-weighted: each eigenvalue may move by
fudge * n * eps * ||(A, B)||_F * kappa_j in the chordal metric, kappa_j from LAPACK's left and right
eigenvectors.  Matching goes both ways (every computed eigenvalue near a LAPACK one and every LAPACK
eigenvalue near a computed one) with nearest neighbours on the Riemann sphere, where the chordal
distance is the Euclidean one; a full assignment over 4800 points would dominate the test's time."""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "gevp-pdhgeqz_amd"))
sys.path.insert(0, HERE)

GOLD = os.path.join(HERE, "golden")
FA, FB = os.path.join(GOLD, "mhd4800a.mtx.gz"), os.path.join(GOLD, "mhd4800b.mtx.gz")
FL = os.path.join(GOLD, "mhd4800_lapack.npz")
EPS = np.finfo(float).eps
RES_TOL = 1e-13
ORTH_TOL = 10.0


def sphere(alpha, beta):
    """Points of the Riemann sphere of diameter 1 for homogeneous eigenvalues (alpha : beta)."""
    alpha = np.asarray(alpha, dtype=complex)
    beta = np.asarray(beta, dtype=complex)
    fin = beta != 0
    lam = np.zeros(len(alpha), dtype=complex)
    lam[fin] = alpha[fin] / beta[fin]
    m2 = np.abs(lam) ** 2
    p = np.stack([lam.real / (1 + m2), lam.imag / (1 + m2), m2 / (1 + m2)], axis=1)
    p[~fin] = (0.0, 0.0, 1.0)
    return p


def test_sphere_distance_is_chordal():
    import qzcheck as QC
    rng = np.random.default_rng(3)
    a = rng.standard_normal(50) + 1j * rng.standard_normal(50)
    b = rng.standard_normal(50).astype(complex)
    b[:3] = 0.0
    p = sphere(a, b)
    d = np.linalg.norm(p[:, None, :] - p[None, :, :], axis=2)
    c = QC.chordal(a[:, None], b[:, None], a[None, :], b[None, :])
    assert np.allclose(d, c, atol=1e-14)


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(FL), reason="mhd4800_lapack.npz not generated")
def test_mhd4800_full_pipeline_against_lapack():
    import torch
    from scipy.spatial import cKDTree
    import gqz
    print("mhd4800 full pipeline", flush=True)
    A, ha = gqz.mm_read(FA)
    B, hb = gqz.mm_read(FB)
    assert A.shape == (4800, 4800) and B.shape == (4800, 4800)
    assert ha["rep"] == "coordinate" and hb["rep"] == "coordinate"
    n = A.shape[0]
    gqz.init(0)
    Ad = gqz.to_colmajor(torch.from_numpy(A))
    Bd = gqz.to_colmajor(torch.from_numpy(B))
    S, P, Q, Z, ar, ai, be, info, st = gqz.schur_pipeline(Ad, Bd)
    assert info == 0
    # residuals with torch's own products (not the library's kernels)
    A0, B0 = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    I = torch.eye(n, dtype=torch.float64, device="cuda")
    ra = float(torch.linalg.norm(Q.t() @ A0 @ Z - S) / torch.linalg.norm(A0))
    rb = float(torch.linalg.norm(Q.t() @ B0 @ Z - P) / torch.linalg.norm(B0))
    rq = float(torch.linalg.norm(Q.t() @ Q - I)) / (n * EPS)
    rz = float(torch.linalg.norm(Z.t() @ Z - I)) / (n * EPS)
    print("residuals", ra, rb, rq, rz, flush=True)
    assert ra < RES_TOL and rb < RES_TOL and rq < ORTH_TOL and rz < ORTH_TOL
    g = np.load(FL)
    a2, b2, cond = g["alpha"], g["beta"], g["cond"]
    a1 = ar.cpu().numpy() + 1j * ai.cpu().numpy()
    b1 = be.cpu().numpy().astype(complex)
    nab = np.hypot(float(g["fro_a"]), float(g["fro_b"]))
    tol = 100.0 * n * EPS * nab * cond
    p1, p2 = sphere(a1, b1), sphere(a2, b2)
    d12, i12 = cKDTree(p2).query(p1)       # computed -> nearest LAPACK
    d21, i21 = cKDTree(p1).query(p2)       # LAPACK -> nearest computed
    r12 = d12 / np.maximum(tol[i12], 1e-300)
    r21 = d21 / np.maximum(tol, 1e-300)
    print("mhd4800: worst ratios %.3g %.3g, max chordal %.2e, infinite %d (LAPACK %d)"
          % (r12.max(), r21.max(), max(d12.max(), d21.max()), int((be.cpu().numpy() == 0).sum()),
             int((np.abs(b2) == 0).sum())), flush=True)
    assert r12.max() < 1.0 and r21.max() < 1.0
