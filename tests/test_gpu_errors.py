"""GPU tests of the failure paths of the QZ (PDHGEQZ.f:496-502: INFO > 0 when the iteration does not
converge, INFO = KBOT + 1 of PDHGEQZ0 / PDHGEQZ1).  A NaN in H keeps every deflation test false, so
the iteration runs into its cap: the top-level loop stops after MAXIT = maxit * NH iterations
(gqz_params.maxit, 30 in the reference) and the LDS-resident QZ of a window or a batched pencil after
40 * n + 20 decisions.  Every loop of the path is bounded by such a cap or by the pencil's order, so
the calls return; the tests check that they return INFO > 0 and leave the other pencils of a batch
intact."""
import ctypes as C

import numpy as np
import pytest

import qzcheck as QC

pytestmark = pytest.mark.gpu

RES_TOL = 1e-13
ORTH_TOL = 10.0


def _mods():
    import torch
    import gqz
    import oracle_ffi as O
    gqz.init(0)
    return torch, gqz, O


def test_nan_input_returns_info_after_the_iteration_cap():
    torch, gqz, O = _mods()
    n = 100
    print("NaN in H, n =", n, flush=True)
    A = O.normal(11, (n, n)); B = O.normal(12, (n, n))
    H, T, _, _ = O.ht_reduce(A, B)
    H = H.copy()
    H[50, 60] = np.nan
    Hd, Td = (gqz.to_colmajor(torch.from_numpy(x)) for x in (H, T))
    Qd = gqz.colmajor(n); Zd = gqz.colmajor(n)
    prm = gqz.GqzParams()
    prm.maxit = 2                     # MAXIT = 2 * NH top-level iterations
    ar, ai, be, info, st = gqz.pdhgeqz("S", "I", "I", Hd, Td, Qd, Zd, params=prm)
    print("INFO =", info, "iterations", st["iters"], flush=True)
    assert 1 <= info <= n
    assert st["iters"] <= 2 * n + 1


def test_batched_pencil_with_nan_reports_info_and_leaves_the_others_intact():
    torch, gqz, O = _mods()
    ns = [40, 60, 50]
    print("batched with a NaN pencil", ns, flush=True)
    rng = np.random.default_rng(21)
    A = [rng.standard_normal((k, k)) for k in ns]
    B = [rng.standard_normal((k, k)) for k in ns]
    A[1][10, 20] = np.nan
    off = np.concatenate([[0], np.cumsum([k * k for k in ns])[:-1]]).astype(np.int64)
    eoff = np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.int64)
    flat = lambda Ms: np.concatenate([M.ravel(order="F") for M in Ms])
    tot = int(sum(k * k for k in ns)); et = int(sum(ns))
    dA = torch.from_numpy(flat(A)).cuda(); dB = torch.from_numpy(flat(B)).cuda()
    dQ = torch.empty(tot, dtype=torch.float64, device="cuda"); dZ = torch.empty_like(dQ)
    ar = torch.zeros(et, dtype=torch.float64, device="cuda"); ai = torch.zeros_like(ar); be = torch.zeros_like(ar)
    info = torch.full((len(ns),), -7, dtype=torch.int32, device="cuda")
    nn = np.ascontiguousarray(ns, dtype=np.int32)
    vp = lambda t: C.c_void_p(t.data_ptr())
    torch.cuda.synchronize()
    r = gqz.lib().gqz_batched_dev(len(ns), nn.ctypes.data_as(C.c_void_p), off.ctypes.data_as(C.c_void_p),
                                  eoff.ctypes.data_as(C.c_void_p), vp(dA), vp(dB), vp(dQ), vp(dZ), vp(ar),
                                  vp(ai), vp(be), vp(info), 0)
    assert r == 0
    infh = info.cpu().numpy()
    print("INFO per pencil", infh, flush=True)
    assert infh[0] == 0 and infh[2] == 0 and infh[1] > 0
    S, P, Qh, Zh = (x.cpu().numpy() for x in (dA, dB, dQ, dZ))
    for i in (0, 2):
        k, o = ns[i], off[i]
        M = lambda x: x[o:o + k * k].reshape(k, k, order="F")
        ra, rb, rq, rz = QC.compres(A[i], B[i], M(S), M(P), M(Qh), M(Zh))
        assert ra < RES_TOL and rb < RES_TOL and rq < ORTH_TOL and rz < ORTH_TOL, (i, ra, rb, rq, rz)
