"""Re-entrancy: the library keeps one context (stream, workspace, pinned staging, settings) per host
thread, so QZ calls from two host threads run at the same time.  The reference keeps its tolerances
and history in COMMON blocks (src/pdhgeqz/PDHGEQZ.f:291-294) and is not re-entrant; the per-call
state here is the `Run` struct, the per-thread state the library context.

Each thread reduces its own pencil (default parameters, so AED windows take the nested QZ, the
reorder, the blocked HT restore and the MFMA slab updates) while the other one runs; each result
must pass the residual checks and give the same eigenvalues as a run of the same pencil alone."""
import threading

import numpy as np
import pytest

import qzcheck as QC

pytestmark = pytest.mark.gpu

RES_TOL = 1e-13
ORTH_TOL = 10.0


def _solve(gqz, torch, H, T):
    Hd, Td = gqz.to_colmajor(torch.from_numpy(H)), gqz.to_colmajor(torch.from_numpy(T))
    n = H.shape[0]
    Q, Z = gqz.colmajor(n), gqz.colmajor(n)
    ar, ai, be, info, st = gqz.pdhgeqz("S", "I", "I", Hd, Td, Q, Z)
    out = [x.cpu().numpy() for x in (Hd, Td, Q, Z, ar, ai, be)]
    return out, info, st


def test_two_host_threads_run_qz_concurrently():
    import torch
    import gqz
    gqz.init(0)
    print("two host threads, n = 700 and 640", flush=True)
    pencils = [QC.pencil_ht_random(700, 21), QC.pencil_ht_random(640, 22)]
    alone = [_solve(gqz, torch, H, T) for H, T in pencils]
    res = [None, None]
    errs = []

    def work(k):
        try:
            res[k] = _solve(gqz, torch, *pencils[k])
        except Exception as e:          # reported below, the test fails
            errs.append(repr(e))

    ths = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in ths:
        t.start()
    for t in ths:
        t.join(timeout=300)
    assert not errs, errs
    assert all(r is not None for r in res)
    for k, ((H, T), r, a) in enumerate(zip(pencils, res, alone)):
        (S, P, Q, Z, ar, ai, be), info, st = r
        assert info == 0 and st["aed_calls"] > 0
        ra, rb, rq, rz = QC.compres(H, T, S, P, Q, Z)
        assert ra < RES_TOL and rb < RES_TOL and rq < ORTH_TOL and rz < ORTH_TOL, (k, ra, rb, rq, rz)
        (_, _, _, _, ar0, ai0, be0), info0, _ = a
        assert info0 == 0
        assert QC.match_two(ar, ai, be, ar0, ai0, be0, H, T) < 1.0
