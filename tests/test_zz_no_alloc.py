"""The allocation-free batched entry point (gqz_batched_noalloc_dev) against gqz_batched_dev: two calls
in a row through the C ABI, the second with fewer and smaller pencils, each bit-identical to the
allocating entry point on the same input (same kernels, same workgroup size), and the size limit."""
import ctypes as C

import numpy as np
import pytest

import qzcheck as QC

pytestmark = pytest.mark.gpu


def _draw(count, nmax, seed):
    rng = np.random.default_rng(seed)
    ns = [int(round(4.0 * (nmax / 4.0) ** u)) for u in rng.uniform(0, 1, count)]
    A = [rng.standard_normal((k, k)) for k in ns]
    B = [rng.standard_normal((k, k)) for k in ns]
    off = np.concatenate([[0], np.cumsum([k * k for k in ns])[:-1]]).astype(np.int64)
    eoff = np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.int64)
    flat = lambda Ms: np.concatenate([M.ravel(order="F") for M in Ms])
    return ns, A, B, off, eoff, flat(A), flat(B)


def _run(entry, torch, ns, fA, fB, off, eoff, nt):
    tot = int(sum(k * k for k in ns)); et = int(sum(ns))
    dA = torch.from_numpy(fA).cuda(); dB = torch.from_numpy(fB).cuda()
    dQ = torch.full((tot,), 7.0, dtype=torch.float64, device="cuda"); dZ = torch.full_like(dQ, 7.0)
    ar = torch.full((et,), 7.0, dtype=torch.float64, device="cuda"); ai = torch.full_like(ar, 7.0)
    be = torch.full_like(ar, 7.0)
    info = torch.full((len(ns),), -7, dtype=torch.int32, device="cuda")
    nn = np.ascontiguousarray(ns, dtype=np.int32)
    vp = lambda t: C.c_void_p(t.data_ptr())
    torch.cuda.synchronize()
    r = entry(len(ns), nn.ctypes.data_as(C.c_void_p), off.ctypes.data_as(C.c_void_p),
              eoff.ctypes.data_as(C.c_void_p), vp(dA), vp(dB), vp(dQ), vp(dZ), vp(ar), vp(ai), vp(be), vp(info), nt)
    torch.cuda.synchronize()
    return r, [x.cpu().numpy() for x in (dA, dB, dQ, dZ, ar, ai, be, info)]


def test_noalloc_batched_matches_allocating_entry_point():
    import torch
    import gqz
    L = gqz.init(0)
    for call, (count, nmax, seed) in enumerate(((40, 180, 1), (17, 60, 2))):
        print("no-alloc batched call", call, "count", count, "max n", nmax, flush=True)
        ns, A, B, off, eoff, fA, fB = _draw(count, nmax, seed)
        r1, ref = _run(L.gqz_batched_dev, torch, ns, fA, fB, off, eoff, 256)
        r2, got = _run(L.gqz_batched_noalloc_dev, torch, ns, fA, fB, off, eoff, 256)
        assert r1 == 0 and r2 == 0, L.gqz_last_error()
        for a, b in zip(ref, got):
            assert np.array_equal(a, b)
        assert (got[7] == 0).all()
        # and the results are right (PCOMPRES on a few pencils)
        S, P, Q, Z = got[:4]
        for i in (0, count // 2, count - 1):
            k, o = ns[i], off[i]
            M = lambda x: x[o:o + k * k].reshape(k, k, order="F")
            ra, rb, rq, rz = QC.compres(A[i], B[i], M(S), M(P), M(Q), M(Z))
            assert ra < 1e-13 and rb < 1e-13 and rq < 10 and rz < 10


def test_noalloc_batched_rejects_counts_above_its_limit():
    import gqz
    L = gqz.init(0)
    count = 8193
    nn = np.full(count, 4, dtype=np.int32)
    off = (16 * np.arange(count)).astype(np.int64); eoff = (4 * np.arange(count)).astype(np.int64)
    dummy = C.c_void_p(16)      # never dereferenced: the size check comes first
    r = L.gqz_batched_noalloc_dev(count, nn.ctypes.data_as(C.c_void_p), off.ctypes.data_as(C.c_void_p),
                                  eoff.ctypes.data_as(C.c_void_p), dummy, dummy, dummy, dummy, dummy, dummy,
                                  dummy, dummy, 256)
    assert r == -1 and b"8192" in L.gqz_last_error()
