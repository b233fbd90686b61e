"""GPU tests on the workloads of the benchmark configurations, at sizes that fit the test budget:

  config 1  dense N(0,1) pencil, QR(B) + HT reduction + multishift QZ, n = 1024, against the oracle's
            pipeline on the same pencil and against LAPACK;
  config 2  dense N(0,1) pencil of order 8000 through the whole pipeline (blocked QR of B, blocked HT
            reduction, multishift QZ with the PIPARMQ defaults NS = 256, NW = 384);
  config 3  the stress family (sparse + 0.5 I, zero diagonal entries in triangular B, graded) at
            n = 1024: infinite-eigenvalue count equal to the oracle's on the same HT pencil;
  config 4  a draw of the batched distribution (n log-uniform in [32, 1024]) including both ends;
  config 5  the learned parameters on a held-out HT pencil of order 12000.
Residuals are PCOMPRES (tools/PQZHLPS.f:19-319) with the bounds the drivers gate on
(testing/EXRAND2.f:404-407: RESA, RESB; RESQ, RESZ < 10 in n eps units), RESA, RESB < 1e-13 here.
Large sizes use the device PCOMPRES, which tests/test_gpu_parity.py checks against numpy."""
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


def _host(t):
    return t.cpu().numpy()


def _note(*a):
    print(*a, flush=True)


def test_config1_dense_pipeline_n1024():
    torch, gqz, O = _mods()
    n = 1024
    _note("config 1: dense N(0,1) pencil, full pipeline, n =", n)
    A, B = QC.pencil_randn(n, 1)
    S, P, Q, Z, ar, ai, be, info, st = gqz.schur_pipeline(gqz.to_colmajor(torch.from_numpy(A)),
                                                          gqz.to_colmajor(torch.from_numpy(B)))
    assert info == 0
    S, P, Q, Z = _host(S), _host(P), _host(Q), _host(Z)
    ra, rb, rq, rz = QC.compres(A, B, S, P, Q, Z)
    assert ra < RES_TOL and rb < RES_TOL and rq < ORTH_TOL and rz < ORTH_TOL
    below, tb, ok2 = QC.schur_structure(S, P)
    assert below == 0 and tb == 0 and ok2
    ar, ai, be = _host(ar), _host(ai), _host(be)
    assert QC.match_eigs(ar, ai, be, A, B)[0] < 1.0
    # the oracle's pipeline on the same pencil (its own HT form)
    Ho, To, _, _ = O.ht_reduce(A, B)
    _, _, oar, oai, obe, _, _, oinfo, _ = O.hgeqz(Ho, To)
    assert oinfo == 0
    assert QC.match_two(ar, ai, be, oar, oai, obe, A, B) < 1.0


def test_config2_full_pipeline_default_parameters_n8000():
    torch, gqz, O = _mods()
    n = 8000
    _note("config 2: dense N(0,1) pencil, QR(B) + HT reduction + QZ with PIPARMQ defaults, n =", n)
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    A0 = torch.randn((n, n), dtype=torch.float64, device="cuda", generator=g).t().contiguous().t()
    B0 = torch.randn((n, n), dtype=torch.float64, device="cuda", generator=g).t().contiguous().t()
    H, T, Q, Z, ar, ai, be, info, st = gqz.schur_pipeline(A0, B0)
    _note("  ", {k: st[k] for k in ("iters", "aed_calls", "sweeps", "swap_fail", "t_total")})
    assert info == 0 and st["sweeps"] > 0 and st["aed_calls"] > 0
    ra, rb, rq, rz = gqz.compres(A0, B0, H, T, Q, Z)
    assert ra < RES_TOL and rb < RES_TOL and rq < ORTH_TOL and rz < ORTH_TOL
    # quasi-triangular, standardised: nothing below the subdiagonal, no two consecutive subdiagonal
    # entries, T upper triangular (checked on the device)
    assert float(torch.tril(H, -2).abs().max()) == 0.0 and float(torch.tril(T, -1).abs().max()) == 0.0
    sub = torch.diagonal(H, -1)
    assert not bool(((sub[:-1] != 0) & (sub[1:] != 0)).any())
    assert bool((be >= 0).all())


def test_config3_stress_pencil_n1024():
    torch, gqz, O = _mods()
    n = 1024
    _note("config 3: stress pencil (sparse + 0.5 I, zero diagonal entries in B, graded), n =", n)
    A, B = QC.pencil_stress(n, 3)
    H, T, Q, Z = O.ht_reduce(A, B)
    Hd, Td, Qd, Zd = (gqz.to_colmajor(torch.from_numpy(x)) for x in (H, T, Q, Z))
    ar, ai, be, info, _ = gqz.pdhgeqz("S", "V", "V", Hd, Td, Qd, Zd)
    assert info == 0
    _, _, oar, oai, obe, _, _, oinfo, _ = O.hgeqz(H, T)
    assert oinfo == 0
    be = _host(be)
    n_inf = int((be == 0).sum())
    _note("   infinite eigenvalues: GPU", n_inf, "oracle", int((obe == 0).sum()))
    assert n_inf == int((obe == 0).sum()) and n_inf > 0
    ra, rb, rq, rz = QC.compres(A, B, _host(Hd), _host(Td), _host(Qd), _host(Zd))
    assert ra < RES_TOL and rb < RES_TOL and rq < ORTH_TOL and rz < ORTH_TOL


def test_config4_batched_draw():
    torch, gqz, O = _mods()
    rng = np.random.default_rng(4)
    ns = [1024, 32] + [int(round(32.0 * 32.0 ** u)) for u in rng.uniform(0, 1, 62)]
    _note("config 4: batched draw of", len(ns), "pencils, n in [%d, %d]" % (min(ns), max(ns)))
    A = [rng.standard_normal((k, k)) for k in ns]
    B = [rng.standard_normal((k, k)) for k in ns]
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
    S, P, Qh, Zh, arh, aih, beh, infh = (_host(x) for x in (dA, dB, dQ, dZ, ar, ai, be, info))
    assert (infh == 0).all()
    for i, k in enumerate(ns):
        o, e = off[i], eoff[i]
        M = lambda x: x[o:o + k * k].reshape(k, k, order="F")
        ra, rb, rq, rz = QC.compres(A[i], B[i], M(S), M(P), M(Qh), M(Zh))
        assert ra < RES_TOL and rb < RES_TOL and rq < ORTH_TOL and rz < ORTH_TOL, (i, k)
        below, tb, ok2 = QC.schur_structure(M(S), M(P))
        assert below == 0 and tb == 0 and ok2, (i, k)
        if i < 2 or i % 8 == 0:
            assert QC.match_eigs(arh[e:e + k], aih[e:e + k], beh[e:e + k], A[i], B[i])[0] < 1.0, (i, k)


def test_config5_learned_parameters_n12000():
    torch, gqz, O = _mods()
    import bench
    import gqz_policy
    n = 12000
    _note("config 5: learned parameters, held-out HT pencil, n =", n)
    gqz_policy.install_learned(gqz.lib())
    H0, T0 = bench.ht_random_device(n, 12, torch, gqz)
    H, T = H0.clone().t().contiguous().t(), T0.clone().t().contiguous().t()
    Q = gqz.colmajor(n); Z = gqz.colmajor(n)
    ar, ai, be, info, st = gqz.pdhgeqz("S", "I", "I", H, T, Q, Z, params=gqz_policy.learned_params())
    _note("  ", {k: st[k] for k in ("iters", "aed_calls", "sweeps", "swap_fail", "t_total")})
    assert info == 0
    ra, rb, rq, rz = gqz.compres(H0, T0, H, T, Q, Z)
    assert ra < RES_TOL and rb < RES_TOL and rq < ORTH_TOL and rz < ORTH_TOL
    assert float(torch.tril(H, -2).abs().max()) == 0.0 and float(torch.tril(T, -1).abs().max()) == 0.0
