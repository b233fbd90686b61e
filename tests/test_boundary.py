"""CPU tests of the drop-in boundary: the C ABI library and the Fortran shim load and export every
entry point declared in include/gqz.h; argument checking, INFO codes and the workspace-query
protocol follow PDHGEQZ.f:384-444 and PDGGHRD.f:300-336; without a GPU the library refuses to
run (there is no CPU fallback).  No compute call reaches a GPU here."""
import ctypes as C
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "gevp-pdhgeqz_amd")
HDR = os.path.join(ROOT, "include", "gqz.h")
LIB = os.path.join(PKG, "libgqz.so")
FLIB = os.path.join(PKG, "libpdhgeqz_f.so")


def declared():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:int|void|const char\s*\*|void\s*\*)\s*\*?\s*(gqz_\w+)\s*\(", txt, flags=re.M)
    return sorted(set(names))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("libgqz.so not built")
    return C.CDLL(LIB)


def test_header_declares_core_entry_points():
    names = declared()
    for f in ("gqz_hgeqz_dev", "gqz_gghrd_dev", "gqz_pdhgeqz_host", "gqz_pdgghrd_host", "gqz_compres_dev",
              "gqz_init", "gqz_batched_dev"):
        assert f in names


def test_library_exports_every_declared_symbol(lib):
    missing = [f for f in declared() if not hasattr(lib, f)]
    assert not missing, missing


def test_fortran_shim_exports():
    if not os.path.exists(FLIB):
        pytest.skip("Fortran shim not built (amdflang missing)")
    F = C.CDLL(FLIB)
    for sym in ("pdhgeqz_", "pdgghrd_", "kkqzconf_", "pilaenvx_"):
        assert hasattr(F, sym)


def _desc(n, nb=64, mb=None, lld=None):
    mb = nb if mb is None else mb
    return (C.c_int * 9)(1, 0, n, n, mb, nb, 0, 0, n if lld is None else lld)


def _host_call(lib, job=b"S", cq=b"I", cz=b"I", n=8, ilo=1, ihi=None, nb=64, mb=None, ldwork=None, liwork=None):
    ihi = n if ihi is None else ihi
    m = max(n, 1)
    H = np.zeros((m, m), order="F"); T = np.eye(m, order="F")
    Q = np.eye(m, order="F"); Z = np.eye(m, order="F")
    ar = np.zeros(m); ai = np.zeros(m); be = np.zeros(m)
    ldwork = max(1, n * nb) if ldwork is None else ldwork
    liwork = max(1, 2 * n) if liwork is None else liwork
    dw = np.zeros(max(1, ldwork)); iw = np.zeros(max(1, liwork), dtype=np.int32)
    d = _desc(max(n, 1), nb, mb)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    lib.gqz_pdhgeqz_host.argtypes = [C.c_char, C.c_char, C.c_char, C.c_int, C.c_int, C.c_int,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_int, C.c_void_p, C.c_int]
    info = lib.gqz_pdhgeqz_host(job, cq, cz, n, ilo, ihi, vp(H), d, vp(T), d, vp(ar), vp(ai), vp(be),
                                vp(Q), d, vp(Z), d, vp(dw), ldwork, vp(iw), liwork)
    return info, dw, iw


@pytest.mark.parametrize("kw,expect", [
    (dict(job=b"X"), -1), (dict(cq=b"Q"), -2), (dict(cz=b"?"), -3), (dict(n=-1), -4),
    (dict(ilo=0), -5), (dict(ihi=9), -6), (dict(nb=4), -7), (dict(mb=32), -8),
    (dict(ldwork=10), -9), (dict(liwork=3), -10)])
def test_pdhgeqz_info_codes(lib, kw, expect):
    """INFO codes of PDHGEQZ.f:384-405 and :435-444 (no GPU needed: checks come first)."""
    info, _, _ = _host_call(lib, **kw)
    assert info == expect


def test_pdhgeqz_workspace_query(lib):
    """LDWORK = -1 / LIWORK = -1 returns the optimum in DWORK(1), IWORK(1) (PDHGEQZ.f:409-430)."""
    info, dw, iw = _host_call(lib, n=50, nb=60, ldwork=-1, liwork=-1)
    assert info == 0 and dw[0] == 50 * 60 and iw[0] == 100


def test_pilaenvx_via_c(lib):
    lib.gqz_piparmq.argtypes = [C.c_int, C.c_int]
    assert lib.gqz_piparmq(50, 1000) == 6000
    assert (lib.gqz_piparmq(53, 8000), lib.gqz_piparmq(51, 8000), lib.gqz_piparmq(52, 8000)) == (256, 384, 7)
    assert (lib.gqz_piparmq(53, 16000), lib.gqz_piparmq(51, 16000), lib.gqz_piparmq(52, 16000)) == (512, 768, 5)


def test_fortran_pilaenvx_and_info():
    """The Fortran entry points pass arguments by reference with hidden character lengths."""
    if not os.path.exists(FLIB):
        pytest.skip("Fortran shim not built")
    F = C.CDLL(FLIB)
    ic = lambda v: C.byref(C.c_int(v))
    F.pilaenvx_.restype = C.c_int
    name = C.c_char_p(b"PDHGEQZ0"); opts = C.c_char_p(b"")
    v = F.pilaenvx_(ic(0), ic(53), name, opts, ic(16000), ic(1), ic(16000), ic(0), C.c_size_t(8), C.c_size_t(0))
    assert v == 512
    # a bad JOB returns INFO = -1 through the Fortran ABI
    n = 4
    d = _desc(n)
    H = np.zeros((n, n), order="F"); T = np.eye(n, order="F")
    e = np.zeros(n); dw = np.zeros(n * 64); iw = np.zeros(2 * n, dtype=np.int32)
    info = C.c_int(123)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    F.pdhgeqz_(C.c_char_p(b"X"), C.c_char_p(b"I"), C.c_char_p(b"I"), ic(n), ic(1), ic(n), vp(H), d, vp(T), d,
               vp(e), vp(e), vp(e), vp(H), d, vp(H), d, vp(dw), ic(n * 64), vp(iw), ic(2 * n), C.byref(info),
               C.c_size_t(1), C.c_size_t(1), C.c_size_t(1))
    assert info.value == -1


def test_no_gpu_fails_loudly(lib):
    """Without a GPU the library reports an error instead of computing on the CPU."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("GPU present")
    except Exception:
        pass
    info, _, _ = _host_call(lib, n=8)
    assert info == -100
    lib.gqz_last_error.restype = C.c_char_p
    assert b"GPU" in lib.gqz_last_error() or b"HIP" in lib.gqz_last_error() or b"hip" in lib.gqz_last_error()


def test_bench_traffic_summary_is_consistent():
    """The counter summary that bench.py reads for roofline.traffic covers the kernel classes it
    names, with calibration factors from kernels of known size."""
    import json
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    sys.path.insert(0, root)
    import bench
    with open(bench.TRAFFIC_FILE) as f:
        t = json.load(f)
    assert t["n"] == 16000 and t["params"] == "learned"
    assert 1.9 < t["calibration"]["fetch_factor"] < 2.1 and 0.99 < t["calibration"]["write_factor"] < 1.01
    for cls, kernels in t["classes"].items():
        for k in kernels:
            if k in t["kernels"]:
                e = t["kernels"][k]
                assert e["dispatches"] > 0 and e["fetch_kib"] >= 0 and e["write_kib"] >= 0
    # chase windows: one kernel, traffic = calibrated FETCH + WRITE per dispatch
    e = t["kernels"]["k_chase"]
    want = 1024.0 * (e["fetch_kib"] * t["calibration"]["fetch_factor"] + e["write_kib"] * t["calibration"]["write_factor"])
    assert abs(bench.class_traffic("chase_window", 16000) - want) < 1e-6 * want
    assert bench.class_traffic("chase_window", 8000) is None       # other workloads: not covered


def _gghrd_call(lib, cq=b"I", cz=b"I", n=8, ilo=1, ihi=None, nb=64, mb=None, lwork=None):
    ihi = n if ihi is None else ihi
    m = max(n, 1)
    A = np.zeros((m, m), order="F"); B = np.eye(m, order="F")
    Q = np.eye(m, order="F"); Z = np.eye(m, order="F")
    lwork = (nb if mb is None else mb) if lwork is None else lwork
    w = np.zeros(max(1, lwork, 1))
    d = _desc(m, nb, mb)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)
    lib.gqz_pdgghrd_host.argtypes = [C.c_char, C.c_char, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_int]
    info = lib.gqz_pdgghrd_host(cq, cz, n, ilo, ihi, vp(A), d, vp(B), d, vp(Q), d, vp(Z), d, vp(w), lwork)
    return info, w


@pytest.mark.parametrize("kw,expect", [
    (dict(cq=b"X"), -1), (dict(cz=b"?"), -2), (dict(n=-1), -3), (dict(ilo=0), -4),
    (dict(ihi=9), -5), (dict(ilo=5, ihi=3), -5), (dict(mb=32), -6), (dict(lwork=10), -7)])
def test_pdgghrd_info_codes(lib, kw, expect):
    """INFO codes of PDGGHRD.f:318-336 (checked before anything reaches a GPU)."""
    info, _ = _gghrd_call(lib, **kw)
    assert info == expect


def test_pdgghrd_workspace_query_comes_first(lib):
    """LWORK = -1 returns DESCA(MB_) in WORK(1) before any argument check (PDGGHRD.f:256-266)."""
    info, w = _gghrd_call(lib, cq=b"X", nb=48, lwork=-1)
    assert info == 0 and w[0] == 48


def test_empty_ranges_are_accepted(lib):
    """ILO = N + 1, IHI = N is a valid empty range for both routines (PDHGEQZ.f:392-395,
    PDGGHRD.f:326-329): no INFO < 0; without a GPU the call then reports the missing device."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("GPU present")
    except Exception:
        pass
    info, _, _ = _host_call(lib, n=8, ilo=9, ihi=8)
    assert info == -100
    info, _ = _gghrd_call(lib, n=8, ilo=9, ihi=8)
    assert info == -100
    info, _ = _gghrd_call(lib, n=0, ilo=1, ihi=0)
    assert info == 0


def _fortran_driver():
    """Compile gevp-pdhgeqz_amd/fortran/exrand_gpu.f90, an EXRAND-style Fortran program, against
    the shim and the library (ordinary Fortran linking, no ctypes)."""
    fc = os.environ.get("GQZ_FC", "/opt/rocm/bin/amdflang")
    if not os.path.exists(FLIB) or not shutil.which(fc):
        pytest.skip("Fortran shim or compiler missing")
    obj = os.path.join(PKG, "_build")
    os.makedirs(obj, exist_ok=True)
    exe = os.path.join(obj, "exrand_gpu")
    src = os.path.join(PKG, "fortran", "exrand_gpu.f90")
    subprocess.run([fc, "-O2", "-J", obj, src, "-o", exe, "-L" + PKG, "-lpdhgeqz_f", "-lgqz",
                    "-Wl,-rpath," + PKG], check=True, capture_output=True, timeout=300)
    return exe


def _key_values(text):
    kv = {}
    for line in text.splitlines():
        f = line.split()
        if len(f) >= 2 and f[0].replace("_", "").isalnum():
            kv[f[0]] = f[1:]
    return kv


def test_fortran_program_argument_checks():
    """A compiled Fortran caller sees the reference's INFO codes, workspace queries and PIPARMQ
    values (PDHGEQZ.f:384-444, PDGGHRD.f:256-336, piparmq.f:440-517) through the shim."""
    r = subprocess.run([_fortran_driver(), "args"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    kv = _key_values(r.stdout)
    assert kv["pdhgeqz_bad_job"] == ["-1"]
    assert kv["pdhgeqz_bad_ilo"] == ["-5"]
    assert kv["pdhgeqz_small_nb"] == ["-7"]
    assert kv["pdhgeqz_query"] == ["0", str(8 * 60), "16"]
    assert kv["pdhgeqz_small_ldwork"] == ["-9"]
    assert kv["pdgghrd_query"] == ["0", "60"]
    assert kv["pdgghrd_bad_compq"] == ["-1"]
    assert kv["pdgghrd_bad_ihi"] == ["-5"]
    assert kv["pilaenvx_16000"] == ["512", "768", "5"]


def test_fortran_program_without_gpu_stops():
    """Without a GPU the Fortran program's first compute call returns INFO = -100 (no CPU path)."""
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("GPU present")
    except Exception:
        pass
    r = subprocess.run([_fortran_driver(), "run", "20"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2
    assert _key_values(r.stdout)["pdgghrd_info"] == ["-100"]


def test_pdtgord_workspace_query_and_info_codes(lib):
    """PDTGORD's descriptor surface (src/reorder/PDTGORD.f:1-4): LDWORK = -1 returns the routine's
    1x1-grid workspace sizes (PDTGORD.f:548-556) before any device work; too small workspaces give the
    routine's INFO codes -31 / -33 (PDTGORD.f:567-572).  No GPU needed: these paths return first."""
    import ctypes as C
    n, nb = 200, 60
    para = (C.c_int * 6)(1, nb // 2, nb, 0, min(nb, 32), nb // 2)
    d = (C.c_int * 9)(1, 0, n, n, nb, nb, 0, 0, n)
    sel = (C.c_int * n)()
    m = C.c_int(0)
    dw = (C.c_double * 1)(); iw = (C.c_int * 1)()
    vp = C.c_void_p
    f = lib.gqz_pdtgord_host
    f.argtypes = [C.c_int, C.c_int, vp, vp, C.c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                  C.POINTER(C.c_int), vp, C.c_int, vp, C.c_int]
    null = vp(0)
    args = lambda ldw, liw: (1, 1, C.cast(sel, vp), C.cast(para, vp), n, null, C.cast(d, vp), null, C.cast(d, vp),
                             null, C.cast(d, vp), null, C.cast(d, vp), null, null, null, C.byref(m),
                             C.cast(dw, vp), ldw, C.cast(iw, vp), liw)
    assert f(*args(-1, 1)) == 0
    lw = 2 * n + 14 * nb * nb + 8 * n * nb
    liw = 5 * 1 + (nb // 2) * nb - (nb // 2) * (nb // 2 + 1) // 2
    assert int(dw[0]) == lw and iw[0] == liw
    assert f(*args(lw - 1, liw)) == -31
    assert f(*args(lw, liw - 1)) == -33
