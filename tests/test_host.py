"""CPU tests of the host-side logic around the kernels: the Python restatement of the PIPARMQ
defaults that the learned policy scales (piparmq.f:445-476) agrees with the library's own, the
committed policy file is plain data and chooses sensible AED windows from the state every QZ call
starts in, and bench.py refuses more than one GPU before it touches a device."""
import ctypes as C
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import gqz_policy as G
import oracle_ffi as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "gevp-pdhgeqz_amd", "libgqz.so")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("libgqz.so not built")
    L = C.CDLL(LIB)
    L.gqz_piparmq.argtypes = [C.c_int, C.c_int]
    return L


def test_policy_defaults_match_library(lib):
    """NS (ISPEC 53) and NW (ISPEC 51) of the training code equal the library's at every order the
    thresholds of piparmq.f:445-476 separate, including the orders on both sides of each threshold."""
    orders = set(range(2, 700)) | {n + d for n in (3000, 6000, 12000, 24000, 48000) for d in (-1, 0, 1)}
    orders |= {1000, 2047, 4096, 8000, 16000, 65536}
    for n in sorted(orders):
        assert G.piparmq_ns(n) == lib.gqz_piparmq(53, n), n
        assert G.piparmq_nw(n) == lib.gqz_piparmq(51, n), n


def _initial_state(n):
    """Features of the first top-level iteration (features() in qz_host.cpp with empty history)."""
    f = np.zeros(12)
    f[0] = math.log2(max(n, 2)) / 16.0
    f[1] = 1.0
    f[11] = f[0]
    return f


def test_committed_policy_is_plain_data():
    with np.load(G.MODEL, allow_pickle=False) as z:
        P = z["params"]
    assert P.dtype == np.float32 and P.shape == (G.NPAR,)
    assert np.all(np.isfinite(P))


def test_policy_nibble_default_matches_library(lib):
    """NIBBLE (ISPEC 52) of the training code equals the library's (14 below 6000, then
    ceil(335 n^-0.44))."""
    for n in list(range(2, 200)) + [5999, 6000, 6001, 8000, 12000, 16000, 24000, 100000]:
        assert G.piparmq_nibble(n) == lib.gqz_piparmq(52, n), n


def test_committed_policy_choices():
    """At the start of a QZ call (empty deflation history) the policy keeps NS at its default, takes
    an AED window below PIPARMQ's NW (the fastest runs of the training grid used 48 or 64 rows, against
    NW = 96..768; the fit of these first-iteration states is loose, up to about 2x) and a sweep-skip
    threshold above PIPARMQ's (the fastest runs used 20..60 %, against 5..14 %).  The multipliers are
    clamped as the library clamps them."""
    P = G.load().astype(np.float64)
    for n in (2000, 4000, 8000, 12000, 16000):
        y = O.mlp_forward(P, _initial_state(n)[None, :])[0]
        assert abs(y[0]) < 0.25, (n, y)
        nw = round(G.piparmq_nw(n) * 2.0 ** min(1.0, max(-5.0, y[1])))
        nib = round(G.piparmq_nibble(n) * 2.0 ** min(4.0, max(-2.0, y[2])))
        assert 32 <= nw <= 160 and 3 * nw <= 2 * G.piparmq_nw(n), (n, nw)
        assert nib > G.piparmq_nibble(n) and 8 <= nib <= 80, (n, nib)


def test_committed_policy_summary():
    """The committed model records the fastest grid point of every training pencil: an AED window of
    48 or 64 and sweep-skip thresholds of 20..60 %, each faster than the PIPARMQ defaults where those
    were run."""
    import json
    with np.load(G.MODEL, allow_pickle=False) as z:
        summary = json.loads(str(z["summary"]))
    assert len(summary) >= 6
    for s in summary:
        nw, nib_e, nib_l = s["best"]
        assert nw in (48, 64) and 20 <= nib_e <= 60 and 20 <= nib_l <= 60, s
        if s.get("default_s"):
            assert s["best_s"] < s["default_s"], s


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
