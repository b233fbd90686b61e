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


def test_committed_policy_windows():
    """At the start of a QZ call the policy keeps NS and NIBBLE at their defaults and shrinks the AED
    window below PIPARMQ's NW, to a few dozen rows at n >= 4000 (the windows that the fastest static
    runs used: about 48 at n = 16000).  The multiplier is clamped as the library clamps it."""
    P = G.load().astype(np.float64)
    prev = None
    for n in (1000, 2000, 4000, 8000, 16000, 32000):
        y = O.mlp_forward(P, _initial_state(n)[None, :])[0]
        assert abs(y[0]) < 0.25 and abs(y[2]) < 0.25, (n, y)
        nw = round(G.piparmq_nw(n) * 2.0 ** min(1.0, max(-5.0, y[1])))
        assert 16 <= nw <= G.piparmq_nw(n), (n, nw)
        if n >= 4000:
            assert 24 <= nw <= 96, (n, nw)
        if prev is not None:
            assert y[1] <= prev + 0.05, (n, y[1], prev)
        prev = y[1]


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
