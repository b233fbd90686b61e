"""Learned parameter policy for the multishift QZ.

The reference chooses the number of shifts NS, the AED window NW and the sweep-skip threshold NIBBLE
from the active block order alone (PIPARMQ, src/pdhgeqz/piparmq.f:440-517, reached through PILAENVX
ISPEC 51-53).  Here a 12 -> 64 -> 64 -> 3 FP32 perceptron (k_mlp.hip) maps the deflation-history
features of each top-level iteration (features() in qz_host.cpp: log2 of the active order, its share
of n, the last deflation fractions, the previous choices, ...) to log2 multipliers of those default
values.  The QZ arithmetic is unchanged: only the three counts move, every result still has to pass
the backward-error check.

Training data come from the QZ itself: for each training pencil a grid of phase policies (an AED
window, and a sweep-skip threshold for the first half of the deflation and another for the second,
PhasePolicy) is run and every top-level iteration is recorded (gqz_trace_get).  The fastest run that
passes the backward-error check fixes the targets: every recorded state of the pencil gets the NW
target log2(W* / NW_def(n)) and the NIBBLE target log2(NIBBLE*(phase of that state) / NIBBLE_def(n))
(the defaults are evaluated once per QZ call, at the order n of the active block) and the NS target 0.
Training uses the library's own FP32 gradient and Adam kernels (gqz_mlp_train_step_dev) on the GPU.
"""
import ctypes as C
import math
import os
import time

import numpy as np

import gqz

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL = os.path.join(HERE, "mlp_policy.npz")
NPAR = 5187
FEATURES = ["log2(nh)/16", "nh/n", "nd/jw last", "nd/jw previous", "nd/jw mean of 3", "inf. deflations/nh",
            "NS/nh last", "NW/nh last", "NIBBLE/100 last", "swept last", "iterations/200", "log2(n)/16"]


# ---- PIPARMQ defaults (piparmq.f:445-476, as in qz_host.cpp) -------------------------------------------
def piparmq_ns(nh):
    ns = 2
    if nh >= 30: ns = 4
    if nh >= 60: ns = 10
    if nh >= 150: ns = max(10, nh // int(round(math.log(nh) / math.log(2.0))))
    for lim, v in ((590, 64), (3000, 128), (6000, 256), (12000, 512), (24000, 1024), (48000, 2048)):
        if nh >= lim: ns = v
    return max(2, ns - ns % 2)


def piparmq_nw(nh):
    ns = piparmq_ns(nh)
    return ns if nh <= 500 else 3 * ns // 2


def piparmq_nibble(nh):
    """NIBBLE of PDHGEQZ0 (PIPARMQ ISPEC 52; 14 below NMIN1 = 6000)."""
    if nh < 6000:
        return 14
    return min(100, int(math.ceil(335.0 * nh ** -0.44)))


# ---- parameters --------------------------------------------------------------------------------------
def init_params(seed=0):
    """Glorot-uniform weights, zero biases, in the k_mlp.hip layout; the last layer starts at zero so
    that an untrained network reproduces the default parameters (multipliers 2^0)."""
    rng = np.random.default_rng(seed)
    P = np.zeros(NPAR, dtype=np.float32)

    def glorot(o, i):
        a = math.sqrt(6.0 / (o + i))
        return rng.uniform(-a, a, size=(o, i)).astype(np.float32)
    P[0:768] = glorot(64, 12).ravel()
    P[832:4928] = glorot(64, 64).ravel()
    return P


def save(P, path=MODEL, **meta):
    np.savez(path, params=np.asarray(P, dtype=np.float32), **{k: np.asarray(v) for k, v in meta.items()})


def load(path=MODEL):
    with np.load(path, allow_pickle=False) as z:
        return np.asarray(z["params"], dtype=np.float32)


def install_learned(L=None, path=MODEL):
    """Upload the trained parameters and enable the policy for calls with use_policy = 1."""
    L = L or gqz.init()
    P = np.ascontiguousarray(load(path))
    r = L.gqz_mlp_policy(P.ctypes.data_as(C.c_void_p), 1)
    if r != 0:
        raise RuntimeError("gqz_mlp_policy failed: " + L.gqz_last_error().decode())
    return P


def disable(L=None):
    L = L or gqz.init()
    L.gqz_mlp_policy(None, 0)


def learned_params():
    """GqzParams that let the installed policy choose NS, NW and NIBBLE."""
    p = gqz.GqzParams()
    p.use_policy = 1
    return p


# ---- pencils (Hessenberg-triangular, generated on the device) ------------------------------------------
def make_pencil(kind, n, seed):
    """HT pencils of three families: 'ht' (testing/EXRAND1.f:243-279: N(0,1) upper parts,
    chi-distributed subdiagonal of H and diagonal of T), 'graded' (rows and columns of the 'ht' pencil
    scaled by 10^U(-2,2), which keeps the HT form and the eigenvalues' conditioning spread) and
    'singular' (3% of diag(T) set to zero: infinite eigenvalues)."""
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    H = gqz.colmajor(n); T = gqz.colmajor(n)
    Hs = torch.randn((n, n), dtype=torch.float64, device="cuda", generator=g)
    H.copy_(Hs.t()); del Hs
    H.triu_()
    Ts = torch.randn((n, n), dtype=torch.float64, device="cuda", generator=g)
    T.copy_(Ts.t()); del Ts
    T.triu_()
    rng = np.random.default_rng(seed)
    sub = np.sqrt(rng.chisquare(np.arange(n - 1, 0, -1)))
    dia = np.sqrt(rng.chisquare(np.arange(n, 0, -1)))
    idx = torch.arange(n - 1, device="cuda")
    H[idx + 1, idx] = torch.from_numpy(sub).cuda()
    ii = torch.arange(n, device="cuda")
    T[ii, ii] = torch.from_numpy(dia).cuda()
    if kind == "graded":
        dr = torch.from_numpy(10.0 ** rng.uniform(-2, 2, n)).cuda()
        dc = torch.from_numpy(10.0 ** rng.uniform(-2, 2, n)).cuda()
        H.mul_(dr[:, None]).mul_(dc[None, :])
        T.mul_(dr[:, None]).mul_(dc[None, :])
    elif kind == "singular":
        z = torch.from_numpy(rng.choice(n, size=max(1, (3 * n) // 100), replace=False)).cuda()
        T[z, z] = 0.0
    elif kind != "ht":
        raise ValueError(kind)
    return H, T


# ---- traces ---------------------------------------------------------------------------------------------
class PhasePolicy:
    """Host callback with an absolute AED window and a sweep-skip threshold that depends on the phase
    of the iteration: nib_early while the active block holds more than half of the pencil (feature 1,
    nh / n), nib_late afterwards.  NS keeps its default (the AED window bounds the shifts anyway)."""

    def __init__(self, L, nw, nib_early, nib_late):
        self.L = L

        def cb(user, feat, nh, ns, nw_, nib):
            nw_[0] = nw
            nib[0] = nib_early if feat[1] > 0.5 else nib_late
        self.fn = gqz.POLICY_FN(cb)

    def __enter__(self):
        self.L.gqz_set_policy(self.fn, None)
        return self

    def __exit__(self, *a):
        self.L.gqz_set_policy(gqz.POLICY_FN(0), None)


class StaticPolicy:
    """Host callback that multiplies the default NS, NW, NIBBLE (passed in by the library) by fixed
    factors (used to explore the parameter space while recording traces)."""

    def __init__(self, L, mns, mnw, mnib):
        self.L, self.m = L, (mns, mnw, mnib)

        def cb(user, feat, nh, ns, nw, nib):
            d_ns, d_nw, d_nib = ns[0], nw[0], nib[0]      # the defaults, passed in
            ns[0] = max(2, int(round(d_ns * self.m[0])) // 2 * 2)
            nw[0] = max(8, int(round(d_nw * self.m[1])))
            nib[0] = min(100, max(1, int(round(d_nib * self.m[2]))))
        self.fn = gqz.POLICY_FN(cb)

    def __enter__(self):
        self.L.gqz_set_policy(self.fn, None)
        return self

    def __exit__(self, *a):
        self.L.gqz_set_policy(gqz.POLICY_FN(0), None)


def run_traced(L, H0, T0, mult=None, learned=False, cap=100000, phase=None):
    """One QZ (JOB='S', COMPQ=COMPZ='I') of a copy of (H0, T0); returns (seconds, info, records,
    residuals), the residuals being PCOMPRES's [resA, resB, resQ/(n eps), resZ/(n eps)] of the result
    (computed after the timed call)."""
    import torch
    n = H0.shape[0]
    H = gqz.colmajor(n); T = gqz.colmajor(n); Q = gqz.colmajor(n); Z = gqz.colmajor(n)
    H.copy_(H0); T.copy_(T0)
    L.gqz_trace_enable(cap)
    prm = gqz.GqzParams()
    if mult is not None or learned or phase is not None:
        prm.use_policy = 1
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if phase is not None:
        with PhasePolicy(L, *phase):
            _, _, _, info, st = gqz.pdhgeqz("S", "I", "I", H, T, Q, Z, params=prm)
    elif mult is not None:
        with StaticPolicy(L, *mult):
            _, _, _, info, st = gqz.pdhgeqz("S", "I", "I", H, T, Q, Z, params=prm)
    else:
        _, _, _, info, st = gqz.pdhgeqz("S", "I", "I", H, T, Q, Z, params=prm)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    res = gqz.compres(H0, T0, H, T, Q, Z)
    recs = (gqz.GqzTraceRec * cap)()
    k = L.gqz_trace_get(recs, cap)
    rows = [(np.array(recs[i].feat[:], dtype=np.float32), recs[i].nh, recs[i].ns, recs[i].nw, recs[i].nibble,
             recs[i].nd, recs[i].t_iter) for i in range(k)]
    L.gqz_trace_enable(0)
    del H, T, Q, Z
    return dt, info, rows, res


def passes(res):
    """The backward-error and orthogonality bar every QZ result is held to."""
    return res[0] < 1e-13 and res[1] < 1e-13 and res[2] < 10.0 and res[3] < 10.0


# ---- training ---------------------------------------------------------------------------------------------
def train(L, X, Y, steps=3000, lr=3e-3, batch=256, seed=0, P0=None, log=None):
    """Fit the perceptron on (X, Y) with the library's gradient + Adam kernels (FP32, on the GPU).
    Returns (params, losses)."""
    import torch
    X = np.ascontiguousarray(X, dtype=np.float32); Y = np.ascontiguousarray(Y, dtype=np.float32)
    m = X.shape[0]
    P = torch.from_numpy(init_params(seed) if P0 is None else np.asarray(P0, dtype=np.float32)).cuda()
    G = torch.zeros(NPAR, dtype=torch.float32, device="cuda")
    M = torch.zeros_like(G); V = torch.zeros_like(G)
    dX = torch.from_numpy(X).cuda(); dY = torch.from_numpy(Y).cuda()
    loss = torch.zeros(1, dtype=torch.float32, device="cuda")
    rng = np.random.default_rng(seed)
    vp = lambda t: C.c_void_p(t.data_ptr())
    losses = []
    for step in range(1, steps + 1):
        sel = torch.from_numpy(rng.integers(0, m, size=min(batch, m))).cuda()
        xb = dX[sel].contiguous(); yb = dY[sel].contiguous()
        torch.cuda.synchronize()
        r = L.gqz_mlp_train_step_dev(vp(P), vp(G), vp(M), vp(V), step, C.c_float(lr), vp(xb), vp(yb),
                                     xb.shape[0], vp(loss))
        if r != 0:
            raise RuntimeError("gqz_mlp_train_step_dev failed: " + L.gqz_last_error().decode())
        if step % 100 == 0 or step == steps:
            L.gqz_sync()
            losses.append(float(loss.item()))
            if log:
                log("train step %d loss %.5f" % (step, losses[-1]))
    L.gqz_sync()
    return P.cpu().numpy(), losses
