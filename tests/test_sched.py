"""CPU tests of the sweep's chain scheduler (the host side of PDHGEQZ5's concurrent windows,
PDHGEQZ5.f:334-346 NUMWIN), replayed through gqz_sched_replay without a device.

With the overlapped updates, round r of a sweep is one chase launch, and B(r) is the rest of round r's
window updates on the second stream:
  * far-left: the rows of each window, right of its near columns;
  * right: the rows above each window, in its columns.
chase(r) runs concurrently with B(r-1), and waits for B(r-2).  A chain's near block N stays pending
from its window's round until its next launch applies it.  N is rows [p, p + w) x columns
[p + w, p + w + nc).

The checks are the conditions under which this order gives the same products as the serial one:
  1. chase(r) touches nothing B(r-1) touches (unless the scheduler drained the second stream first);
  2. no update and no other chain's chase touches a pending N: a right update would mix columns with
     and without the chain's U, and another left update does not commute with it;
  3. every chain sees the window sequence of a lone chain, and applies exactly the near block its
     previous window left;
  4. the windows of a round are disjoint and ordered.
A checker that is fed a planted violation must report it."""
import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "gevp-pdhgeqz_amd", "libgqz.so")
WMAX = 64


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("libgqz.so not built")
    L = C.CDLL(LIB)
    L.gqz_sched_replay.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int]
    return L


def replay(L, n, ktop, kbot, nb, ovl):
    nb = np.ascontiguousarray(nb, dtype=np.int32)
    cap = 400000
    out = np.zeros(8 * cap, dtype=np.int32)
    r = L.gqz_sched_replay(n, ktop, kbot, len(nb), nb.ctypes.data_as(C.c_void_p), int(ovl),
                           out.ctypes.data_as(C.c_void_p), cap)
    assert 0 < r <= cap, r
    return out[:8 * r].reshape(-1, 8)


def _meet(a, b):
    """rectangles (r0, r1, c0, c1), half open"""
    return a[0] < b[1] and b[0] < a[1] and a[2] < b[3] and b[2] < a[3]


def check(rec, n, ktop, kbot, nb, ovl):
    """Returns a list of violations (empty when the schedule is safe)."""
    bad = []
    rounds = {}
    drained = set()
    for row in rec:
        rnd, c, p, w, steps, last, nc, near = (int(x) for x in row)
        if c < 0:
            drained.add(rnd)
            continue
        rounds.setdefault(rnd, []).append((c, p, w, steps, last, nc, near))
    prev = {}                                    # chain -> its previous window
    chase, upd, pend = {}, {}, []                # round -> rectangles; pending blocks
    for rnd in sorted(rounds):
        ws = rounds[rnd]
        # 4. disjoint and ordered: a chain ahead (smaller index) lies below
        srt = sorted(ws, key=lambda x: x[0])
        for a, b in zip(srt, srt[1:]):
            if not b[1] + b[2] <= a[1]:
                bad.append(("overlapping windows", rnd, a[0], b[0]))
        ch_r, up_r = [], []
        for c, p, w, steps, last, nc, near in ws:
            # 3. the lone chain's window sequence and near blocks
            if c not in prev:
                if p != ktop or near != 0:
                    bad.append(("intro window", rnd, c, p))
            else:
                pp, pw, pst, pnc = prev[c]
                if p != pp + pst:
                    bad.append(("window sequence", rnd, c, p, pp + pst))
                if near != pnc or (near and pp + pw + near != p + w):
                    bad.append(("near block", rnd, c, near, pnc))
            if w != min(WMAX, kbot - p + 1):
                bad.append(("window order", rnd, c, w))
            if last != int(p + w - 1 == kbot):
                bad.append(("last flag", rnd, c))
            if last and steps != w:
                bad.append(("push-off steps", rnd, c))
            if not last and steps != w - 3 * nb[c] - 1:
                bad.append(("steps", rnd, c, steps))
            if ovl and not last:
                np_ = p + steps
                want = min(np_ + min(WMAX, kbot - np_ + 1), n) - (p + w)
                if nc != want:
                    bad.append(("near columns", rnd, c, nc, want))
            if not ovl and (nc or near):
                bad.append(("near without overlap", rnd, c))
            ch_r.append((c, (p, p + w, p, p + w)))
            if near:
                pp, pw = prev[c][0], prev[c][1]
                ch_r.append((c, (pp, pp + pw, pp + pw, pp + pw + near)))
            up_r.append((c, (p, p + w, p + w + nc, n)))         # far-left
            up_r.append((c, (0, p, p, p + w)))                  # right (rows above the window)
            prev[c] = (p, w, steps, nc)
        chase[rnd], upd[rnd] = ch_r, up_r
        # 1. chase(r) against B(r-1)
        if ovl and rnd - 1 in upd and rnd not in drained:
            for c1, a in ch_r:
                for c2, b in upd[rnd - 1]:
                    if _meet(a, b):
                        bad.append(("chase meets the previous round's updates", rnd, c1, c2))
    # 2. pending near blocks: created at round r by chain c, applied at the chain's next launch r2
    if ovl:
        nxt = {}
        for rnd in sorted(rounds, reverse=True):
            for c, p, w, steps, last, nc, near in rounds[rnd]:
                if nc:
                    pend.append((c, rnd, nxt.get(c), (p, p + w, p + w, p + w + nc)))
                nxt[c] = rnd
        for c, r, r2, N in pend:
            if r2 is None:
                bad.append(("near block never applied", c, r))
                continue
            for t in range(r, r2):
                for c2, b in upd.get(t, []):
                    if (c2 != c or t != r) and _meet(N, b):
                        bad.append(("update meets a pending near block", c, r, t, c2))
            for t in range(r + 1, r2 + 1):
                for c2, a in chase.get(t, []):
                    if c2 != c and _meet(N, a):
                        bad.append(("chase meets a pending near block", c, r, t, c2))
    # every chain ends with its push-off window at kbot
    for c in range(len(nb)):
        if c not in prev or prev[c][0] + prev[c][1] - 1 != kbot:
            bad.append(("chain not finished", c))
    return bad


CASES = [
    # (n, ktop, kbot, bulges per chain): the learned-parameter sweeps (chains of 10 shifts and a
    # remainder), PIPARMQ's 512 shifts in chains of 20, short chains, single bulges, sub-blocks
    (16000, 0, 15999, [5] * 6 + [2]),
    (16000, 0, 15999, [10] * 25 + [6]),
    (16000, 3000, 9000, [2] * 16),
    (8000, 0, 7999, [1] * 20),
    (6000, 17, 5980, [20] * 3 + [3]),
    (1500, 0, 1499, [5] * 6 + [2]),
    (3000, 2800, 2999, [5, 5, 5, 2]),
    (3000, 2900, 2999, [5, 5]),
    (2100, 0, 2099, [7, 3, 9, 1, 5, 20, 2, 11]),
]


@pytest.mark.parametrize("ovl", [1, 0])
@pytest.mark.parametrize("n,ktop,kbot,nb", CASES)
def test_schedule_is_safe(lib, n, ktop, kbot, nb, ovl):
    rec = replay(lib, n, ktop, kbot, nb, ovl)
    assert check(rec, n, ktop, kbot, nb, ovl) == []
    # the replay really runs chains side by side (the checks are not vacuous)
    if len(nb) > 1 and kbot - ktop > 400:
        per_round = np.bincount(rec[rec[:, 1] >= 0, 0])
        assert per_round.max() >= min(len(nb), 3)


def test_random_chain_mixes(lib):
    rng = np.random.default_rng(5)
    for _ in range(40):
        n = int(rng.integers(200, 5000))
        ktop = int(rng.integers(0, n // 3))
        kbot = int(rng.integers(min(ktop + 10, n - 1), n))
        nb = list(rng.integers(1, 21, size=int(rng.integers(1, 17))))
        for ovl in (0, 1):
            rec = replay(lib, n, ktop, kbot, nb, ovl)
            assert check(rec, n, ktop, kbot, nb, ovl) == [], (n, ktop, kbot, nb, ovl)


def test_checker_reports_planted_violations(lib):
    n, ktop, kbot, nb = 16000, 0, 15999, [5] * 6 + [2]
    rec = replay(lib, n, ktop, kbot, nb, 1)
    assert check(rec, n, ktop, kbot, nb, 1) == []
    # a chain one round early: its window reaches into the rows of the chain ahead or its updates
    k = int(np.flatnonzero((rec[:, 1] == 1) & (rec[:, 0] > 10))[0])
    bad = rec.copy()
    bad[k, 0] -= 1
    assert check(bad, n, ktop, kbot, nb, 1) != []
    # a near block applied one column short
    k = int(np.flatnonzero((rec[:, 1] == 0) & (rec[:, 7] > 0))[3])
    bad = rec.copy()
    bad[k, 7] -= 1
    assert any(v[0] == "near block" for v in check(bad, n, ktop, kbot, nb, 1))
    # two chains packed closer than the pending near blocks allow: chain 1 moved 48 rows down in
    # every round where it runs beside chain 0
    bad = rec.copy()
    for k in np.flatnonzero(bad[:, 1] == 1):
        if bad[k, 2] + bad[k, 3] + 48 <= kbot - 200:
            bad[k, 2] += 48
    assert check(bad, n, ktop, kbot, nb, 1) != []
