"""Generate QZ traces on the GPU and train the learned parameter policy (writes mlp_policy.npz).

    python gevp-pdhgeqz_amd/train_policy.py [--sizes 2000,4000,8000] [--families ht,graded,singular]
                                            [--seeds 1] [--steps 3000] [--out PATH]

For every training pencil the QZ runs once per multiplier triple of the grid (NS x NW x NIBBLE
relative to the PIPARMQ defaults, which the QZ evaluates once per call at the order of the active
block).  The first AED window W* of the fastest run is the target: every state recorded in the runs
of that pencil is labelled log2(W* / NW_def(n)) for NW and 0 for NS and NIBBLE.  Larger pencils (--large-sizes, one family, a reduced grid) extend the
range of active orders seen in training.  The perceptron is then fitted with
the library's own FP32 gradient and Adam kernels.  Each run's backward error and orthogonality are
checked (PCOMPRES on the device): a run that fails them, or returns INFO /= 0, is never a target.
"""
import argparse
import itertools
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import gqz  # noqa: E402
import gqz_policy as GP  # noqa: E402

# NS beyond the shifts an AED window delivers has no effect, so the grid spans the window size: from
# windows that fit the LDS-resident AED kernel to windows with a nested QZ
GRID_NS = (1.0,)
GRID_NW = (0.03125, 0.0625, 0.125, 0.25)
GRID_NIB = (1.0,)
GRID_NW_LARGE = (0.0625, 0.125)


def log(*a):
    print(*a, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", default="2000,4000,8000")
    ap.add_argument("--families", default="ht,graded,singular")
    ap.add_argument("--seeds", default="1")
    ap.add_argument("--large-sizes", default="12000,16000")
    ap.add_argument("--large-families", default="ht")
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--out", default=GP.MODEL)
    args = ap.parse_args()
    import torch
    L = gqz.init(0)
    GP.disable(L)
    sizes = [int(x) for x in args.sizes.split(",")]
    fams = args.families.split(",")
    seeds = [int(x) for x in args.seeds.split(",")]
    X, Y, summary = [], [], []
    cases = [(fam, n, seed, GRID_NW) for fam, n, seed in itertools.product(fams, sizes, seeds)]
    large = [int(x) for x in args.large_sizes.split(",") if x]
    cases += [(fam, n, seeds[0], GRID_NW_LARGE) for fam in args.large_families.split(",") for n in large]
    for fam, n, seed, grid_nw in cases:
        H0, T0 = GP.make_pencil(fam, n, 1000 * seed + n)
        runs = []
        for mult in itertools.product(GRID_NS, grid_nw, GRID_NIB):
            dt, info, rows, res = GP.run_traced(L, H0, T0, mult=mult)
            runs.append((dt, info, mult, rows, res))
            log("trace %s n=%d seed=%d mult=%s: %.2f s info=%d iterations=%d resA=%.1e resQ=%.2f"
                % (fam, n, seed, mult, dt, info, len(rows), res[0], res[2]))
        ok = [r for r in runs if r[1] == 0 and GP.passes(r[4])]
        if not ok:
            log("no run of %s n=%d seed=%d passed the checks: pencil skipped" % (fam, n, seed))
            continue
        best = min(ok, key=lambda r: r[0])
        dflt = [r for r in ok if r[2] == (1.0, 1.0, 1.0)]
        wstar = best[3][0][3] if best[3] else GP.piparmq_nw(n)     # first AED window of the fastest run
        for r in ok:                     # every recorded state of every run of this pencil
            for row in r[3]:
                t_nw = float(np.clip(np.log2(wstar / GP.piparmq_nw(n)), -5.0, 1.0))
                X.append(row[0]); Y.append(np.array([0.0, t_nw, 0.0], dtype=np.float32))
        summary.append(dict(family=fam, n=n, seed=seed, best=list(best[2]), best_s=best[0],
                            default_s=dflt[0][0] if dflt else None))
        log("best %s n=%d: mult=%s %.2f s (default %.2f s)" % (fam, n, best[2], best[0],
                                                               dflt[0][0] if dflt else float("nan")))
        del H0, T0
        torch.cuda.empty_cache()
    X = np.array(X, dtype=np.float32); Y = np.array(Y, dtype=np.float32)
    log("dataset: %d rows" % X.shape[0])
    t0 = time.time()
    P, losses = GP.train(L, X, Y, steps=args.steps, log=log)
    log("trained in %.1f s, final loss %.5f" % (time.time() - t0, losses[-1]))
    GP.save(P, args.out, X=X, Y=Y, losses=np.array(losses), summary=json.dumps(summary))
    log("wrote", args.out)


if __name__ == "__main__":
    main()
