"""Training of the mixed solver's step controller (the 16-64-64-2 MLP).

Labels come from a search, per training problem, over piecewise-constant
schedules: the run is split into three phases by log10(TW / TW_0) (> -3,
> -7, below; TW is the weight removed by a step), and coordinate descent picks
for each phase the action (N_2 fraction, log10 threshold) from a small grid
that minimises the device time to convergence.  A schedule is admissible only
if its singular values stay within 4x the spread between the reference's two
FMA builds (+64 N eps) of the plain solver's.  The features recorded along the
best schedule's run, paired with the action of the following step, form the
training set; the MLP is fitted with the library's own GPU kernels
(jkg_ctl_fit).  Then, for a few rounds, the network itself drives the training
problems and the states it visits are labelled with the best schedule's action
for their phase (imitation with data aggregation), and it is refitted on the
union.  The network kept is the candidate with the smallest total device time
on validation problems (other seeds) among those that meet the accuracy bound.  Features 9 (the size), 10 and 11 (the action just used) are
zeroed in training and their first-layer weights are zeroed after it: the
network cannot learn to repeat its last action, nor extrapolate in N.

Usage (on a GPU):  python -m jkogb_amd.ctl_train --out jkogb_amd/ctl_params.npy
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PARAMS = os.path.join(HERE, "ctl_params.npy")

ACTIONS = [(1.0, -16.0), (1.0, -12.0), (1.0, -9.0), (1.0, -6.0), (0.75, -16.0), (0.75, -12.0), (0.75, -9.0),
           (0.875, -12.0)]
# features the network must not use: 9 (the size: training sizes are smaller than the
# deployment's), 10 and 11 (the action just used)
MASKED = (9, 10, 11)
DEFAULT = (1.0, -16.0)
BOUNDS = (-3.0, -7.0)
EPS = np.finfo(np.float64).eps
IN, H, OUT = 16, 64, 2
NPARAM = H * IN + H + H * H + H + OUT * H + OUT


def table(acts):
    t = [BOUNDS[0], BOUNDS[1]]
    for a in acts:
        t += [a[0], a[1]]
    return np.array(t, np.float32)


def init_params(seed=0):
    """He-style initial weights in the blob layout of jkg.h (W1, b1, W2, b2, W3, b3)."""
    rng = np.random.default_rng(seed)
    parts = [rng.standard_normal((H, IN)) * np.sqrt(2.0 / IN), np.zeros(H), rng.standard_normal((H, H)) *
             np.sqrt(2.0 / H), np.zeros(H), rng.standard_normal((OUT, H)) * np.sqrt(1.0 / H), np.array([0.9, 0.1])]
    return np.concatenate([p.ravel() for p in parts]).astype(np.float32)


def problems(sizes, seeds):
    from jkogb_amd import inputs
    out = []
    for n in sizes:
        for cfg in ("c1", "c2", "c5"):
            for s in seeds:
                Y, J = inputs.make(cfg, n=n, seed=s)
                out.append(("%s n=%d seed=%d" % (cfg, n, s), Y, J))
    return out


def accuracy_guard(ctx, Y, J, n2):
    plain = ctx.dsolve(Y, J, n2)
    fused = ctx.dsolve(Y, J, n2, fused=True)
    n = Y.shape[0]
    spread = float(np.max(np.abs(fused["SY"] - plain["SY"]) / np.abs(plain["SY"])))
    return plain, 4 * spread + 64 * n * EPS


def search(ctx, Y, J, n2, log):
    plain, bound = accuracy_guard(ctx, Y, J, n2)

    def run(acts, trace=False):
        out = ctx.dsolve_mixed(Y, J, n2, ctl_table=table(acts), ctl_trace=trace)
        rel = float(np.max(np.abs(out["SY"] - plain["SY"]) / np.abs(plain["SY"])))
        return out, rel <= bound

    best = [DEFAULT, DEFAULT, DEFAULT]
    base, ok = run(best)
    t_base = base["seconds"]
    t_best = t_base
    for ph in range(3):
        for a in ACTIONS:
            if a == best[ph]:
                continue
            cand = list(best)
            cand[ph] = a
            out, ok = run(cand)
            if ok and out["seconds"] < 0.995 * t_best:
                best, t_best = cand, out["seconds"]
    final, _ = run(best, trace=True)
    log("    schedule %s: %.3f s vs %.3f s default (x%.3f), plain %.3f s"
        % (best, final["seconds"], t_base, t_base / final["seconds"], plain["seconds"]))
    return best, final["ctl"], t_base, final["seconds"]


def phase_action(sched, lrel):
    ph = 0 if lrel > BOUNDS[0] else (1 if lrel > BOUNDS[1] else 2)
    return sched[ph]


def samples(ctl, sched=None):
    """(features after step s, normalised action of step s+1) for the controlled steps; with
    sched, the action is the schedule's for the phase of the state (relabelling)."""
    X, T = [], []
    for s in range(len(ctl) - 1):
        f = ctl[s, :16].astype(np.float32).copy()
        if f[1] * 20 <= -10:  # the fallback region: the plain ordering, not the controller
            continue
        if sched is None:
            n2f, lt = float(ctl[s + 1, 16]), float(ctl[s + 1, 17])
        else:
            n2f, lt = phase_action(sched, float(f[1]) * 20)
        f[list(MASKED)] = 0.0
        X.append(f)
        T.append([(n2f - 0.75) / 0.25, (lt + 16.0) / 10.0])
    return np.array(X, np.float32).reshape(-1, IN), np.array(T, np.float32).reshape(-1, OUT)


def validate(ctx, cands, vprobs, log):
    """Total device time of each candidate (None: no controller) over the validation problems;
    a candidate that breaks the accuracy bound on any problem is out."""
    tot = [0.0] * len(cands)
    ok = [True] * len(cands)
    for name, Y, J in vprobs:
        n2 = Y.shape[0] // 2
        plain, bound = accuracy_guard(ctx, Y, J, n2)
        for i, p in enumerate(cands):
            out = ctx.dsolve_mixed(Y, J, n2) if p is None else ctx.dsolve_mixed(Y, J, n2, ctl_params=p)
            rel = float(np.max(np.abs(out["SY"] - plain["SY"]) / np.abs(plain["SY"])))
            tot[i] += out["seconds"]
            ok[i] = ok[i] and rel <= bound
        log("  validation %s: %s" % (name, ", ".join("%.3f" % t for t in tot)))
    return tot, ok


def train(ctx, X, T, W, epochs, seed, log):
    p0 = init_params(seed)
    p, hist = ctx.ctl_fit(X, T, p0, W=W, epochs=epochs, batch=256, lr=2e-3, weight_decay=1e-5, seed=seed)
    # the network must not depend on the masked features
    w1 = p[:H * IN].reshape(H, IN)
    w1[:, list(MASKED)] = 0.0
    loss, _, _ = ctx.ctl_grad(X, T, p, W)
    log("  fit: %d samples, loss %.4g -> %.4g (after zeroing the masked features' weights: %.4g)"
        % (len(X), hist[0], hist[-1], loss))
    return p


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=PARAMS)
    ap.add_argument("--sizes", default="256,384,512")
    ap.add_argument("--seeds", default="101,102")
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--rounds", type=int, default=2, help="imitation rounds with data aggregation")
    ap.add_argument("--val-seeds", default="201")
    ap.add_argument("--val-size", type=int, default=384)
    ap.add_argument("--report", default=None, help="write a JSON summary of the search here")
    args = ap.parse_args(argv)
    from jkogb_amd import jkg

    def log(msg):
        print(msg, flush=True)

    sizes = [int(x) for x in args.sizes.split(",")]
    seeds = [int(x) for x in args.seeds.split(",")]
    ctx = jkg.Context(0)
    Xs, Ts, Ws, rep = [], [], [], []
    t0 = time.time()
    train_probs = problems(sizes, seeds)
    scheds = []
    for name, Y, J in train_probs:
        log("%s (%.0f s)" % (name, time.time() - t0))
        n2 = Y.shape[0] // 2
        best, ctl, tb, tf = search(ctx, Y, J, n2, log)
        scheds.append(best)
        X, T = samples(ctl)
        if len(X):
            Xs.append(X)
            Ts.append(T)
            Ws.append(np.full(len(X), 1.0 / len(X), np.float32))
        rep.append({"problem": name, "schedule": best, "seconds_default": tb, "seconds_best": tf})

    def fit(seed):
        X, T, W = np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Ws)
        return train(ctx, X, T, W, args.epochs, seed, log)

    cands = [fit(1)]
    for r in range(args.rounds):  # the network drives, the schedules label
        for (name, Y, J), sched in zip(train_probs, scheds):
            out = ctx.dsolve_mixed(Y, J, Y.shape[0] // 2, ctl_params=cands[-1], ctl_trace=True)
            X, T = samples(out["ctl"], sched)
            if len(X):
                Xs.append(X)
                Ts.append(T)
                Ws.append(np.full(len(X), 1.0 / len(X), np.float32))
        log("round %d (%.0f s)" % (r + 1, time.time() - t0))
        cands.append(fit(2 + r))
    vprobs = problems([args.val_size], [int(x) for x in args.val_seeds.split(",")])
    tot, ok = validate(ctx, [None] + cands, vprobs, log)
    good = [i for i in range(1, len(tot)) if ok[i] and tot[i] < tot[0]]
    pick = min(good, key=lambda i: tot[i]) if good else None
    if pick is None:
        log("no candidate beats the mixed solver without a controller on validation; keeping the best-fitting one")
        pick = len(cands)
    p = cands[pick - 1]
    log("picked candidate %d: validation %.3f s vs %.3f s without a controller" % (pick, tot[pick], tot[0]))
    np.save(args.out, p)
    # the same parameters as raw little-endian FP32, for the drop-in programs (JKG_SOLVER=learned)
    np.ascontiguousarray(p, dtype="<f4").tofile(os.path.splitext(args.out)[0] + ".f32")
    log("wrote %s" % args.out)
    rep.append({"validation_seconds": [float(t) for t in tot], "validation_ok": [bool(o) for o in ok],
                "picked": int(pick)})
    if args.report:
        with open(args.report, "w") as f:
            json.dump(rep, f, indent=1)
    ctx.close()


if __name__ == "__main__":
    sys.exit(main())
