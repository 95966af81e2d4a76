#!/usr/bin/env python3
"""Benchmark of the dynamic-ordering HSVD step at N = 4096 on one MI355X.

One "step" is one pass of the hot path (weights -> sort -> greedy select ->
2x2 HSVD -> rotations) on the N = 4096 real input of the C5 configuration
(Gaussian, column scales 10**U(-6,6), J half +1 / half -1).  The timed
region starts with Y already in HBM and covers K consecutive steps from the
start of the solve (device events around the graph launches).

Printed as one JSON line:
  value        algorithmic HBM GB/s of the step (bytes defined below / time)
  ms_per_step  device time per step
  roofline     the dominant kernel (the hyperbolic pair weights: FP64 VALU)
  roofline_rotate   the rotation kernel against HBM
  cpu_baseline the reference (oracle/_ref/plain/djk.exe, built from the
               reference's own sources) on the same input, host cores
  full_solve_c2     a complete HSVD at N = 1024 (C2), wall time on the device

Algorithmic bytes of one step (N, NN = N(N-1)/2, IT finite weights, SL
pivots, b2/b4 pivots with a row/column transform):
  (a) weights  8 N^2 (read A) + 8 NN (one key per pair)
  (b) sort     24 NN (read + write one 8-byte key and 4-byte value)
  (c) select   12 IT
  (d) 2x2      112 SL
  (e) rotate   16 N^2 (A read + written once) + 32 N (b2 + b4) (U, Z columns)
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def step_bytes(n, it, sl, b2, b4):
    nn = n * (n - 1) // 2
    a = 8.0 * n * n + 8.0 * nn
    b = 24.0 * nn
    c = 12.0 * it
    d = 112.0 * sl
    e = 16.0 * n * n + 32.0 * n * (b2 + b4)
    return a + b + c + d + e, e


def cpu_reference_sample(y, j, n2, budget_s, threads):
    """Runs the reference djk.exe on the same input for at most budget_s and
    returns the per-step seconds of the steps it completed (its own CSV)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "plain", "djk.exe")
    if not os.path.exists(exe):
        return None
    from jkogb_amd import inputs
    n = y.shape[0]
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "c5")
        inputs.write_yj(fn, y, j)
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close")
        p = subprocess.Popen([exe, fn, str(n), str(n2)], stdout=subprocess.PIPE, stderr=subprocess.PIPE, env=env,
                             text=True)
        t0 = time.time()
        lines = []
        try:
            import selectors
            sel = selectors.DefaultSelector()
            sel.register(p.stderr, selectors.EVENT_READ)
            buf = ""
            while time.time() - t0 < budget_s and p.poll() is None:
                for key, _ in sel.select(timeout=1.0):
                    chunk = os.read(key.fileobj.fileno(), 65536).decode()
                    buf += chunk
                    while "\n" in buf:
                        line, buf = buf.split("\n", 1)
                        lines.append(line)
        finally:
            if p.poll() is None:
                p.kill()
            p.wait()
    steps = []
    for line in lines:
        f = [x.strip() for x in line.split(",")]
        if len(f) == 8 and re.match(r"^-?\d+$", f[0]):
            steps.append((int(f[0]), int(f[1]), float(f[2]) + float(f[3])))
    return steps


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=4096)
    ap.add_argument("--fused", action="store_true", help="FMAD = fma (default: the reference's unfused build)")
    ap.add_argument("--profile-steps", type=int, default=3)
    ap.add_argument("--cpu-budget", type=float, default=40.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-full", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the step is single-GPU; see DESIGN.md for how it "
              "would shard); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from jkogb_amd import inputs, jkg
    n = args.n
    y, j = inputs.make("c5", n=n, seed=0)
    n2 = n // 2
    ctx = jkg.Context(0)
    # warmup: captures the step graph and touches every buffer
    ctx.dsolve(y, j, n2, max_steps=args.warmup, fused=args.fused)
    t0 = time.time()
    r = ctx.dsolve(y, j, n2, max_steps=args.steps, fused=args.fused)
    wall = time.time() - t0
    k = r["steps"] if r["info"] == args.steps else r["steps"]
    ms_step = 1e3 * r["seconds"] / max(k, 1)
    bytes_step, rot_bytes = step_bytes(n, r["sum_it"] / k, r["sum_sl"] / k, r["sum_b2"] / k, r["sum_b4"] / k)
    gbs = bytes_step / (ms_step * 1e-3) / 1e9

    # per-stage device times (each stage launched and timed on the library's stream)
    prof = ctx.dprofile(y, j, n2, steps=args.profile_steps, fused=args.fused)
    ps = prof["steps"]
    st_ms = {s: v / ps for s, v in prof["stage_ms"].items()}
    flops_w = 14.0 * (n - 2) * prof["sum_abod"] / ps
    tflops_w = flops_w / (st_ms["weights"] * 1e-3) / 1e12
    _, rot_b = step_bytes(n, 0, 0, prof["sum_b2"] / ps, prof["sum_b4"] / ps)
    rot_gbs = rot_b / (st_ms["rotate"] * 1e-3) / 1e9

    full = None
    if not args.no_full:
        yc, jc = inputs.make("c2", n=1024, seed=0)
        rc = ctx.dsolve(yc, jc, 512, fused=args.fused)
        full = {"workload": "C2: N=1024 real, columns 10**U(-4,4), J balanced, N_2=512", "steps": rc["info"],
                "seconds": round(rc["seconds"], 4), "ms_per_step": round(1e3 * rc["seconds"] / max(rc["steps"], 1), 4)}
    ctx.close()

    cpu = None
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        steps = cpu_reference_sample(y, j, n2, args.cpu_budget, threads)
        if steps:
            per = float(np.mean([s[2] for s in steps]))
            cpu = {"value": round(bytes_step / per / 1e9, 4), "unit": "GB/s", "cores": threads, "kind": "reference",
                   "sample": "the reference djk.exe (flang build of its own sources) on the same N=4096 input, "
                             "%d completed step(s) within %.0f s; %.3f s per step (its BUILDs+TRANSFs CSV)"
                             % (len(steps), args.cpu_budget, per),
                   "s_per_step": round(per, 4)}
        else:
            cpu = {"value": None, "unit": "GB/s", "cores": threads, "kind": "reference",
                   "sample": "no step of the reference completed within %.0f s at N=4096" % args.cpu_budget}

    out = {
        "metric": "HSVD wall time + HBM GB/s at N=4096, mixed/learned vs plain FP64, same accuracy",
        "value": round(gbs, 3),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": k,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "C5 input at N=4096: real Gaussian Y, column scales 10**U(-6,6), J 2048 (+1) / "
                               "2048 (-1) permuted, N_2=2048; plain FP64 solver (faithful weights, %s FMAD)"
                               % ("fused" if args.fused else "unfused"),
                   "n": n, "n2": n2, "solver": "plain-fp64"},
        "roofline": {"bound": "mfma", "kernel": "k_hyp_weights (FP64 VALU, ABODNDF2 per pair)",
                     "achieved": round(tflops_w, 3), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(tflops_w / FP64_PEAK_TFLOPS, 4), "traffic": None},
        "roofline_rotate": {"bound": "hbm", "kernel": "k_rotate", "achieved": round(rot_gbs, 1),
                            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(rot_gbs / HBM_PEAK_GBS, 4),
                            "traffic": None, "bytes_per_launch": int(rot_b),
                            "ms_per_launch": round(st_ms["rotate"], 4)},
        "stage_ms_per_step": {s: round(v, 4) for s, v in st_ms.items()},
        "cpu_baseline": cpu,
        "full_solve_c2": full,
        "host_wall_s": round(wall, 3),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
