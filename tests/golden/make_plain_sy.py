#!/usr/bin/env python3
"""Writes the hyperbolic singular values of complete plain-solver runs at full size, the
accuracy references of the fast solvers' full-size tests.  The plain solver's arithmetic and
trajectory are the reference's own (bit-identical to djk.exe / zjk.exe in the parity tests,
for both FMAD policies); these sizes take the reference itself days on the CPU, so the runs
are made on the GPU and kept as vectors.

  c5      solve_c5_n4096_plain_sy.npy: N = 4096 real input of the learned-controller
          configuration (jkogb_amd.inputs.make("c5", seed=0): Gaussian Y, column scales
          10**U(-6,6), J 2048/2048), unfused FMAD (the reference's default build).  At this
          size the reference's termination (no pivot or no effective transform) is not
          reached: after about 47 000 steps the transforms only move entries at the rounding
          level (NORMF2D falls to 1e-74), so the run stops after 52 000 steps, where NORMF2D is
          5.8e-75.  About 12 minutes on one MI355X.
  c5_fma  solve_c5_n4096_plain_fma_sy.npy: the same with FMAD = fma (the reference's fused
          build), also 52 000 steps: how far two valid builds of the reference land apart.
  c3      solve_c3_n2048_plain_sy.npy: N = 2048 complex input of the complex configuration
          (inputs.make("c3", seed=0): Re, Im ~ N(0, 1/2), J 1024/1024), unfused, run to the
          reference's own termination: 23 019 steps, about 9 minutes.

    python3 tests/golden/make_plain_sy.py c5 c5_fma c3 [--out DIR]   # on a machine with the GPU
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
import torch  # noqa: E402,F401  (torch's HIP runtime first: the library links the same soname)
from jkogb_amd import inputs, jkg  # noqa: E402

RUNS = {
    "c5": dict(cfg="c5", n=4096, fused=False, steps=52000, file="solve_c5_n4096_plain_sy.npy"),
    "c5_fma": dict(cfg="c5", n=4096, fused=True, steps=52000, file="solve_c5_n4096_plain_fma_sy.npy"),
    "c3": dict(cfg="c3", n=2048, fused=False, steps=-1, file="solve_c3_n2048_plain_sy.npy"),
}


def run(ctx, name, out_dir):
    r = RUNS[name]
    n = r["n"]
    y, j = inputs.make(r["cfg"], n=n, seed=0)
    t0 = time.time()
    last = [0]

    def prog(step, done):
        if done or step - last[0] >= 2000:
            last[0] = step
            print("%s: plain step %d (%.0f s)" % (name, step, time.time() - t0), flush=True)

    cap = r["steps"] + 1 if r["steps"] > 0 else 64 * n + 64
    o = ctx.dsolve(y, j, n // 2, fused=r["fused"], max_steps=r["steps"], trace_cap=cap, progress=prog)
    np.save(os.path.join(out_dir, r["file"]), o["SY"])
    print(json.dumps({"run": name, "info": o["info"], "steps": o["steps"], "seconds": o["seconds"],
                      "tw_last": float(o["trace"]["tw"][-1])}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("runs", nargs="+", choices=sorted(RUNS))
    ap.add_argument("--out", default=HERE)
    a = ap.parse_args()
    os.makedirs(a.out, exist_ok=True)
    with jkg.Context(0) as ctx:
        for name in a.runs:
            run(ctx, name, a.out)


if __name__ == "__main__":
    main()
