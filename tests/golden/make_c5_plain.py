#!/usr/bin/env python3
"""Writes tests/golden/solve_c5_n4096_plain_sy.npy: the hyperbolic singular values of the
N = 4096 real input of the learned-controller configuration (jkogb_amd.inputs.make("c5",
seed=0): Gaussian Y, column scales 10**U(-6,6), J 2048/2048) from the plain solver, whose
arithmetic and trajectory are the reference's own (bit-identical to djk.exe in the parity
tests; unfused FMAD, the reference's default build).  At this size the reference's
termination (no pivot or no effective transform) is not reached: after about 47 000 steps
the transforms only move entries at the rounding level (NORMF2D falls to 1e-74), so the run
stops after 52 000 steps, where NORMF2D is 5.8e-75.  On one MI355X it takes about 12 minutes
(the reference itself on 16 CPU cores: about 2 days).

    python3 tests/golden/make_c5_plain.py            # on a machine with the GPU
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from jkogb_amd import inputs, jkg  # noqa: E402

STEPS = 52000


def main():
    n = 4096
    y, j = inputs.make("c5", n=n, seed=0)
    t0 = time.time()
    last = [0]

    def prog(step, done):
        if done or step - last[0] >= 2000:
            last[0] = step
            print("plain step %d (%.0f s)" % (step, time.time() - t0), flush=True)

    with jkg.Context(0) as ctx:
        r = ctx.dsolve(y, j, n // 2, max_steps=STEPS, trace_cap=STEPS + 1, progress=prog)
    np.save(os.path.join(HERE, "solve_c5_n4096_plain_sy.npy"), r["SY"])
    print(json.dumps({"steps": r["steps"], "seconds": r["seconds"], "tw_last": float(r["trace"]["tw"][-1])}))


if __name__ == "__main__":
    main()
