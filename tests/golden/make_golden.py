#!/usr/bin/env python3
"""Writes the golden vectors under tests/golden/ from the Fortran reference
built from its own sources by oracle/ref.mk (oracle/_ref/, this container
only).  The reference ships no test vectors of its own, so these are the
only known answers for the bits of the 2x2 HSVD, the weights, the sort and
selection, and whole trajectories.

    make -f oracle/ref.mk && python3 tests/golden/make_golden.py

Fixtures (numpy .npz, no pickles):
  solve_<case>.npz   Y, J, n2, id_ncp, id_tru -> YU, ZZ, SY, TT, PN, steps and
                     the per-step CSV trace (djk.exe/zjk.exe, 1 OpenMP thread so
                     the NORMF2D sums run in pivot order)
  build_<case>.npz   the debug build's DSTEP_BUILD.<k> dumps for the first steps:
                     the sorted (W,P,Q,B) list and the chosen (position,W,P,Q,B)
  kat_dhsvd2.npz     dtest.exe 2x2 known answers (INFO, U, V, Z, E)
  kat_zhsvd2.npz     ztest.exe 2x2 known answers
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from jkogb_amd import inputs  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref")
ENV = dict(os.environ, OMP_NUM_THREADS="1")

# name, config, n, n2, seed, complex, variant, id_ncp, id_tru
SOLVE_CASES = [
    ("c1_r64", "c1", 64, 32, 1, False, "plain", 0, 0),
    ("c1_r64_fma", "c1", 64, 32, 1, False, "fma", 0, 0),
    ("c2_r32", "c2", 32, 16, 2, False, "plain", 0, 0),
    ("c5_r40", "c5", 40, 20, 3, False, "plain", 0, 0),
    ("c1_r33_n7", "c1", 33, 7, 4, False, "plain", 0, 0),
    ("c1_r24_ncp3", "c1", 24, 12, 5, False, "plain", 3, 2),
    ("c3_z16", "c3", 16, 8, 6, True, "plain", 0, 0),
    ("c3_z32", "c3", 32, 16, 7, True, "plain", 0, 0),
    ("c3_z24_fma", "c3", 24, 12, 8, True, "fma", 0, 0),
]
BUILD_CASES = [("c1_r16", "c1", 16, 8, 3, False, 4), ("c3_z12", "c3", 12, 6, 9, True, 3)]


def parse_trace(stderr):
    rows = []
    for line in stderr.strip().split("\n"):
        f = [x.strip() for x in line.split(",")]
        if len(f) != 8 or not f[0].lstrip("-").isdigit():
            continue
        rows.append((int(f[0]), int(f[1]), int(f[4]), float(f[5]), int(f[6]), int(f[7])))
    a = np.array(rows, dtype=[("step", "i4"), ("oldlen", "i4"), ("newlen", "i4"), ("tw", "f8"),
                              ("trig", "i4"), ("hyp", "i4")])
    return a


def solve_case(name, cfg, n, n2, seed, cplx, variant, id_ncp, id_tru):
    y, j = inputs.make(cfg, n=n, seed=seed)
    exe = os.path.join(REF, variant, "zjk.exe" if cplx else "djk.exe")
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "a")
        inputs.write_yj(fn, y, j)
        p = subprocess.run([exe, fn, str(n), str(n2), str(id_ncp), str(id_tru)], capture_output=True,
                           text=True, env=ENV, check=True)
        yu, zz, sy = inputs.read_uzs(fn, n, cplx)
    out = p.stdout.split()
    tt_line = [l for l in p.stdout.split("\n") if l.startswith("J has")][0].split()
    pn = np.array([int(tt_line[2]), int(tt_line[5])], np.int32)
    tt = int(tt_line[10])
    steps = int(out[out.index("Executed") + 1])
    np.savez(os.path.join(HERE, "solve_%s.npz" % name), Y=y, J=j, n2=n2, id_ncp=id_ncp, id_tru=id_tru,
             YU=yu, ZZ=zz, SY=sy, TT=tt, PN=pn, steps=steps, trace=parse_trace(p.stderr),
             variant=variant)
    print(name, "steps", steps)


def build_case(name, cfg, n, n2, seed, cplx, nsteps):
    y, j = inputs.make(cfg, n=n, seed=seed)
    exe = os.path.join(REF, "dbg", "zjk.exe" if cplx else "djk.exe")
    lists, sels = [], []
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "a")
        inputs.write_yj(fn, y, j)
        subprocess.run([exe, "a", str(n), str(n2)], capture_output=True, text=True, env=ENV, cwd=d,
                       check=True)
        for k in range(nsteps):
            txt = open(os.path.join(d, "%s_BUILD.%011d" % ("ZSTEP" if cplx else "DSTEP", k))).read()
            part = txt.split('"I","J","W","P","Q","B"')
            head, sel = part[0], part[1]
            lst = []
            for line in head.split('"J","W","P","Q","B"')[1].strip().split("\n"):
                f = [x.strip() for x in line.split(",")]
                if len(f) == 5 and f[0].isdigit():
                    lst.append((int(f[0]), float(f[1]), int(f[2]), int(f[3]), int(f[4])))
                else:
                    break
            sl = []
            for line in sel.strip().split("\n"):
                f = [x.strip() for x in line.split(",")]
                if len(f) == 6:
                    sl.append((int(f[0]), int(f[1]), float(f[2]), int(f[3]), int(f[4]), int(f[5])))
            lists.append(np.array(lst, dtype=[("pos", "i4"), ("w", "f8"), ("p", "i4"), ("q", "i4"),
                                              ("b", "i4")]))
            sels.append(np.array(sl, dtype=[("i", "i4"), ("pos", "i4"), ("w", "f8"), ("p", "i4"),
                                            ("q", "i4"), ("b", "i4")]))
    kw = {}
    for k in range(nsteps):
        kw["list%d" % k] = lists[k]
        kw["sel%d" % k] = sels[k]
    np.savez(os.path.join(HERE, "build_%s.npz" % name), Y=y, J=j, n2=n2, nsteps=nsteps, **kw)
    print(name, "build dumps", [len(l) for l in lists], [len(s) for s in sels])


def kat_inputs(cplx):
    rng = np.random.default_rng(12345)
    cases = []
    fixed = [
        ([3, 1, 2, 4], (1, -1)), ([3, 1, 2, 4], (1, 1)), ([3, 1, 2, 4], (-1, -1)), ([3, 1, 2, 4], (-1, 1)),
        ([1, 0, 0, 2], (1, 1)), ([2, 0, 0, 1], (1, 1)), ([1, 0, 0, 2], (-1, -1)), ([1, 0, 0, 2], (1, -1)),
        ([-1, 0, 0, 2], (1, -1)), ([0, 0, 0, 0], (1, 1)), ([0, 0, 0, 0], (1, -1)), ([0, 1, 0, 0], (1, -1)),
        ([0, 0, 1, 0], (1, -1)), ([0, 0, 1, 0], (1, 1)), ([1, 1, 1, 1], (1, 1)), ([1, 1, 1, 1], (1, -1)),
        ([1, 1, -1, 1], (1, -1)), ([1, -1, 1, 1], (1, -1)), ([1e-300, 2e-300, 3e-300, 1e-300], (1, -1)),
        ([1e300, 2e300, -3e300, 1e300], (1, -1)), ([5e-320, 1e-310, 0, 1], (1, 1)),
        ([1, 0.9, 0.2, 1], (1, -1)), ([2, 1.999, 0.001, 2], (1, -1)), ([1, 1e-20, 1e-20, 1], (1, -1)),
        ([0, 1, 1, 0], (1, 1)), ([0, 1, 1, 0], (1, -1)), ([1, 2, 3, 6.0000001], (1, -1)),
        ([-0.0, 0.0, 0.0, 1], (1, 1)), ([4, -3, 2, -1], (-1, 1)), ([1, 3, 0, 2], (-1, 1)),
    ]
    for a, jj in fixed:
        a = np.array(a, dtype=np.float64)
        cases.append((a.astype(np.complex128) if cplx else a, jj))
    for k in range(60):
        e = rng.uniform(-8, 8, 4) if k % 3 == 0 else np.zeros(4)
        a = rng.standard_normal(4) * 10.0 ** e
        if cplx:
            a = a + 1j * rng.standard_normal(4) * 10.0 ** e * (k % 4 != 1)
        jj = (1 if rng.random() < 0.5 else -1, 1 if rng.random() < 0.5 else -1)
        cases.append((a, jj))
    return cases


def kat(cplx):
    exe = os.path.join(REF, "plain", "ztest.exe" if cplx else "dtest.exe")
    recs = []
    for a, jj in kat_inputs(cplx):
        if cplx:
            fmt = lambda z: "(%r,%r)" % (float(z.real), float(z.imag))
        else:
            fmt = lambda x: repr(float(x))
        inp = "\n".join([fmt(v) for v in a] + [str(jj[0]), str(jj[1])]) + "\n"
        p = subprocess.run([exe], input=inp, capture_output=True, text=True, env=ENV, check=True)
        vals = {}
        info = None
        for line in p.stdout.replace("=", "= ").split("\n"):
            line = line.strip()
            if line.startswith("INFO="):
                info = int(line.split()[-1])
        txt = p.stdout
        for key in ["U11", "U21", "U12", "U22", "V11", "V21", "V12", "V22", "Z11", "Z21", "Z12", "Z22",
                    "E11", "E21", "E12", "E22"]:
            s = txt.split(key + "=")[-1].split("\n")[0].strip()
            if cplx:
                re, im = s.strip("()").split(",")
                vals[key] = complex(float(re), float(im))
            else:
                vals[key] = float(s)
        recs.append((a, jj, info, vals))
    dt = np.complex128 if cplx else np.float64
    A = np.array([r[0] for r in recs], dtype=dt)
    Jm = np.array([r[1] for r in recs], dtype=np.int32)
    INFO = np.array([r[2] for r in recs], dtype=np.int32)
    get = lambda p: np.array([[r[3][p + "11"], r[3][p + "21"], r[3][p + "12"], r[3][p + "22"]] for r in recs],
                             dtype=dt)
    np.savez(os.path.join(HERE, "kat_%s.npz" % ("zhsvd2" if cplx else "dhsvd2")), A=A, J=Jm, INFO=INFO,
             U=get("U"), V=get("V"), Z=get("Z"), E=get("E"))
    print("kat", "complex" if cplx else "real", len(recs), "cases; INFO values", sorted(set(INFO.tolist())))


if __name__ == "__main__":
    if not os.path.isdir(REF):
        sys.exit("oracle/_ref is not built: make -f oracle/ref.mk")
    for c in SOLVE_CASES:
        solve_case(*c)
    for c in BUILD_CASES:
        build_case(*c)
    kat(False)
    kat(True)
