"""The drop-in programs host/djk.exe and host/zjk.exe: the reference's process-level
interface (src/D/djk.F90, src/Z/zjk.F90, src/shared/readcl.F90, src/shared/binio.F90)
with the HSVD on the GPU.  On the golden inputs (the reference's own runs, written by
tests/golden/make_golden.py from oracle/_ref): FN.YU, FN.ZZ and FN.SY byte for byte, the
two stdout lines (src/D/djk.F90:79, 86-87) and the stderr CSV line by line, every field
but the two timing columns (src/D/dstep.F90:388-440, 559-562).  The fast solvers behind
JKG_SOLVER are checked by their accuracy against the plain run of the same program."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden

pytestmark = pytest.mark.gpu

DJK = os.path.join(ROOT, "host", "djk.exe")
ZJK = os.path.join(ROOT, "host", "zjk.exe")
HEADER = '"STEP","OLDLEN","BUILDs","TRANSFs","NEWLEN","NORMF2D","TRIG","HYP"'
CSV_ROW = re.compile(r"^ *(-?\d+), *(-?\d+), *(\d+\.\d{6}), *(\d+\.\d{6}), *(-?\d+), *(\S+), *(-?\d+), *(-?\d+)$")


def write_inputs(fn, y, j):
    np.asfortranarray(y).reshape(-1, order="F").tofile(fn + ".Y")
    np.asarray(j, dtype=np.int64).tofile(fn + ".J")


def read_outputs(fn, n, cplx):
    dt = np.complex128 if cplx else np.float64
    yu = np.fromfile(fn + ".YU", dtype=dt)
    zz = np.fromfile(fn + ".ZZ", dtype=dt)
    sy = np.fromfile(fn + ".SY", dtype=np.float64)
    assert yu.size == n * n and zz.size == n * n and sy.size == n
    return yu.reshape(n, n, order="F"), zz.reshape(n, n, order="F"), sy


def run(exe, fn, args, **env):
    e = dict(os.environ, JKG_DEVICE="0")
    e.update(env)
    return subprocess.run([exe, fn] + [str(a) for a in args], capture_output=True, text=True, env=e, timeout=300)


def csv_rows(stderr):
    lines = stderr.rstrip("\n").split("\n")
    assert lines[0] == HEADER, lines[0]
    rows = []
    for line in lines[1:]:
        m = CSV_ROW.match(line)
        if not m:
            break
        rows.append(m.groups())
    return rows, lines


def census(pn, tt):
    return "J has%11d +  and%11d - signs for <=%11d trig. transfs." % (pn[0], pn[1], tt)


def check_trace(rows, tr):
    assert len(rows) == len(tr)
    for r, t in zip(rows, tr):
        assert (int(r[0]), int(r[1]), int(r[4]), int(r[6]), int(r[7])) == \
            (t["step"], t["oldlen"], t["newlen"], t["trig"], t["hyp"])
        assert np.float64(float(r[5])).tobytes() == np.float64(t["tw"]).tobytes()
        assert float(r[2]) >= 0.0 and float(r[3]) >= 0.0
    # the last line is the step without a pivot or without an effective transform
    assert int(rows[-1][4]) <= 0


def run_case(tmp_path, g, fused):
    y, j = g["Y"], g["J"]
    cplx = np.iscomplexobj(y)
    n = y.shape[0]
    fn = str(tmp_path / "a")
    write_inputs(fn, y, j)
    p = run(ZJK if cplx else DJK, fn, [n, int(g["n2"]), int(g["id_ncp"]), int(g["id_tru"])], JKG_FUSED=str(fused))
    assert p.returncode == 0, p.stderr[-2000:]
    out = p.stdout.rstrip("\n").split("\n")
    assert out[0] == census(g["PN"], int(g["TT"]))
    assert len(out) == 2
    rows, lines = csv_rows(p.stderr)
    assert len(lines) == len(rows) + 1  # nothing else on stderr
    return out, rows, read_outputs(fn, n, cplx)


@pytest.mark.parametrize("case", ["c1_r64", "c2_r32", "c5_r40", "c1_r33_n7", "c1_r24_ncp3", "c3_z16", "c3_z32"])
def test_dropin_matches_the_reference_run(tmp_path, case):
    g = golden("solve_%s.npz" % case)
    out, rows, (yu, zz, sy) = run_case(tmp_path, g, 0)
    assert re.fullmatch(r"Executed %11d steps with transformations in +\d+\.\d{6} s" % int(g["steps"]), out[1]), out[1]
    for a, b in ((yu, g["YU"]), (zz, g["ZZ"]), (sy, g["SY"])):
        assert np.ascontiguousarray(a).tobytes() == np.ascontiguousarray(b).tobytes()
    check_trace(rows, g["trace"])


@pytest.mark.parametrize("case", ["c1_r64_fma", "c3_z24_fma"])
def test_dropin_fused_policy(tmp_path, orc, case):
    """JKG_FUSED=1 (FMAD = fma): the bits of the oracle's fused policy, and the accuracy of the
    reference's build with contracted multiply-adds (which fuses more than the FMAD sites)."""
    g = golden("solve_%s.npz" % case)
    out, rows, (yu, zz, sy) = run_case(tmp_path, g, 1)
    o = orc.solve(g["Y"], g["J"], int(g["n2"]), fused=True)
    assert re.fullmatch(r"Executed %11d steps with transformations in +\d+\.\d{6} s" % o["info"], out[1]), out[1]
    for a, b in ((yu, o["YU"]), (zz, o["ZZ"]), (sy, o["SY"])):
        assert np.ascontiguousarray(a).tobytes() == np.ascontiguousarray(b).tobytes()
    ot = o["trace"]
    tr = np.zeros(len(ot), dtype=g["trace"].dtype)
    for f in ("step", "oldlen", "newlen", "tw", "trig", "hyp"):
        tr[f] = ot[f]
    check_trace(rows, tr)
    n = sy.size
    assert np.max(np.abs(sy - g["SY"]) / np.abs(g["SY"])) <= 64 * n * np.finfo(np.float64).eps


@pytest.mark.parametrize("solver", ["mixed", "learned", "warm", "warm_learned"])
def test_dropin_fast_solvers(tmp_path, solver):
    """JKG_SOLVER selects the mixed solver, the trained controller, the FP32 warm start or both;
    their singular values agree with the plain run of the same program to 64 N eps, the factors
    are orthogonal and J-orthogonal to the same order, and the CSV has one line per step."""
    from jkogb_amd import inputs
    y, j = inputs.make("c5", n=96, seed=11)
    n = 96
    eps = np.finfo(np.float64).eps
    fn0, fn1 = str(tmp_path / "plain"), str(tmp_path / solver)
    write_inputs(fn0, y, j)
    write_inputs(fn1, y, j)
    p0 = run(DJK, fn0, [n, n // 2])
    p1 = run(DJK, fn1, [n, n // 2], JKG_SOLVER=solver)
    assert p0.returncode == 0 and p1.returncode == 0, p1.stderr[-2000:]
    assert p1.stdout.split("\n")[0] == p0.stdout.split("\n")[0]
    steps = int(re.search(r"Executed +(\d+) steps", p1.stdout).group(1))
    rows, _ = csv_rows(p1.stderr)
    assert len(rows) == steps + 1
    _, _, sy0 = read_outputs(fn0, n, False)
    yu, zz, sy = read_outputs(fn1, n, False)
    assert np.max(np.abs(sy - sy0) / np.abs(sy0)) <= 64 * n * eps
    assert np.linalg.norm(yu.T @ yu - np.eye(n)) <= 64 * n * eps
    jd = np.diag(j.astype(float))
    assert np.linalg.norm(zz @ jd @ zz.T - jd) / np.linalg.norm(zz) ** 2 <= 64 * n * eps
    assert np.linalg.norm(y - yu @ np.diag(sy) @ zz) / np.linalg.norm(y) <= 64 * n * eps


def test_dropin_complex_fast_solvers(tmp_path):
    from jkogb_amd import inputs
    y, j = inputs.make("c3", n=64, seed=12)
    n = 64
    eps = np.finfo(np.float64).eps
    fn0 = str(tmp_path / "plain")
    write_inputs(fn0, y, j)
    assert run(ZJK, fn0, [n, n // 2]).returncode == 0
    _, _, sy0 = read_outputs(fn0, n, True)
    for solver in ("mixed", "warm"):
        fn = str(tmp_path / solver)
        write_inputs(fn, y, j)
        p = run(ZJK, fn, [n, n // 2], JKG_SOLVER=solver)
        assert p.returncode == 0, p.stderr[-2000:]
        yu, zz, sy = read_outputs(fn, n, True)
        assert np.max(np.abs(sy - sy0) / np.abs(sy0)) <= 64 * n * eps, solver
        assert np.linalg.norm(y - yu @ np.diag(sy) @ zz) / np.linalg.norm(y) <= 64 * n * eps, solver


def test_dropin_error_writes_the_factors(tmp_path, orc):
    """A loop that ends with an error (a 2x2 of an infinite entry fails, INFO < 0): the ERROR
    line on stderr after the CSV with the loop's INFO, exit status 0 and FN.YU, FN.ZZ, FN.SY
    written all the same, as the reference does (src/D/djk.F90:86-114)."""
    from jkogb_amd import inputs
    y, j = inputs.make("c1", n=16, seed=41)
    y[3, 3] = np.inf
    o = orc.solve(y, j, 8, max_steps=3000)
    assert o["info"] < 0
    fn = str(tmp_path / "a")
    write_inputs(fn, y, j)
    p = run(DJK, fn, [16, 8])
    assert p.returncode == 0, p.stderr[-2000:]
    read_outputs(fn, 16, False)
    assert re.search(r"^ERROR +%d after +\d+\.\d{6} s$" % o["info"], p.stderr, re.M), p.stderr[-500:]
    assert "Executed" not in p.stdout
    rows, _ = csv_rows(p.stderr)
    assert int(rows[-1][4]) == o["info"]  # the failing step's NEWLEN is the 2x2's INFO


def test_dropin_rejects_a_bad_solver_name(tmp_path):
    from jkogb_amd import inputs
    y, j = inputs.make("c1", n=8, seed=1)
    fn = str(tmp_path / "a")
    write_inputs(fn, y, j)
    p = run(DJK, fn, [8, 4], JKG_SOLVER="fastest")
    assert p.returncode != 0
    assert "JKG_SOLVER" in p.stderr
    assert not os.path.exists(fn + ".YU")


@pytest.mark.parametrize("case,k", [("c1_r64", 150), ("c3_z16", 60)])
def test_dropin_checkpoint_and_resume(tmp_path, case, k):
    """JKG_CHECKPOINT=K: the loop stops every K steps, writes FN.CA/CU/CZ/CS and goes on; the
    outputs, stdout and CSV are the reference's run.  JKG_RESUME=1 from the last checkpoint
    left behind: the same outputs, and the CSV from the checkpoint's step on."""
    g = golden("solve_%s.npz" % case)
    y, j = g["Y"], g["J"]
    cplx = np.iscomplexobj(y)
    n, steps = y.shape[0], int(g["steps"])
    assert steps > k
    fn = str(tmp_path / "a")
    write_inputs(fn, y, j)
    exe = ZJK if cplx else DJK
    args = [n, int(g["n2"]), int(g["id_ncp"]), int(g["id_tru"])]
    p = run(exe, fn, args, JKG_CHECKPOINT=str(k))
    assert p.returncode == 0, p.stderr[-2000:]
    out = p.stdout.rstrip("\n").split("\n")
    assert re.fullmatch(r"Executed %11d steps with transformations in +\d+\.\d{6} s" % steps, out[1]), out[1]
    rows, lines = csv_rows(p.stderr)
    assert len(lines) == len(rows) + 1
    check_trace(rows, g["trace"])
    outs = read_outputs(fn, n, cplx)
    for a, b in zip(outs, (g["YU"], g["ZZ"], g["SY"])):
        assert np.ascontiguousarray(a).tobytes() == np.ascontiguousarray(b).tobytes()
    # the last checkpoint: after the last whole part of K steps
    s0 = int(np.fromfile(fn + ".CS", dtype=np.int64)[0])
    assert s0 == (steps // k) * k
    dt = np.complex128 if cplx else np.float64
    for name in ("CA", "CU", "CZ"):
        assert np.fromfile(fn + "." + name, dtype=dt).size == n * n
    for name in ("YU", "ZZ", "SY"):
        os.remove(fn + "." + name)
    os.remove(fn + ".Y")  # the checkpoint replaces it
    p = run(exe, fn, args, JKG_RESUME="1")
    assert p.returncode == 0, p.stderr[-2000:]
    out = p.stdout.rstrip("\n").split("\n")
    assert out[0] == census(g["PN"], int(g["TT"]))
    assert re.fullmatch(r"Executed %11d steps with transformations in +\d+\.\d{6} s" % steps, out[1]), out[1]
    rows, _ = csv_rows(p.stderr)
    check_trace(rows, g["trace"][s0:])
    outs = read_outputs(fn, n, cplx)
    for a, b in zip(outs, (g["YU"], g["ZZ"], g["SY"])):
        assert np.ascontiguousarray(a).tobytes() == np.ascontiguousarray(b).tobytes()


def test_dropin_frames(tmp_path):
    """JKG_FRAMES=E,S: FN.FRM holds the S x S pictures of |A| at the steps 0 mod E, as the
    library records them (jkg_set_frames), also when the loop runs in checkpointed parts."""
    from jkogb_amd import jkg
    g = golden("solve_c1_r64.npz")
    y, j = g["Y"], g["J"]
    n, steps = y.shape[0], int(g["steps"])
    with jkg.Context(0) as ctx:
        ctx.set_frames(every=5, size=16, max_frames=4096)
        ctx.dsolve(y, j, int(g["n2"]))
        ref = ctx.frames()
    assert ref.shape == (steps // 5 + 1, 16, 16)
    for env in ({"JKG_FRAMES": "5,16"}, {"JKG_FRAMES": "5,16,64", "JKG_CHECKPOINT": "77"}):
        fn = str(tmp_path / "f")
        write_inputs(fn, y, j)
        p = run(DJK, fn, [n, int(g["n2"])], **env)
        assert p.returncode == 0, p.stderr[-2000:]
        frm = np.fromfile(fn + ".FRM", dtype=np.uint8)
        assert frm.size == ref.size and np.array_equal(frm.reshape(ref.shape), ref), env
