"""Checkpoint and resume of the loop (jkg_dsolve_from / jkg_zsolve_from) through the C ABI.

The loop's state between two steps is (A, U, Z) and the next step's index: the reference's
DSTEP_LOOP carries nothing else from one step to the next (src/D/dstep.F90:549-592).  So a
run cut into parts, each resumed from the state the previous one returned, must be the
uninterrupted run bit for bit:
  - against the reference's own runs (the goldens): YU, ZZ, SY and every trace field but the
    two timing columns, for cuts at several part lengths, real and complex;
  - the state after K steps against the oracle's loop stopped after K steps;
  - frames taken at the same absolute steps as in one run.
"""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

FIELDS = ("step", "oldlen", "newlen", "tw", "trig", "hyp")


@pytest.fixture(scope="module")
def ctx():
    from jkogb_amd import jkg
    c = jkg.Context(0)
    yield c
    c.close()


def bits(x):
    return np.ascontiguousarray(x).tobytes()


def run_in_parts(ctx, Y, J, n2, k, **kw):
    n = Y.shape[0]
    dt = Y.dtype
    A, U, Z = Y, np.eye(n, dtype=dt), np.eye(n, dtype=dt)
    step0, parts, traces = 0, 0, []
    while True:
        r = ctx.dsolve_from(A, U, Z, J, step0=step0, n2=n2, max_steps=k, **kw)
        parts += 1
        traces.append(r["trace"])
        assert r["steps"] == len(r["trace"]) and r["trace"]["step"][0] == step0
        if r["ended"]:
            return r, np.concatenate(traces), parts
        assert r["info"] == step0 + k and r["steps"] == k, (r["info"], step0, k)
        A, U, Z, step0 = r["A"], r["U"], r["Z"], r["info"]


@pytest.mark.parametrize("case,k", [("c1_r64", 97), ("c1_r64", 1), ("c2_r32", 50), ("c5_r40", 33),
                                    ("c1_r24_ncp3", 40), ("c3_z16", 40), ("c3_z32", 7)])
def test_resumed_parts_are_the_reference_run(ctx, case, k):
    g = golden("solve_%s.npz" % case)
    Y, J = g["Y"], g["J"]
    r, tr, parts = run_in_parts(ctx, Y, J, int(g["n2"]), k, id_ncp=int(g["id_ncp"]), id_tru=int(g["id_tru"]))
    assert parts > 1
    assert r["info"] == int(g["steps"])
    for a, b in ((r["YU"], g["YU"]), (r["ZZ"], g["ZZ"]), (r["SY"], g["SY"])):
        assert bits(a) == bits(b)
    gt = g["trace"]
    assert len(tr) == len(gt)
    for f in FIELDS:
        assert bits(tr[f]) == bits(gt[f]), f


def test_state_after_k_steps_is_the_oracle_loop(ctx, orc):
    from jkogb_amd import inputs
    n, k = 48, 40
    Y, J = inputs.make("c2", n=n, seed=5)
    r = ctx.dsolve_from(Y, np.eye(n), np.eye(n), J, n2=n // 2, max_steps=k)
    o = orc.solve(Y, J, n // 2, max_steps=k)
    assert not r["ended"] and r["info"] == k and o["info"] == k
    # the finalised outputs of the state are the oracle's after K steps
    for a, b in ((r["YU"], o["YU"]), (r["ZZ"], o["ZZ"]), (r["SY"], o["SY"])):
        assert bits(a) == bits(b)
    # the state itself: U as is, A with its diagonal (the oracle's finalisation moved it to SY)
    assert bits(r["U"]) == bits(o["YU"])
    off = ~np.eye(n, dtype=bool)
    assert bits(r["A"][off]) == bits(o["A"][off])
    assert bits(np.diag(r["A"]).copy()) == bits(o["SY"])
    # and Z is J Z^T J of the finalised ZZ: the finalisation undone
    jz = np.asarray(J, dtype=float)
    assert bits(r["Z"]) == bits(np.asfortranarray(jz[:, None] * o["ZZ"].T * jz[None, :]))


def test_frames_follow_absolute_steps(ctx):
    from jkogb_amd import inputs
    n = 40
    Y, J = inputs.make("c1", n=n, seed=8)
    ctx.set_frames(every=3, size=8, max_frames=4096)
    try:
        whole = ctx.dsolve(Y, J, n // 2)
        f_whole = ctx.frames()
        A, U, Z, step0, fr = Y, np.eye(n), np.eye(n), 0, []
        while True:
            r = ctx.dsolve_from(A, U, Z, J, step0=step0, n2=n // 2, max_steps=10)
            fr.append(ctx.frames())
            if r["ended"]:
                break
            A, U, Z, step0 = r["A"], r["U"], r["Z"], r["info"]
    finally:
        ctx.set_frames(0)
    assert r["info"] == whole["info"]
    f_parts = np.concatenate(fr)
    assert f_whole.shape[0] == (whole["steps"] - 1) // 3 + 1
    assert f_parts.shape == f_whole.shape and np.array_equal(f_parts, f_whole)


def test_resume_rejects_bad_arguments(ctx):
    import ctypes as C
    from jkogb_amd import jkg
    n = 8
    A = np.zeros((n, n), order="F")
    J = np.ones(n, dtype=np.int64)
    res = jkg.Result()
    o = jkg.Opts(0, 0, 0, -1, 0, 0, None)
    keep = [A, A.copy(), A.copy(), J, A.copy(), A.copy(), np.zeros(n)]  # alive for the calls
    p = [x.ctypes.data for x in keep]
    base = [ctx._h, n, 4, p[0], p[1], p[2], p[3], 0, p[4], p[5], p[6]]
    for pos, val in ((4, None), (5, None), (6, None), (7, None), (8, -1), (9, None), (10, None), (11, None)):
        args = list(base)
        args[pos - 1] = val
        assert jkg._dsolve_from(*args, C.byref(o), C.byref(res)) == -pos, pos
    with pytest.raises(ValueError):
        ctx.dsolve_from(A[:, :4], A, A, J)
