"""Parity of the gfx950 path (libjkg.so, through its C ABI) with the oracle
and with the reference's golden vectors.  The FP64 path keeps the reference's
operation order, so every comparison here is bit-exact: the 2x2 HSVD, each
pair weight, the sorted candidate list, the chosen pivots, one step, and
whole trajectories (YU, ZZ, SY and the per-step trace)."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

SOLVES_REAL = ["c1_r64", "c2_r32", "c5_r40", "c1_r33_n7"]


@pytest.fixture(scope="module")
def ctx():
    from jkogb_amd import jkg
    c = jkg.Context(0)
    yield c
    c.close()


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def same_bits(a, b):
    return np.array_equal(bits(a), bits(b))


def inputs(cfg, n, seed):
    from jkogb_amd import inputs as gen
    return gen.make(cfg, n=n, seed=seed)


# --------------------------------------------------------------- (d) 2x2 HSVD
@pytest.mark.parametrize("fused", [False, True])
def test_dhsvd2_batch_bit_exact(ctx, orc, fused):
    rng = np.random.default_rng(1)
    g = golden("kat_dhsvd2.npz")
    blocks = [a for a in g["A"]] + [rng.standard_normal(4) * 10.0 ** rng.uniform(-30, 30) for _ in range(3000)]
    blocks += [np.array([1, 0, 0, 2.0]), np.array([0, 0, 0, 0.0]), np.array([-0.0, 0, 0, 1]),
               np.array([1, 1e-300, 0, 1]), np.array([np.nan, 1, 1, 1]), np.array([np.inf, 0, 0, 1])]
    A = np.array(blocks)
    J = np.array([(1, -1) if k % 3 else ((1, 1) if k % 2 else (-1, -1)) for k in range(len(A))], np.int32)
    J[: len(g["J"])] = g["J"]
    for th1 in (0.75, 1.0):
        U, Z, D, info = ctx.dhsvd2(A, J, th1=th1, fused=fused)
        for k in range(len(A)):
            Ao, Uo, Zo, io = orc.dhsvd2(A[k].reshape(2, 2, order="F"), J[k], th1=th1, fused=fused)
            assert info[k] == io, k
            if io < 0:
                continue
            assert same_bits(U[k], Uo.reshape(4, order="F")), k
            assert same_bits(Z[k], Zo.reshape(4, order="F")), k
            assert same_bits(D[k], np.array([Ao[0, 0], Ao[1, 1]])), k


def test_hypot_sqrt_div_are_the_reference_ones(ctx, orc):
    """The 2x2 path rests on glibc's hypot, IEEE sqrt and division; exercised
    through blocks whose first column norm is a hypot of graded entries."""
    rng = np.random.default_rng(2)
    A = np.zeros((20000, 4))
    A[:, 0] = rng.standard_normal(20000) * 10.0 ** rng.uniform(-150, 150, 20000)
    A[:, 1] = rng.standard_normal(20000) * 10.0 ** rng.uniform(-150, 150, 20000)
    A[:, 3] = rng.standard_normal(20000)
    J = np.tile(np.array([1, 1], np.int32), (20000, 1))
    U, Z, D, info = ctx.dhsvd2(A, J, th1=1.0)
    for k in range(0, 20000, 7):
        Ao, Uo, Zo, io = orc.dhsvd2(A[k].reshape(2, 2, order="F"), J[k], th1=1.0)
        assert io == info[k]
        assert same_bits(D[k], np.array([Ao[0, 0], Ao[1, 1]]))
        assert same_bits(U[k], Uo.reshape(4, order="F"))


# --------------------------------------------------------------- (a) weights
@pytest.mark.parametrize("cfg,n,seed", [("c1", 64, 1), ("c2", 100, 2), ("c5", 129, 3), ("c1", 3, 4),
                                        ("c1", 37, 5)])
@pytest.mark.parametrize("fused", [False, True])
def test_pair_weights_bit_exact(ctx, orc, cfg, n, seed, fused):
    from jkogb_amd import jkg
    y, j = inputs(cfg, n, seed)
    keys = ctx.dweights(y, j, fused=fused)
    w_gpu = jkg.keys_to_weights(keys)
    w_orc = orc.pair_weights(y, j, fused=fused)
    nan = np.isnan(w_orc)
    assert np.array_equal(np.isnan(w_gpu), nan)
    w_orc0 = np.where(w_orc == 0, 0.0, w_orc)  # the key of -0 is the key of +0
    assert same_bits(w_gpu[~nan], w_orc0[~nan])


def test_weights_midrun_state(ctx, orc):
    """Weights of a half-converged A (partial blocks, zeros, TH1FIX cases)."""
    from jkogb_amd import jkg
    y, j = inputs("c5", 48, 7)
    r = orc.solve(y, j, 24, max_steps=150)
    a = np.array(r["A"], order="F")
    for i in range(48):
        a[i, i] = r["SY"][i]
    keys = ctx.dweights(a, j)
    w_gpu = jkg.keys_to_weights(keys)
    w_orc = orc.pair_weights(a, j)
    nan = np.isnan(w_orc)
    assert np.array_equal(np.isnan(w_gpu), nan)
    assert same_bits(w_gpu[~nan], np.where(w_orc == 0, 0.0, w_orc)[~nan])


# --------------------------------------------------------------- (b)+(c) build
def check_build(ctx, orc, a, j, n2, fused=False):
    srt, step = ctx.dbuild(a, j, n2, fused=fused)
    tt, P, Q, _ = orc.tru(j)
    info, dz, sl, ostep, it = orc.step_build(a, j, n2, P, Q, tt, fused=fused)
    assert info == 0
    assert len(srt) == it
    w0 = np.where(dz["w"][:it] == 0, 0.0, dz["w"][:it])
    assert same_bits(srt["w"], w0)
    assert np.array_equal(srt["p"], dz["p"][:it]) and np.array_equal(srt["q"], dz["q"][:it])
    assert np.array_equal(step, ostep)
    return len(step)


@pytest.mark.parametrize("cfg,n,seed,n2", [("c1", 64, 1, 32), ("c2", 160, 2, 80), ("c5", 257, 3, 128),
                                           ("c1", 300, 4, 7), ("c1", 12, 5, 6)])
def test_build_bit_exact_vs_oracle(ctx, orc, cfg, n, seed, n2):
    y, j = inputs(cfg, n, seed)
    assert check_build(ctx, orc, y, j, n2) > 0


@pytest.mark.parametrize("n", [300, 700])
def test_build_with_dropped_pairs(ctx, orc, n):
    """Most pairs are dropped before the sort (src/D/dstep.F90 DSTEP_BUILD keeps the
    finite weights only): a diagonal that is already ordered in each sign class
    (non-increasing for J = +1, non-decreasing for J = -1, both positive) and
    about 10 % of the off-diagonal entries nonzero, so a pair with both of its
    off-diagonal entries zero has a NaN weight.  The step's sort leaves those keys
    out after its first pass; the finite part must come out in the oracle's order
    and give the oracle's selection."""
    rng = np.random.default_rng(41 + n)
    npl = n // 2 + 20
    j = np.array([1] * npl + [-1] * (n - npl), dtype=np.int64)
    d = np.concatenate([np.sort(rng.uniform(1.0, 2.0, npl))[::-1], np.sort(rng.uniform(1.0, 2.0, n - npl))])
    a = np.zeros((n, n))
    mask = rng.random((n, n)) < 0.1
    a[mask] = rng.standard_normal(int(mask.sum())) * 1e-2
    a[np.diag_indices(n)] = d
    a = np.asfortranarray(a)
    nn = n * (n - 1) // 2
    srt, _ = ctx.dbuild(a, j, n // 2)
    assert 0.1 * nn < len(srt) < 0.3 * nn  # most keys are NaN
    assert check_build(ctx, orc, a, j, n // 2) > 0


def test_build_deep_walk_uses_filter_rounds(ctx, orc):
    """N = 700: NN = 244650 candidates, more than one walk round (65536
    entries), so the filter/compaction rounds decide the tail of the walk."""
    y, j = inputs("c2", 700, 9)
    assert check_build(ctx, orc, y, j, 350) == 350


@pytest.mark.parametrize("rounds,limit", [(2, 2048), (5, 2049), (16, 3001), (3, 100000)])
def test_build_any_walk_schedule(ctx, orc, rounds, limit):
    """The pivots do not depend on the walk's schedule: rounds that stop after
    `limit` entries (rounded up to the walk's 4096-entry super-chunks), each
    followed by a filter pass, and a last unlimited round give the oracle's
    selection."""
    y, j = inputs("c2", 700, 9)
    try:
        ctx.set_select(rounds, limit)
        assert check_build(ctx, orc, y, j, 350) == 350
    finally:
        ctx.set_select()


def test_solve_same_under_any_walk_schedule(ctx):
    y, j = inputs("c5", 300, 31)
    base = ctx.dsolve(y, j, 150, max_steps=40)
    try:
        ctx.set_select(7, 2051)
        other = ctx.dsolve(y, j, 150, max_steps=40)
    finally:
        ctx.set_select()
    for key in ("SY", "YU", "ZZ"):
        assert same_bits(base[key], other[key]), key
    assert base["info"] == other["info"]


def test_build_vs_reference_debug_dumps(ctx):
    g = golden("build_c1_r16.npz")
    srt, step = ctx.dbuild(g["Y"], g["J"], int(g["n2"]))
    lst, sel = g["list0"], g["sel0"]
    assert same_bits(srt["w"], lst["w"])
    assert np.array_equal(srt["p"], lst["p"]) and np.array_equal(srt["q"], lst["q"])
    assert np.array_equal(step, sel["pos"])


def test_plugin_srt_ncp(ctx, orc):
    """The DPROC plug-ins ASRT/ANCP on lists with many exact ties and +-0."""
    import ctypes as C
    rng = np.random.default_rng(3)
    for n in (5, 17, 64, 200):
        pairs = [(p, q) for q in range(2, n + 1) for p in range(1, q)]
        w = np.round(rng.standard_normal(len(pairs)), 1)
        w[rng.random(len(pairs)) < 0.2] = -0.0
        w[rng.random(len(pairs)) < 0.1] = 0.0
        perm = rng.permutation(len(pairs))
        dz = np.zeros(len(pairs), orc.AW_DTYPE)
        for k, i in enumerate(perm):
            p, q = pairs[i]
            dz[k] = (w[i], p, q, q - p, k + 1)
        a = dz.copy()
        ctx.srt(a)
        b = np.zeros(2 * len(dz), orc.AW_DTYPE)
        b[:len(dz)] = dz
        orc.lib().orc_srt.restype = C.c_int64
        orc.lib().orc_srt(len(dz), len(b), orc._p(b))
        assert np.array_equal(a["p"], b["p"][:len(dz)]) and np.array_equal(a["q"], b["q"][:len(dz)])
        n2 = n // 2
        step = ctx.ncp(n, a, n2)
        sl = C.c_int32(0)
        st = np.zeros(n2, np.int32)
        orc.lib().orc_ncp.restype = C.c_int64
        orc.lib().orc_ncp(2, n, len(dz), len(b), orc._p(b), n2, C.byref(sl), orc._p(st))
        assert np.array_equal(step, st[:sl.value])


# --------------------------------------------------------------- whole runs
def check_solve(res, ref_sy, ref_yu, ref_zz, ref_trace=None, steps=None):
    if steps is not None:
        assert res["info"] == steps
    assert same_bits(res["SY"], ref_sy)
    assert same_bits(res["YU"], ref_yu)
    assert same_bits(res["ZZ"], ref_zz)
    if ref_trace is not None:
        t = res["trace"]
        assert len(t) == len(ref_trace)
        for f in ["step", "oldlen", "newlen", "trig", "hyp"]:
            assert np.array_equal(t[f], ref_trace[f]), f
        assert same_bits(t["tw"], ref_trace["tw"])


@pytest.mark.parametrize("name", SOLVES_REAL)
def test_solve_bit_exact_vs_reference(ctx, name):
    g = golden("solve_%s.npz" % name)
    r = ctx.dsolve(g["Y"], g["J"], int(g["n2"]))
    assert r["tt"] == int(g["TT"]) and tuple(r["pn"]) == tuple(g["PN"])
    check_solve(r, g["SY"], g["YU"], g["ZZ"], g["trace"], int(g["steps"]))


def test_solve_fused_bit_exact_vs_oracle(ctx, orc):
    y, j = inputs("c1", 64, 1)
    r = ctx.dsolve(y, j, 32, fused=True)
    o = orc.solve(y, j, 32, fused=True)
    check_solve(r, o["SY"], o["YU"], o["ZZ"], steps=o["info"])


@pytest.mark.parametrize("cfg,n,seed,n2", [("c2", 128, 11, 64), ("c1", 96, 12, 17), ("c5", 90, 13, 45)])
def test_solve_full_bit_exact_vs_oracle(ctx, orc, cfg, n, seed, n2):
    y, j = inputs(cfg, n, seed)
    r = ctx.dsolve(y, j, n2)
    o = orc.solve(y, j, n2)
    check_solve(r, o["SY"], o["YU"], o["ZZ"], steps=o["info"])
    t, ot = r["trace"], o["trace"]
    assert np.array_equal(t["newlen"], ot["newlen"]) and same_bits(t["tw"], ot["tw"])


@pytest.mark.parametrize("cfg,n,seed,steps", [("c2", 384, 21, 12), ("c5", 515, 22, 6)])
def test_solve_prefix_bit_exact_vs_oracle(ctx, orc, cfg, n, seed, steps):
    """Larger N, a bounded number of steps: the state after k steps matches."""
    y, j = inputs(cfg, n, seed)
    r = ctx.dsolve(y, j, 0, max_steps=steps)
    o = orc.solve(y, j, 0, max_steps=steps)
    assert r["info"] == o["info"] == steps
    check_solve(r, o["SY"], o["YU"], o["ZZ"])
    assert np.array_equal(r["trace"]["newlen"], o["trace"]["newlen"][:steps])


def test_solve_edge_cases(ctx, orc):
    # smallest size, all-trigonometric and all-hyperbolic sign patterns
    for y, j, n2 in [(np.array([[1.0, 2, 3], [4, 5, 6], [7, 8, 10]]), np.array([1, -1, 1]), 1),
                     (inputs("c1", 20, 31)[0], np.ones(20, np.int64), 10),
                     (inputs("c1", 21, 32)[0], -np.ones(21, np.int64), 10),
                     (np.diag([3.0, 1, 2, 5]), np.array([1, 1, -1, -1]), 2),
                     (np.eye(6), np.array([1, -1, 1, -1, 1, -1]), 3)]:
        r = ctx.dsolve(y, j, n2)
        o = orc.solve(y, j, n2)
        check_solve(r, o["SY"], o["YU"], o["ZZ"], steps=o["info"])


def test_solve_rejects_bad_j_and_sizes(ctx):
    from jkogb_amd import jkg
    with pytest.raises(jkg.JKGError):
        ctx.dsolve(np.eye(4), np.array([1, 2, -1, 1]))
    with pytest.raises(jkg.JKGError):
        ctx.dsolve(np.eye(2), np.array([1, -1]))


def test_solve_nonfinite_input_reports_error(ctx, orc):
    y, j = inputs("c1", 16, 41)
    y[3, 5] = np.nan
    r = ctx.dsolve(y, j, 8, max_steps=3000)
    o = orc.solve(y, j, 8, max_steps=3000)
    assert r["info"] == o["info"]


def test_one_step_entry_point(ctx, orc):
    y, j = inputs("c1", 50, 51)
    n = 50
    a, u, z, rec = ctx.dstep(y, j, np.eye(n), np.eye(n), 25)
    o = orc.solve(y, j, 25, max_steps=1)
    assert rec["oldlen"] == o["trace"]["oldlen"][0] and rec["newlen"] == o["trace"]["newlen"][0]
    assert same_bits(np.diag(a), o["SY"])
    assert same_bits(u, o["YU"])


# --------------------------------------------------------------- complex (zjk)
@pytest.mark.parametrize("fused", [False, True])
def test_zhsvd2_batch_bit_exact(ctx, orc, fused):
    rng = np.random.default_rng(5)
    g = golden("kat_zhsvd2.npz")
    blocks = [a for a in g["A"]]
    for k in range(2000):
        e = 10.0 ** rng.uniform(-20, 20)
        a = (rng.standard_normal(4) + 1j * rng.standard_normal(4) * (k % 5 != 0)) * e
        if k % 7 == 0:
            a[rng.integers(0, 4)] = 0
        blocks.append(a)
    A = np.array(blocks, dtype=np.complex128)
    J = np.array([(1, -1) if k % 3 else ((1, 1) if k % 2 else (-1, -1)) for k in range(len(A))], np.int32)
    J[: len(g["J"])] = g["J"]
    for th1 in (0.75, 1.0):
        U, Z, D, info = ctx.dhsvd2(A, J, th1=th1, fused=fused)
        for k in range(len(A)):
            Ao, Uo, Zo, io = orc.zhsvd2(A[k].reshape(2, 2, order="F"), J[k], th1=th1, fused=fused)
            assert info[k] == io, k
            if io < 0:
                continue
            assert same_bits(U[k], Uo.reshape(4, order="F")), k
            assert same_bits(Z[k], Zo.reshape(4, order="F")), k
            assert same_bits(D[k], np.array([Ao[0, 0].real, Ao[1, 1].real])), k


@pytest.mark.parametrize("n,seed", [(24, 1), (41, 2)])
@pytest.mark.parametrize("fused", [False, True])
def test_zweights_bit_exact(ctx, orc, n, seed, fused):
    from jkogb_amd import jkg
    y, j = inputs("c3", n, seed)
    w_gpu = jkg.keys_to_weights(ctx.dweights(y, j, fused=fused))
    w_orc = orc.pair_weights(y, j, fused=fused)
    nan = np.isnan(w_orc)
    assert np.array_equal(np.isnan(w_gpu), nan)
    assert same_bits(w_gpu[~nan], np.where(w_orc == 0, 0.0, w_orc)[~nan])


@pytest.mark.parametrize("name", ["c3_z16", "c3_z32"])
def test_zsolve_bit_exact_vs_reference(ctx, name):
    g = golden("solve_%s.npz" % name)
    r = ctx.dsolve(g["Y"], g["J"], int(g["n2"]))
    assert r["tt"] == int(g["TT"])
    check_solve(r, g["SY"], g["YU"], g["ZZ"], g["trace"], int(g["steps"]))


def test_zsolve_fused_and_larger_vs_oracle(ctx, orc):
    y, j = inputs("c3", 40, 61)
    r = ctx.dsolve(y, j, 20, fused=True)
    o = orc.solve(y, j, 20, fused=True)
    check_solve(r, o["SY"], o["YU"], o["ZZ"], steps=o["info"])
    y, j = inputs("c3", 200, 62)
    r = ctx.dsolve(y, j, 100, max_steps=5)
    o = orc.solve(y, j, 100, max_steps=5)
    check_solve(r, o["SY"], o["YU"], o["ZZ"], steps=5)


def test_context_reuse_with_another_j_split(ctx, orc):
    # a second problem of the same N with a different number of +1 signs in the same context
    # (the captured step must follow the new split)
    Y1, J1 = inputs("c2", 48, 4)
    Y2 = np.random.default_rng(8).standard_normal((48, 48))
    J2 = np.array([1] * 30 + [-1] * 18)
    np.random.default_rng(9).shuffle(J2)
    for Y, J in [(Y1, J1), (Y2, J2), (Y1, J1)]:
        out = ctx.dsolve(Y, J, 24)
        ref = orc.solve(Y, J, 24)
        assert out["info"] == ref["info"]
        assert same_bits(out["SY"], ref["SY"])
        assert same_bits(out["YU"], ref["YU"])


def test_progress_hook(ctx):
    # the hook sees the steps executed so far at every poll, increasing, with their trace records
    # already in the caller's buffer, and the end exactly once
    calls = []
    Y, J = inputs("c2", 64, 5)

    def hook(s, d):
        calls.append((s, d))
    r = ctx.dsolve(Y, J, 32, steps_per_sync=8, progress=hook)
    assert calls and calls[-1][1] and sum(d for _, d in calls) == 1
    assert all(a[0] <= b[0] for a, b in zip(calls, calls[1:]))
    assert calls[-1][0] == r["steps"] == r["info"] + 1
    assert all(s % 8 == 0 for s, d in calls if not d)


# --------------------------------------------------------------- ID_NCP=3
def test_ncp3_plugin_bit_exact(ctx, orc):
    # AW_NCP3 behind the ANCP plug-in against the oracle, with ties, +-0 and negative weights
    import ctypes as C
    rng = np.random.default_rng(33)
    for n in (5, 12, 40, 97):
        pairs = [(p, q) for q in range(2, n + 1) for p in range(1, q)]
        for trial in range(3):
            w = np.round(rng.standard_normal(len(pairs)), 1)
            w[rng.random(len(pairs)) < 0.15] = -0.0
            perm = rng.permutation(len(pairs))
            dz = np.zeros(len(pairs), orc.AW_DTYPE)
            for k, i in enumerate(perm):
                p, q = pairs[i]
                dz[k] = (w[i], p, q, q - p, k + 1)
            a = dz.copy()
            ctx.srt(a)
            n2 = [n // 2, max(1, n // 3), 1][trial]
            step = ctx.ncp(n, a, n2, id_ncp=3)
            b = np.zeros(2 * len(a), orc.AW_DTYPE)
            b[:len(a)] = a
            sl = C.c_int32(0)
            st = np.zeros(n2, np.int32)
            orc.lib().orc_ncp.restype = C.c_int64
            orc.lib().orc_ncp(3, n, len(a), len(b), orc._p(b), n2, C.byref(sl), orc._p(st))
            assert np.array_equal(step, st[:sl.value]), (n, trial)


def test_ncp3_solve_bit_exact_vs_reference(ctx):
    # a whole djk.exe run with ID_NCP=3 (the reference's own build)
    g = golden("solve_c1_r24_ncp3.npz")
    r = ctx.dsolve(g["Y"], g["J"], int(g["n2"]), id_ncp=int(g["id_ncp"]), id_tru=int(g["id_tru"]))
    check_solve(r, g["SY"], g["YU"], g["ZZ"], g["trace"], int(g["steps"]))


@pytest.mark.parametrize("cfg,n,fused", [("c1", 32, False), ("c2", 40, True), ("c5", 37, False)])
def test_ncp3_solve_bit_exact_vs_oracle(ctx, orc, cfg, n, fused):
    y, j = inputs(cfg, n, 41)
    r = ctx.dsolve(y, j, n // 2, id_ncp=3, fused=fused)
    o = orc.solve(y, j, n // 2, id_ncp=3, fused=fused)
    check_solve(r, o["SY"], o["YU"], o["ZZ"], steps=o["info"])


def test_ncp3_complex_solve_bit_exact_vs_oracle(ctx, orc):
    y, j = inputs("c3", 20, 42)
    r = ctx.dsolve(y, j, 10, id_ncp=3)
    o = orc.solve(y, j, 10, id_ncp=3)
    check_solve(r, o["SY"], o["YU"], o["ZZ"], steps=o["info"])
