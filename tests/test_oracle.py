"""Pins the CPU oracle (oracle/libjkoracle.so) before it is trusted as the
checker of the GPU path:

* bit-exact against the Fortran reference's own outputs (tests/golden/,
  written by tests/golden/make_golden.py from oracle/_ref),
* against the mathematics the reference does not state as tests: the HSVD
  identities, sigma^2 = |eig(Y J Y^T)|, the canonical output order, brute-force
  greedy matching, the closed form of the pair enumeration.
"""
import itertools
import math

import numpy as np
import pytest

from conftest import golden

EPS = np.finfo(np.float64).eps
SOLVES = ["c1_r64", "c2_r32", "c5_r40", "c1_r33_n7", "c1_r24_ncp3", "c3_z16", "c3_z32"]


def bits(a):
    """Bit patterns of an FP64 or complex128 array (as uint64 words)."""
    return np.ascontiguousarray(a).view(np.uint64)


@pytest.mark.parametrize("name", SOLVES)
def test_solve_bit_exact_vs_reference(orc, name):
    g = golden("solve_%s.npz" % name)
    r = orc.solve(g["Y"], g["J"], int(g["n2"]), id_ncp=int(g["id_ncp"]), id_tru=int(g["id_tru"]))
    assert r["info"] == int(g["steps"])
    assert r["tt"] == int(g["TT"]) and list(r["pn"]) == list(g["PN"])
    assert np.array_equal(bits(r["SY"]), bits(g["SY"]))
    assert np.array_equal(bits(r["YU"]), bits(g["YU"]))
    assert np.array_equal(bits(r["ZZ"]), bits(g["ZZ"]))
    t, gt = r["trace"], g["trace"]
    assert len(gt) == r["info"] + 1
    for f in ["step", "oldlen", "newlen", "trig", "hyp"]:
        assert np.array_equal(t[f][:len(gt)], gt[f]), f
    # the CSV prints NORMF2D with 17 digits after the point: exact for FP64
    assert np.array_equal(t["tw"][:len(gt)], gt["tw"])


@pytest.mark.parametrize("name", ["c1_r64_fma", "c3_z24_fma"])
def test_fused_policy_accuracy_vs_fma_reference(orc, name):
    """The reference built with -ffp-contract=fast contracts more than the FMAD
    sites, so the fused policy agrees with it in accuracy, not bits."""
    g = golden("solve_%s.npz" % name)
    n = g["Y"].shape[0]
    r = orc.solve(g["Y"], g["J"], int(g["n2"]), fused=True)
    rel = np.max(np.abs(r["SY"] - g["SY"]) / np.abs(g["SY"]))
    assert rel <= 64 * n * EPS


@pytest.mark.parametrize("name", ["c1_r16", "c3_z12"])
def test_build_sorted_list_and_selection_vs_debug_dumps(orc, name):
    g = golden("build_%s.npz" % name)
    y, j, n2 = g["Y"], g["J"], int(g["n2"])
    tt, P, Q, _ = orc.tru(j)
    a = np.array(y, order="F", copy=True)
    n = a.shape[0]
    for k in range(int(g["nsteps"])):
        info, dz, sl, step, it = orc.step_build(a, j, n2, P, Q, tt, id_ncp=2, s=k)
        lst, sel = g["list%d" % k], g["sel%d" % k]
        assert info == 0 and it == len(lst)
        assert np.array_equal(dz["w"][:it], lst["w"])
        assert np.array_equal(dz["p"][:it], lst["p"]) and np.array_equal(dz["q"][:it], lst["q"])
        assert np.array_equal(dz["b"][:it], lst["b"])
        assert sl == len(sel) and np.array_equal(step, sel["pos"])
        # advance A by one whole step through the oracle's own loop
        r = orc.solve(y, j, n2, max_steps=k + 1)
        a = np.array(r["A"], order="F", copy=True)
        for i in range(n):
            a[i, i] = r["SY"][i]


@pytest.mark.parametrize("cplx", [False, True])
def test_2x2_known_answers(orc, cplx):
    """dtest/ztest (TH1 = 1): U^T (resp. U^H), V = Z from DHSVD2, Z after JZTJ2."""
    g = golden("kat_%s.npz" % ("zhsvd2" if cplx else "dhsvd2"))
    for a, jj, info, U, V, Z in zip(g["A"], g["J"], g["INFO"], g["U"], g["V"], g["Z"]):
        f = orc.zhsvd2 if cplx else orc.dhsvd2
        A2 = np.array(a).reshape(2, 2, order="F")
        Ao, Uo, Vo, inf = f(A2, jj, th1=1.0)
        assert inf == info
        if inf < 0:
            continue
        Ut = Uo.conj().T if cplx else Uo.T
        Zt = Vo.conj().T.copy() if cplx else Vo.T.copy()
        for s in (0, 1):
            if jj[s] == -1:
                Zt[1, 0] = -Zt[1, 0]
                Zt[0, 1] = -Zt[0, 1]
        f2 = lambda m: np.ascontiguousarray(np.array(m).reshape(4, order="F"))
        for mine, ref in ((Ut, U), (Vo, V), (Zt, Z)):
            mine = f2(mine)
            # the driver prints NaN without its sign bit: compare NaN as a value
            nan = np.isnan(mine.view(np.float64)) & np.isnan(np.ascontiguousarray(ref).view(np.float64))
            assert np.array_equal(bits(mine)[~nan], bits(ref)[~nan])


def test_2x2_identities_random(orc):
    """A = U^T diag(s) Z^{-1} (the 2x2 HSVD), U orthogonal, Z J-unitary."""
    rng = np.random.default_rng(7)
    for k in range(3000):
        a = rng.standard_normal((2, 2)) * 10.0 ** rng.uniform(-5, 5)
        jj = (1, -1) if k % 2 else (1, 1)
        Ao, U, Z, info = orc.dhsvd2(a, jj, th1=1.0)
        assert info >= 0
        J = np.diag(jj).astype(float)
        s = np.diag([Ao[0, 0], Ao[1, 1]])
        rec = U.T @ s @ np.linalg.inv(Z)
        assert np.linalg.norm(rec - a) <= 1e-13 * np.linalg.norm(a)
        assert np.linalg.norm(U.T @ U - np.eye(2)) <= 1e-14
        assert np.linalg.norm(Z.T @ J @ Z - J) <= 1e-12 * max(1.0, np.linalg.norm(Z) ** 2)
        assert Ao[0, 0] >= 0 and Ao[1, 1] >= 0
        if jj == (1, 1):
            assert Ao[0, 0] >= Ao[1, 1]


@pytest.mark.parametrize("name", ["c1_r64", "c5_r40", "c3_z32"])
def test_golden_outputs_satisfy_hsvd_invariants(name):
    """Independent of any restatement: Y = YU diag(SY) ZZ, YU unitary,
    ZZ J ZZ^H = J, SY^2 = |eig(Y J Y^H)|, canonical SY order."""
    g = golden("solve_%s.npz" % name)
    y, j, yu, zz, sy = g["Y"], g["J"], g["YU"], g["ZZ"], g["SY"]
    n = y.shape[0]
    J = np.diag(j.astype(float))
    assert np.linalg.norm(yu @ np.diag(sy) @ zz - y) <= 1e3 * n * EPS * np.linalg.norm(y)
    assert np.linalg.norm(yu.conj().T @ yu - np.eye(n)) <= 1e3 * n * EPS
    assert np.linalg.norm(zz @ J @ zz.conj().T - J) <= 1e4 * n * EPS * np.linalg.norm(zz) ** 2
    ev = np.sort(np.abs(np.linalg.eigvals(y @ J @ y.conj().T)))
    s2 = np.sort(sy ** 2)
    # eig() of Y J Y^H is only normwise accurate for graded Y: compare normwise
    assert np.max(np.abs(ev - s2)) <= 1e-10 * np.max(s2)
    assert np.all(sy > 0)
    plus, minus = sy[j == 1], sy[j == -1]
    assert np.all(np.diff(plus) <= 0) and np.all(np.diff(minus) >= 0)


def test_hypot_matches_libm(orc):
    """The reference calls libm's hypot (CPython's math.hypot is its own,
    correctly rounded, routine, so libm is called through ctypes)."""
    import ctypes
    libm = ctypes.CDLL("libm.so.6")
    libm.hypot.restype = ctypes.c_double
    libm.hypot.argtypes = [ctypes.c_double, ctypes.c_double]
    rng = np.random.default_rng(3)
    x = rng.standard_normal(20000) * 10.0 ** rng.integers(-300, 300, 20000)
    y = rng.standard_normal(20000) * 10.0 ** rng.integers(-300, 300, 20000)
    x[:7] = [0.0, -0.0, np.inf, 5e-324, 1e308, 3.0, 1e-310]
    y[:7] = [0.0, 2.0, 1.0, 5e-324, 1e308, 4.0, 1e-310]
    for a, b in zip(x, y):
        assert orc.hypot(a, b) == libm.hypot(a, b)


def test_hypot_is_within_one_ulp_of_exact(orc):
    """What the mathematics fixes regardless of libm: |h - sqrt(x^2+y^2)| < 1 ulp."""
    from fractions import Fraction
    rng = np.random.default_rng(4)
    for _ in range(300):
        a, b = rng.standard_normal(2) * 10.0 ** rng.uniform(-20, 20, 2)
        h = orc.hypot(a, b)
        exact2 = Fraction(a) ** 2 + Fraction(b) ** 2
        lo, hi = np.nextafter(h, 0), np.nextafter(h, np.inf)
        assert Fraction(lo) ** 2 < exact2 < Fraction(hi) ** 2


@pytest.mark.parametrize("n", [3, 4, 7, 12])
def test_tru_closed_form(orc, n):
    rng = np.random.default_rng(n)
    j = np.where(rng.random(n) < 0.5, 1, -1)
    for id_tru in (1, 2):
        tt, P, Q, PN = orc.tru(j, id_tru)
        pairs = [(p, q) for q in range(2, n + 1) for p in range(1, q)] if id_tru == 1 else \
            [(p, q) for p in range(1, n) for q in range(p + 1, n + 1)]
        trig = [pq for pq in pairs if j[pq[0] - 1] == j[pq[1] - 1]]
        hyp = [pq for pq in pairs if j[pq[0] - 1] != j[pq[1] - 1]]
        assert tt == len(trig) and list(PN) == [int((j == 1).sum()), int((j == -1).sum())]
        assert list(zip(P[:tt], Q[:tt])) == trig
        assert list(zip(P[tt:], Q[tt:]))[::-1] == hyp
    assert orc.tru(np.array([1, 2, -1]), 1)[0] == -2


def test_jstep_len(orc):
    for n in range(0, 12):
        for n2 in range(-3, 9):
            want = max(n // 2, 0) if n2 == 0 else min(max(n2, 0), n // 2)
            assert orc.jstep_len(n, n2) == want


def test_sort_and_greedy_match_brute_force(orc):
    """AW order = (W desc, B desc, P desc); the greedy equals a brute-force
    walk; on tiny lists it is the lexicographically-first maximal matching."""
    rng = np.random.default_rng(11)
    for trial in range(200):
        n = int(rng.integers(3, 9))
        pairs = [(p, q) for q in range(2, n + 1) for p in range(1, q)]
        w = np.round(rng.standard_normal(len(pairs)), 1)  # many ties
        w[rng.random(len(pairs)) < 0.1] = -0.0
        dz = np.zeros(2 * len(pairs), orc.AW_DTYPE)
        for k, (p, q) in enumerate(pairs):
            dz[k] = (w[k], p, q, q - p, k + 1)
        perm = rng.permutation(len(pairs))
        dz[:len(pairs)] = dz[:len(pairs)][perm]
        import ctypes as C
        orc.lib().orc_srt.restype = C.c_int64
        orc.lib().orc_srt(len(pairs), len(dz), orc._p(dz))
        key = sorted(range(len(pairs)), key=lambda k: (-w[k], -(pairs[k][1] - pairs[k][0]), -pairs[k][0]))
        assert [tuple(x) for x in zip(dz["p"][:len(pairs)], dz["q"][:len(pairs)])] == [pairs[k] for k in key]
        n2 = int(rng.integers(1, n // 2 + 1))
        sl = C.c_int32(0)
        step = np.zeros(n2, np.int32)
        orc.lib().orc_ncp.restype = C.c_int64
        orc.lib().orc_ncp(2, n, len(pairs), len(dz), orc._p(dz), n2, C.byref(sl), orc._p(step))
        used, want = set(), []
        for pos, k in enumerate(key):
            p, q = pairs[k]
            if p not in used and q not in used:
                want.append(pos + 1)
                used |= {p, q}
            if len(want) == n2:
                break
        assert list(step[:sl.value]) == want


def test_ncp3_is_best_total_weight_over_start_offsets(orc):
    import ctypes as C
    rng = np.random.default_rng(5)
    for trial in range(50):
        n = int(rng.integers(4, 8))
        pairs = [(p, q) for q in range(2, n + 1) for p in range(1, q)]
        w = np.sort(rng.standard_normal(len(pairs)))[::-1]
        dz = np.zeros(2 * len(pairs), orc.AW_DTYPE)
        for k, (p, q) in enumerate(rng.permutation(pairs)):
            dz[k] = (w[k], p, q, q - p, k + 1)
        n2 = n // 2
        sl = C.c_int32(0)
        step = np.zeros(n2, np.int32)
        orc.lib().orc_ncp(3, n, len(pairs), len(dz), orc._p(dz), n2, C.byref(sl), orc._p(step))
        best = None
        for s0 in range(len(pairs)):
            used, sel = set(), []
            for pos in range(s0, len(pairs)):
                p, q = int(dz["p"][pos]), int(dz["q"][pos])
                if p not in used and q not in used:
                    sel.append(pos)
                    used |= {p, q}
                    if len(sel) == n2:
                        break
            tot = sum(dz["w"][k] for k in sel)
            if best is None or tot > best[0] + 1e-12:
                best = (tot, sel)
        got = sum(dz["w"][k - 1] for k in step[:sl.value])
        assert abs(got - best[0]) <= 1e-12


def test_loop_max_steps_prefix_matches_full_run(orc):
    g = golden("solve_c1_r64.npz")
    full = orc.solve(g["Y"], g["J"], 32)
    part = orc.solve(g["Y"], g["J"], 32, max_steps=10)
    assert part["info"] == 10
    for f in ["oldlen", "newlen", "trig", "hyp", "tw"]:
        assert np.array_equal(part["trace"][f][:10], full["trace"][f][:10])


# ---------------------------------------------------------------- cyclic sweeps (dcyc_loop)
def reference_sweep_pairs(n, order):
    """The pair order of the reference's pointwise sweeps, enumerated as its loops do
    (src/old/jk1.f90:105-116 row-cyclic, :125-136 column-cyclic), 1-based."""
    if order == 2:
        return [(i, j) for i in range(1, n) for j in range(i + 1, n + 1)]
    return [(i, j) for j in range(2, n + 1) for i in range(1, j)]


@pytest.mark.parametrize("order", [2, 3])
def test_cyclic_pointwise_visits_the_reference_order(orc, order):
    """Step s of the pointwise sweeps transforms pair s mod NN of the reference's order:
    the loop stopped after s + 1 steps has that pair's off-diagonal entries annihilated
    (DSTEP_TRANSF zeroes them) -- for steps across the first sweep and into the second."""
    n = 9
    rng = np.random.default_rng(3)
    y = rng.standard_normal((n, n))
    j = np.ones(n, dtype=np.int64)
    pairs = reference_sweep_pairs(n, order)
    nn = len(pairs)
    for s in list(range(0, nn, 5)) + [nn - 1, nn, nn + 7]:
        r = orc.solve_cyclic(y, j, max_steps=s + 1, order=order)
        p, q = pairs[s % nn]
        a = r["A"]
        assert a[p - 1, q - 1] == 0.0 and a[q - 1, p - 1] == 0.0, (s, p, q)
        # a pair the step did not touch keeps a non-zero off-diagonal entry somewhere
        assert np.count_nonzero(a) > 0


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("n,seed", [(24, 1), (37, 2)])
def test_cyclic_svd_matches_lapack(orc, order, n, seed):
    """J = +1: every pair is trigonometric and the HSVD is the SVD; the singular values of
    every cyclic order agree with LAPACK's (numpy.linalg.svd) to 8 N eps, relatively, and
    the factors are orthogonal with Y = YU diag(SY) ZZ to the same order."""
    rng = np.random.default_rng(seed)
    y = rng.standard_normal((n, n))
    j = np.ones(n, dtype=np.int64)
    r = orc.solve_cyclic(y, j, order=order)
    assert r["info"] > 0
    sv = np.linalg.svd(y, compute_uv=False)
    got = np.sort(np.abs(r["SY"]))[::-1]
    assert np.max(np.abs(got - sv) / sv) <= 8 * n * EPS
    yu, zz = r["YU"], r["ZZ"]
    assert np.linalg.norm(yu.T @ yu - np.eye(n), 2) <= 8 * n * EPS
    assert np.linalg.norm(y - (yu * r["SY"]) @ zz, 2) <= 8 * n * EPS * sv[0]


@pytest.mark.parametrize("order", [1, 2, 3])
def test_cyclic_hyperbolic_matches_eig(orc, order):
    """Mixed signs: sigma^2 = |eig(Y J Y^T)| per sign class, within eig's own error
    N eps (sigma + ||Y||^2 / sigma), and the J-orthogonality of ZZ.  The input is
    diagonally dominant: a fixed cyclic order has no guard against the growth of the
    hyperbolic transforms, and on a general input it can diverge (the dynamic
    ordering's weights are that guard)."""
    n = 30
    rng = np.random.default_rng(11)
    y = np.diag(rng.uniform(2.0, 4.0, n)) + 0.2 * rng.standard_normal((n, n))
    j = np.where(rng.random(n) < 0.5, 1, -1).astype(np.int64)
    r = orc.solve_cyclic(y, j, order=order)
    assert r["info"] > 0
    ev = np.linalg.eigvalsh(y @ np.diag(j.astype(float)) @ y.T)
    y2 = np.linalg.norm(y, 2) ** 2
    for cls in (1, -1):
        a = np.sort(np.abs(r["SY"][j == cls]))
        e = np.sort(np.sqrt(np.abs(ev[ev > 0] if cls > 0 else ev[ev < 0])))
        assert a.size == e.size
        assert np.max(np.abs(a - e) / (n * EPS * (e + y2 / e))) <= 1.0
    jd = np.diag(j.astype(float))
    zz = r["ZZ"]
    assert np.linalg.norm(zz @ jd @ zz.T - jd, 2) <= 64 * n * EPS * np.linalg.norm(zz, 2) ** 2


def test_cyclic_orders_differ(orc):
    """The three orders are three trajectories: the same values, other step counts or bits."""
    n = 20
    y = np.random.default_rng(4).standard_normal((n, n))
    j = np.ones(n, dtype=np.int64)
    rs = [orc.solve_cyclic(y, j, order=o) for o in (1, 2, 3)]
    assert rs[1]["info"] != rs[2]["info"] or not np.array_equal(bits(rs[1]["YU"]), bits(rs[2]["YU"]))
    for r in rs:
        assert np.allclose(np.sort(np.abs(r["SY"])), np.sort(np.abs(rs[0]["SY"])), rtol=1e-12)
