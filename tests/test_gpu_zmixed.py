"""The complex mixed solver (jkg_zsolve_mixed) through the C ABI.

Tolerances, as for the real mixed solver:
  - pair weights: trigonometric pairs and ZMAGF2H's early returns bit-identical
    to the oracle; the Gram-form hyperbolic weights within 32 N eps of the
    per-pair scale s_pq = (c_p + c_q)(1 + |W|_F^2) + |A_qp|^2 + |A_pq|^2,
    c = off-diagonal column norms^2, W = the pair's 2x2 column factor;
  - whole solves: the hyperbolic singular values within 64 N eps of the plain
    solver's (the reference's arithmetic), index by index and sorted within each
    J class (the spread between the reference's unfused and fused builds is
    printed beside it); reconstruction and unitarity residuals within 4x the plain
    solver's + 2 N eps and 64 N eps, the J-unitarity residual within 4x the
    plain solver's + 2 N eps.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps


@pytest.fixture(scope="module")
def ctx():
    from jkogb_amd import jkg
    c = jkg.Context(0)
    yield c
    c.close()


def inputs(cfg, n, seed):
    from jkogb_amd import inputs as gen
    return gen.make(cfg, n=n, seed=seed)


def check_zweights(ctx, orc, A, J):
    from jkogb_amd import jkg
    A = np.asfortranarray(A, dtype=np.complex128)
    n = A.shape[0]
    wf = orc.pair_weights(A, J)
    wm = jkg.keys_to_weights(ctx.dweights_mixed(A, J))
    O = A.copy()
    np.fill_diagonal(O, 0.0)
    c = np.sum(np.abs(O) ** 2, axis=0)
    ngram = 0
    for m in range(1, n):
        for p in range(m, 0, -1):
            q = p + n - m
            k = m * (m - 1) // 2 + (m - p)
            if np.isnan(wf[k]):
                assert np.isnan(wm[k]), (p, q)
                continue
            P, Q = p - 1, q - 1
            if J[P] == J[Q]:
                assert wm[k] == wf[k] and np.signbit(wm[k]) == np.signbit(wf[k]), (p, q)
                continue
            a2, u2, z2, info = orc.zhsvd2([[A[P, P], A[P, Q]], [A[Q, P], A[Q, Q]]], [J[P], J[Q]])
            if info < 0 or (info & 1) == 0 or (info & 4) == 0:
                assert wm[k] == wf[k], (p, q, info)
                continue
            ngram += 1
            s = (c[P] + c[Q]) * (1.0 + np.sum(np.abs(z2) ** 2)) + abs(A[Q, P]) ** 2 + abs(A[P, Q]) ** 2
            assert abs(wm[k] - wf[k]) <= 32 * n * EPS * s, (p, q, wm[k], wf[k], s)
    return ngram


@pytest.mark.parametrize("n,steps", [(40, 0), (48, 30), (37, 120)])
def test_zgram_weights_vs_oracle(ctx, orc, n, steps):
    Y, J = inputs("c3", n, 11)
    A = orc.solve(Y, J, n // 2, max_steps=steps)["A"] if steps else Y
    assert check_zweights(ctx, orc, A, J) > 0


def test_zgram_weights_tile_edges(ctx, orc):
    # J classes of sizes that leave partial tiles of the product and of the weight kernel
    rng = np.random.default_rng(5)
    n = 150
    Y = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * 10.0 ** rng.uniform(-3, 3, n)
    J = np.where(rng.random(n) < 0.8, 1, -1)
    assert check_zweights(ctx, orc, Y, J) > 0


def zresiduals(Y, J, out):
    n = Y.shape[0]
    YU, ZZ, SY = out["YU"], out["ZZ"], out["SY"]
    rec = np.linalg.norm(Y - (YU * SY) @ ZZ) / np.linalg.norm(Y)
    orth = np.linalg.norm(YU.conj().T @ YU - np.eye(n))
    Jd = np.diag(np.asarray(J, dtype=np.float64))
    jorth = np.linalg.norm(ZZ @ Jd @ ZZ.conj().T - Jd)
    return rec, orth, jorth


def sorted_by_class(SY, J):
    J = np.asarray(J)
    return np.sort(SY[J > 0]), np.sort(SY[J < 0])


@pytest.mark.parametrize("tol", [0.0, EPS])
@pytest.mark.parametrize("n", [64, 97, 160])
def test_zmixed_solve_matches_plain(ctx, n, tol):
    Y, J = inputs("c3", n, 3)
    plain = ctx.dsolve(Y, J, n // 2)
    fused = ctx.dsolve(Y, J, n // 2, fused=True)
    mix = ctx.dsolve_mixed(Y, J, n // 2, tol=tol)
    assert plain["info"] >= 0 and mix["info"] >= 0
    spread = np.max(np.abs(fused["SY"] - plain["SY"]) / np.abs(plain["SY"]))
    bound = 64 * n * EPS
    rel = np.abs(mix["SY"] - plain["SY"]) / np.abs(plain["SY"])
    print("N=%d: %.3f N eps from the plain solver; the reference builds' spread %.3f N eps"
          % (n, rel.max() / (n * EPS), spread / (n * EPS)))
    assert rel.max() <= bound, (rel.max(), spread)
    for a, b in zip(sorted_by_class(mix["SY"], J), sorted_by_class(plain["SY"], J)):
        assert np.max(np.abs(a - b) / np.abs(b)) <= bound
    rec, orth, jorth = zresiduals(Y, J, mix)
    prec, porth, pjorth = zresiduals(Y, J, plain)
    assert prec <= 64 * n * EPS and porth <= 64 * n * EPS, (prec, porth)  # the plain solver's own level
    assert rec <= min(4 * prec + 2 * n * EPS, 64 * n * EPS), (rec, prec)
    assert orth <= min(4 * porth + 2 * n * EPS, 64 * n * EPS), (orth, porth)
    assert jorth <= 4 * pjorth + 2 * n * EPS, (jorth, pjorth)
    if tol == 0.0:
        assert abs(mix["steps"] - plain["steps"]) <= 0.25 * plain["steps"], (mix["steps"], plain["steps"])
    else:
        assert mix["steps"] < plain["steps"], (mix["steps"], plain["steps"])


def test_zmixed_pair_weights_mode_is_the_plain_solver(ctx):
    Y, J = inputs("c3", 48, 9)
    plain = ctx.dsolve(Y, J, 24)
    mix = ctx.dsolve_mixed(Y, J, 24, weights=0, tol=0.0)
    bits = lambda a: np.ascontiguousarray(a).view(np.uint64)
    assert np.array_equal(bits(plain["SY"]), bits(mix["SY"]))
    assert np.array_equal(bits(plain["YU"]), bits(mix["YU"]))
    assert np.array_equal(bits(plain["ZZ"]), bits(mix["ZZ"]))
    assert plain["steps"] == mix["steps"]


def test_zmixed_rejects_controller(ctx):
    from jkogb_amd import jkg
    Y, J = inputs("c3", 16, 1)
    with pytest.raises(jkg.JKGError):
        ctx.dsolve_mixed(Y, J, 8, ctl_table=np.zeros(8, np.float32))


def test_zmixed_profile_stages(ctx):
    Y, J = inputs("c3", 128, 2)
    p = ctx.dprofile_mixed(Y, J, 64, steps=3)
    q = ctx.dprofile(Y, J, 64, steps=3)
    assert p["steps"] == 3 and q["steps"] == 3
    assert all(v >= 0.0 for v in p["stage_ms"].values())
    assert p["sum_abod"] > 0


# ---------------------------------------------------------------- FP32 warm start
def check_against_plain(ctx, Y, J, out):
    n = Y.shape[0]
    plain = ctx.dsolve(Y, J, n // 2)
    fused = ctx.dsolve(Y, J, n // 2, fused=True)
    spread = np.max(np.abs(fused["SY"] - plain["SY"]) / np.abs(plain["SY"]))
    bound = 64 * n * EPS
    rel = np.max(np.abs(out["SY"] - plain["SY"]) / np.abs(plain["SY"]))
    print("N=%d: %.3f N eps from the plain solver; the reference builds' spread %.3f N eps"
          % (n, rel / (n * EPS), spread / (n * EPS)))
    assert rel <= bound, (rel, spread)
    rec, orth, jorth = zresiduals(Y, J, out)
    prec, porth, pjorth = zresiduals(Y, J, plain)
    assert rec <= min(4 * prec + 2 * n * EPS, 64 * n * EPS), (rec, prec)
    assert orth <= min(4 * porth + 2 * n * EPS, 64 * n * EPS), (orth, porth)
    assert jorth <= 4 * pjorth + 2 * n * EPS, (jorth, pjorth)
    return plain


@pytest.mark.parametrize("key_bits", [64, 32])
@pytest.mark.parametrize("n", [96, 131, 192])
def test_warm_start_matches_plain(ctx, n, key_bits):
    # FP32 steps to a relative off-norm of 1e-5, polar factors of U and Z in FP64, A = U^H Y Z, FP64 steps:
    # the result is an HSVD of Y to the same accuracy as the FP64 solvers'
    Y, J = inputs("c3", n, 21)
    out = ctx.dsolve_mixed(Y, J, n // 2, fp32_warm=1, key_bits=key_bits)
    w = out["warm"]
    assert out["info"] >= 0
    assert w["fallback"] == 0 and w["warm_end"] in (1, 3), w
    assert w["steps32"] > 0 and w["ns_u"] >= 1 and w["ns_z"] >= 1, w
    assert w["off_rel"] < 1e-5 or w["warm_end"] == 3, w
    assert w["res_u"] < 1e-3 and w["res_z"] < 1e-2, w  # FP32-accurate factors before the first update
    check_against_plain(ctx, Y, J, out)


def test_warm_start_to_the_end_of_fp32(ctx):
    # a tolerance at the edge of what FP32 steps reach: the phase ends on it, on a stall or on convergence in
    # FP32, and the hand-over still gives the FP64 solvers' accuracy
    n = 80
    Y, J = inputs("c3", n, 22)
    out = ctx.dsolve_mixed(Y, J, n // 2, fp32_warm=1, switch_tol=1e-300)
    w = out["warm"]
    assert w["warm_end"] in (1, 2, 3) and w["fallback"] == 0, w
    assert w["off_rel"] < 1e-5, w
    check_against_plain(ctx, Y, J, out)


def test_warm_start_step_limit(ctx):
    # a step limit inside the FP32 phase: exactly that many steps, then the hand-over
    n = 64
    Y, J = inputs("c3", n, 23)
    out = ctx.dsolve_mixed(Y, J, n // 2, fp32_warm=1, max_steps=5)
    assert out["steps"] == 5 and out["warm"]["warm_end"] == 4 and out["warm"]["steps32"] == 5
    full = ctx.dsolve_mixed(Y, J, n // 2, fp32_warm=1)
    assert full["steps"] > 5


def test_warm_start_rejects(ctx):
    from jkogb_amd import jkg
    Y, J = inputs("c3", 16, 1)
    for kw in ({"fp32_warm": 1, "switch_tol": 0.0}, {"fp32_warm": 1, "switch_tol": 1.0}, {"fp32_warm": 2},
               {"key_bits": 16}):
        with pytest.raises(jkg.JKGError) as e:
            ctx.dsolve_mixed(Y, J, 8, **kw)
        assert e.value.code == -10, kw
    with pytest.raises(jkg.JKGError):  # 32-bit keys only with the warm start
        ctx.dsolve_mixed(Y, J, 8, key_bits=32)
    with pytest.raises(jkg.JKGError):  # the warm start needs the Gram-form weights
        ctx.dsolve_mixed(Y, J, 8, fp32_warm=1, weights=0, tol=0.0)
