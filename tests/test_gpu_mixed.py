"""The mixed solver (jkg_dsolve_mixed) through the C ABI.

Its weights are the reference's up to rounding, so the checks here are
tolerances, written in each test:
  - pair weights: every pair whose weight the mixed solver computes with the
    reference's own formula (trigonometric pairs, DMAGF2H's early returns) is
    bit-identical to the oracle; the Gram-form hyperbolic weights are within
    32 N eps of a per-pair scale s_pq = (c_p + c_q)(1 + |B|_F^2) + A_qp^2 + A_pq^2,
    c = off-diagonal column norms^2, B = the pair's 2x2 column factor;
  - whole solves: the hyperbolic singular values agree with the plain
    solver's (bit-exact with the reference) to 64 N eps relatively, index by
    index and sorted within each J class; the reconstruction, orthogonality and J-orthogonality residuals
    stay within 4x the plain solver's + 2 N eps (the first two also within
    64 N eps); with tol = 0 the step count stays within 25 % of the plain
    solver's, with tol = eps it is smaller.
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


def check_weights(ctx, orc, A, J):
    from jkogb_amd import jkg
    A = np.asfortranarray(A)
    n = A.shape[0]
    wf = orc.pair_weights(A, J)
    km = ctx.dweights_mixed(A, J)
    wm = jkg.keys_to_weights(km)
    O = A.copy()
    np.fill_diagonal(O, 0.0)
    c = np.einsum("ij,ij->j", O, O)
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
            a2, u2, z2, info = orc.dhsvd2([[A[P, P], A[P, Q]], [A[Q, P], A[Q, Q]]], [J[P], J[Q]])
            if info < 0 or (info & 1) == 0 or (info & 4) == 0:
                assert wm[k] == wf[k], (p, q, info)
                continue
            ngram += 1
            s = (c[P] + c[Q]) * (1.0 + np.sum(z2 * z2)) + A[Q, P] ** 2 + A[P, Q] ** 2
            assert abs(wm[k] - wf[k]) <= 32 * n * EPS * s, (p, q, wm[k], wf[k], s)
    return ngram


@pytest.mark.parametrize("cfg,n,steps", [("c1", 40, 0), ("c2", 48, 30), ("c5", 56, 80), ("c5", 33, 200)])
def test_gram_weights_vs_oracle(ctx, orc, cfg, n, steps):
    Y, J = inputs(cfg, n, 11)
    A = orc.solve(Y, J, n // 2, max_steps=steps)["A"] if steps else Y
    ngram = check_weights(ctx, orc, A, J)
    assert ngram > 0 or steps > 100


def test_gram_weights_tile_edges(ctx, orc):
    # J classes of sizes that leave partial 128-column tiles and 64-lane rows
    rng = np.random.default_rng(5)
    n = 150
    Y = rng.standard_normal((n, n)) * 10.0 ** rng.uniform(-3, 3, n)
    J = np.where(rng.random(n) < 0.8, 1, -1)
    assert check_weights(ctx, orc, Y, J) > 0


def residuals(Y, J, out):
    n = Y.shape[0]
    YU, ZZ, SY = out["YU"], out["ZZ"], out["SY"]
    rec = np.linalg.norm(Y - (YU * SY) @ ZZ) / np.linalg.norm(Y)
    orth = np.linalg.norm(YU.T @ YU - np.eye(n))
    Jd = np.diag(np.asarray(J, dtype=np.float64))
    jorth = np.linalg.norm(ZZ @ Jd @ ZZ.T - Jd)
    return rec, orth, jorth


def sorted_by_class(SY, J):
    J = np.asarray(J)
    return np.sort(SY[J > 0]), np.sort(SY[J < 0])


@pytest.mark.parametrize("tol", [0.0, EPS])
@pytest.mark.parametrize("cfg,n", [("c1", 96), ("c2", 128), ("c5", 160), ("c1", 203)])
def test_mixed_solve_matches_plain(ctx, cfg, n, tol):
    Y, J = inputs(cfg, n, 3)
    plain = ctx.dsolve(Y, J, n // 2)
    mix = ctx.dsolve_mixed(Y, J, n // 2, tol=tol)
    assert plain["info"] >= 0 and mix["info"] >= 0
    # each J class ends sorted as the reference sorts it, so SY compares index by index
    rel = np.abs(mix["SY"] - plain["SY"]) / np.abs(plain["SY"])
    assert rel.max() <= 64 * n * EPS, rel.max()
    for a, b in zip(sorted_by_class(mix["SY"], J), sorted_by_class(plain["SY"], J)):
        rel = np.abs(a - b) / np.abs(b)
        assert rel.max() <= 64 * n * EPS, rel.max()
    rec, orth, jorth = residuals(Y, J, mix)
    prec, porth, pjorth = residuals(Y, J, plain)
    # as good as the plain solver's (the reference's own level), to a factor, and 64 N eps at most
    assert rec <= min(4 * prec + 2 * n * EPS, 64 * n * EPS), (rec, prec)
    assert orth <= min(4 * porth + 2 * n * EPS, 64 * n * EPS), (orth, porth)
    # J-orthogonality: hyperbolic factors may be ill-conditioned, so relative to the plain solver's only
    assert jorth <= 4 * pjorth + 2 * n * EPS, (jorth, pjorth)
    if tol == 0.0:  # the reference's termination: about as many steps
        assert abs(mix["steps"] - plain["steps"]) <= 0.25 * plain["steps"], (mix["steps"], plain["steps"])
    else:  # the convergence test ends the rounding-level tail
        assert mix["steps"] < plain["steps"], (mix["steps"], plain["steps"])


def test_mixed_pair_weights_mode_is_the_plain_solver(ctx):
    # weights=JKG_W_PAIR runs the reference's weights: the plain trajectory bit for bit
    Y, J = inputs("c2", 64, 9)
    plain = ctx.dsolve(Y, J, 32)
    mix = ctx.dsolve_mixed(Y, J, 32, weights=0, tol=0.0)
    assert np.array_equal(plain["SY"].view(np.uint64), mix["SY"].view(np.uint64))
    assert np.array_equal(plain["YU"].view(np.uint64), mix["YU"].view(np.uint64))
    assert plain["steps"] == mix["steps"]


def test_mixed_rejects_unsupported(ctx):
    from jkogb_amd import jkg
    Y, J = inputs("c1", 16, 1)
    with pytest.raises(jkg.JKGError) as e:
        ctx.dsolve_mixed(Y, J, 8, weights=7)
    assert e.value.code == -10


def test_mixed_graph_follows_options(ctx):
    # successive solves of one size with other options must not reuse a step graph captured for the first
    Y, J = inputs("c2", 96, 2)
    a = ctx.dsolve_mixed(Y, J, 48, tol=0.0)
    b = ctx.dsolve_mixed(Y, J, 48, tol=EPS)
    c = ctx.dsolve_mixed(Y, J, 48, tol=0.0)
    assert b["steps"] < a["steps"]
    assert a["steps"] == c["steps"] and np.array_equal(a["SY"], c["SY"])


@pytest.mark.parametrize("cfg,n", [("c1", 128), ("c2", 160)])
def test_mixed_threshold(ctx, cfg, n):
    # a walk cut at tau * (largest weight) changes the trajectory; the result stays within 64 N eps of
    # the plain solver (the reference's arithmetic)
    Y, J = inputs(cfg, n, 6)
    plain = ctx.dsolve(Y, J, n // 2)
    fused = ctx.dsolve(Y, J, n // 2, fused=True)
    base = ctx.dsolve_mixed(Y, J, n // 2)
    thr = ctx.dsolve_mixed(Y, J, n // 2, threshold=1e-6)
    assert thr["info"] >= 0
    assert thr["sum_it"] != base["sum_it"] or thr["steps"] != base["steps"]
    spread = np.max(np.abs(fused["SY"] - plain["SY"]) / np.abs(plain["SY"]))
    rel = np.max(np.abs(thr["SY"] - plain["SY"]) / np.abs(plain["SY"]))
    print("%s N=%d: %.2f N eps from the plain solver; the reference builds' spread %.2f N eps"
          % (cfg, n, rel / (n * EPS), spread / (n * EPS)))
    assert rel <= 64 * n * EPS, (rel, spread)
    rec, orth, jorth = residuals(Y, J, thr)
    prec, porth, pjorth = residuals(Y, J, plain)
    assert rec <= 4 * prec + 2 * n * EPS and orth <= 4 * porth + 2 * n * EPS, (rec, orth)


def test_mixed_threshold_rejects_out_of_range(ctx):
    from jkogb_amd import jkg
    Y, J = inputs("c1", 16, 1)
    for t in (-1.0, 1.0, float("nan")):
        with pytest.raises(jkg.JKGError):
            ctx.dsolve_mixed(Y, J, 8, threshold=t)


# ---------------------------------------------------------------- FP32 warm start (real)
@pytest.mark.parametrize("cfg,n,key_bits", [("c1", 64, 64), ("c5", 80, 32), ("c2", 96, 64)])
def test_real_warm_start_vs_oracle(ctx, orc, cfg, n, key_bits):
    """FP32 steps until the relative off-norm is below 1e-5, polar factors of U and Z in FP64,
    A = U^T Y Z, FP64 steps: every singular value within 64 N eps of the oracle's (the
    reference's arithmetic, run on the CPU), orthogonality and reconstruction to the same order."""
    Y, J = inputs(cfg, n, 17)
    out = ctx.dsolve_mixed(Y, J, n // 2, fp32_warm=1, key_bits=key_bits)
    w = out["warm"]
    assert out["info"] >= 0
    assert w["fallback"] == 0 and w["warm_end"] in (1, 2, 3), w
    assert w["steps32"] > 0 and w["ns_u"] >= 1 and w["ns_z"] >= 1, w
    o = orc.solve(Y, J, n // 2)
    rel = np.abs(out["SY"] - o["SY"]) / np.abs(o["SY"])
    assert rel.max() <= 64 * n * EPS, (rel.max() / (n * EPS))
    rec, orth, jorth = residuals(Y, J, out)
    orec, oorth, ojorth = residuals(Y, J, o)
    assert rec <= 64 * n * EPS and orth <= 64 * n * EPS, (rec, orth)
    assert jorth <= 4 * ojorth + 2 * n * EPS, (jorth, ojorth)


@pytest.mark.parametrize("cfg,n", [("c5", 160), ("c2", 203)])
def test_real_warm_start_vs_plain(ctx, cfg, n):
    """Larger sizes against the plain solver (bit-identical to the oracle, see the parity tests):
    64 N eps index by index, residuals within 64 N eps."""
    Y, J = inputs(cfg, n, 18)
    out = ctx.dsolve_mixed(Y, J, n // 2, fp32_warm=1, key_bits=32)
    plain = ctx.dsolve(Y, J, n // 2)
    assert out["warm"]["fallback"] == 0
    rel = np.abs(out["SY"] - plain["SY"]) / np.abs(plain["SY"])
    assert rel.max() <= 64 * n * EPS, (rel.max() / (n * EPS))
    rec, orth, jorth = residuals(Y, J, out)
    prec, porth, pjorth = residuals(Y, J, plain)
    assert rec <= 64 * n * EPS and orth <= 64 * n * EPS, (rec, orth)
    assert jorth <= 4 * pjorth + 2 * n * EPS, (jorth, pjorth)


def test_real_warm_start_step_limit_and_frames(ctx):
    n = 64
    Y, J = inputs("c2", n, 19)
    out = ctx.dsolve_mixed(Y, J, n // 2, fp32_warm=1, max_steps=5)
    assert out["steps"] == 5 and out["warm"]["warm_end"] == 4 and out["warm"]["steps32"] == 5
    ctx.set_frames(every=1, size=8, max_frames=4096)
    try:
        full = ctx.dsolve_mixed(Y, J, n // 2, fp32_warm=1)
        assert ctx.frames().shape[0] == full["steps"]
    finally:
        ctx.set_frames(0)
    assert full["steps"] > 5
