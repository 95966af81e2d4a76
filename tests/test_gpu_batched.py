"""The batched solver (jkg_dsolve_batched) through the C ABI: every instance
follows the plain solver's arithmetic, so each one's YU, ZZ, SY and INFO must
equal, bit for bit, what jkg_dsolve returns for it alone (which the parity
tests pin to the reference), whatever the launch grid, kernel variant and
schedule."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from jkogb_amd import jkg
    c = jkg.Context(0)
    yield c
    c.close()


def batch(seed, count, lo, hi):
    rng = np.random.default_rng(seed)
    ns = rng.integers(lo, hi + 1, count)
    ns[0] = 3
    Ys, Js = [], []
    for n in ns:
        Ys.append(rng.standard_normal((n, n)) * (10.0 ** rng.uniform(-2, 2, n) if n % 3 == 0 else 1.0))
        J = np.where(rng.random(n) < 0.5, 1, -1)
        if n % 7 == 0:
            J[:] = 1
        Js.append(J)
    return Ys, Js


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def check(ctx, Ys, Js, outs, fused=False, max_steps=-1):
    for i, (Y, J, o) in enumerate(zip(Ys, Js, outs)):
        n = Y.shape[0]
        r = ctx.dsolve(Y, J, n // 2, fused=fused, max_steps=max_steps)
        assert o["info"] == r["info"], (i, n, o["info"], r["info"])
        assert same(o["SY"], r["SY"]), (i, n)
        assert same(o["YU"], r["YU"]), (i, n)
        assert same(o["ZZ"], r["ZZ"]), (i, n)


def test_batched_matches_single_solves(ctx):
    Ys, Js = batch(1, 48, 3, 96)
    outs, run = ctx.dsolve_batched(Ys, Js)
    print("batched: %d instances, %.3f s, %d steps, grid %d, variant %d" % (len(Ys), run["seconds"], run["steps"],
                                                                          run["grid"], run["variant"]))
    check(ctx, Ys, Js, outs)


@pytest.mark.parametrize("grid,variant", [(256, 0), (512, 1), (1024, 2), (2048, 3)])
def test_batched_every_variant(ctx, grid, variant):
    Ys, Js = batch(2, 24, 3, 70)
    outs, run = ctx.dsolve_batched(Ys, Js, grid=grid, variant=variant)
    assert run["grid"] == grid and run["variant"] == variant
    check(ctx, Ys, Js, outs)


def test_batched_prefix_and_fused(ctx):
    Ys, Js = batch(3, 16, 3, 60)
    outs, _ = ctx.dsolve_batched(Ys, Js, max_steps=7)
    check(ctx, Ys, Js, outs, max_steps=7)
    outs, _ = ctx.dsolve_batched(Ys, Js, fused=True)
    check(ctx, Ys, Js, outs, fused=True)


def test_batched_against_the_oracle(ctx, orc):
    Ys, Js = batch(4, 6, 8, 40)
    outs, _ = ctx.dsolve_batched(Ys, Js)
    for Y, J, o in zip(Ys, Js, outs):
        r = orc.solve(Y, J, Y.shape[0] // 2)
        assert o["info"] == r["info"] and same(o["SY"], r["SY"]) and same(o["YU"], r["YU"])


def test_batched_larger_sizes(ctx):
    # sizes that use several weight tiles and walk chunks
    Ys, Js = batch(5, 6, 150, 256)
    outs, _ = ctx.dsolve_batched(Ys, Js)
    check(ctx, Ys, Js, outs)


def test_batched_rejects_bad_j(ctx):
    from jkogb_amd import jkg
    Ys, Js = batch(6, 3, 5, 10)
    Js[1] = Js[1].copy()
    Js[1][0] = 2
    with pytest.raises(jkg.JKGError) as e:
        ctx.dsolve_batched(Ys, Js)
    assert e.value.code == jkg.JKG_EJ


def test_batched_rejects_a_grid_beyond_the_maximum(ctx):
    """grid above JKG_BATCH_GRID_MAX (4096) is argument 11 (the options), before any device work."""
    from jkogb_amd import jkg
    Ys, Js = batch(7, 3, 5, 10)
    for grid in (4097, 1 << 20, -1):
        with pytest.raises(jkg.JKGError) as e:
            ctx.dsolve_batched(Ys, Js, grid=grid)
        assert e.value.code == -11
    outs, run = ctx.dsolve_batched(Ys, Js, grid=4096, variant=0)
    assert run["grid"] == 4096
    check(ctx, Ys, Js, outs)


def test_batched_convergence_test(ctx):
    # with the mixed solver's convergence test: each instance within the spread of the reference's two
    # FMA builds (+64 N eps) of its plain solve, in no more steps
    eps = np.finfo(np.float64).eps
    Ys, Js = batch(7, 12, 20, 90)
    outs, _ = ctx.dsolve_batched(Ys, Js, tol=eps)
    for Y, J, o in zip(Ys, Js, outs):
        n = Y.shape[0]
        p = ctx.dsolve(Y, J, n // 2)
        f = ctx.dsolve(Y, J, n // 2, fused=True)
        m = ctx.dsolve_mixed(Y, J, n // 2)
        spread = np.max(np.abs(f["SY"] - p["SY"]) / np.abs(p["SY"]))
        rel = np.max(np.abs(o["SY"] - p["SY"]) / np.abs(p["SY"]))
        assert rel <= 4 * spread + 64 * n * eps, (n, rel, spread)
        assert 0 <= o["info"] <= p["info"]


def test_convergence_test_on_a_stagnating_instance(ctx):
    # an instance of the batched configuration (N=208) on which the reference's termination never
    # triggers: its weights underflow to 0 after ~15000 steps while transforms continue, and the
    # diagonal is then still far from the HSVD (smallest value 0.0566 against the exact 0.00632).
    # With the convergence test both the batched and the mixed solver stop near 2000 steps at the
    # exact values, sqrt|eig(Y J Y^T)| per class -- up to the float64 eigenvalues' own error,
    # |d sigma| <= N eps |Y|^2 / sigma, so each value is checked to 64 N eps (sigma + |Y|^2 / sigma).
    from jkogb_amd import inputs
    eps = np.finfo(np.float64).eps
    n = 208
    y, j = inputs.make("c4", n=n, seed=12370)
    ev = np.linalg.eigvalsh(y @ np.diag(j.astype(float)) @ y.T)
    y2 = np.linalg.norm(y, 2) ** 2
    b, _ = ctx.dsolve_batched([y], [j], tol=eps)
    m = ctx.dsolve_mixed(y, j, n // 2)
    for out in (b[0], m):
        assert 0 < out["info"] < 4000
        for cls in (1, -1):
            a = np.sort(np.abs(out["SY"][j == cls]))
            e = np.sort(np.sqrt(np.abs(ev[ev > 0] if cls > 0 else ev[ev < 0])))
            assert np.all(np.abs(a - e) <= 64 * n * eps * (e + y2 / e))
    # the two agree with each other to rounding
    assert np.max(np.abs(b[0]["SY"] - m["SY"]) / np.abs(m["SY"])) <= 64 * n * eps
    p = ctx.dsolve(y, j, n // 2, max_steps=5000)
    assert p["info"] == 5000  # still running where the convergence test has long ended
