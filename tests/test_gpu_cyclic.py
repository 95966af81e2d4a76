"""The cyclic sweep baselines (jkg_mixed_opts.ordering = 1 round robin, 2 row-cyclic
and 3 column-cyclic pointwise, the reference's sweeps of src/old/jk1.f90:101-146)
through the C ABI.

Without a convergence tolerance their steps are the oracle's dcyc_loop bit for bit
(YU, ZZ, SY, INFO and the trace), on inputs where cyclic sweeps converge; on the
ordinary SVD (J = +1) the values match LAPACK's to 64 N eps.  (On hyperbolic
inputs cyclic sweeps can be far less accurate than the dynamic ordering, or
diverge -- the pointwise sweeps on the Gaussian C1/C2 inputs do: that is the
baseline's point, and not tested here.)
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


def bits(a):
    return np.ascontiguousarray(a).view(np.uint64)


def svd_case(n, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, n)), np.ones(n, dtype=np.int64)


def dominant_case(n, seed):
    """Diagonally dominant, mixed signs: the pointwise sweeps converge on it."""
    rng = np.random.default_rng(seed)
    y = np.diag(rng.uniform(2.0, 4.0, n)) + 0.2 * rng.standard_normal((n, n))
    return y, np.where(rng.random(n) < 0.5, 1, -1).astype(np.int64)


def cases(order):
    from jkogb_amd import inputs
    out = [("svd32", *svd_case(32, 1)), ("svd47", *svd_case(47, 2)), ("dom30", *dominant_case(30, 11))]
    cfgs = [("c2", 48), ("c5", 40), ("c1", 33)] if order == 1 else [("c5", 40)]
    for cfg, n in cfgs:
        out.append(("%s%d" % (cfg, n), *inputs.make(cfg, n=n, seed=3)))
    return out


@pytest.mark.parametrize("order", [1, 2, 3])
@pytest.mark.parametrize("fused", [False, True])
def test_cyclic_bit_exact_vs_oracle(ctx, orc, fused, order):
    for name, y, j in cases(order):
        n = y.shape[0]
        r = ctx.dsolve_mixed(y, j, n // 2, ordering=order, tol=0.0, fused=fused, trace_cap=40000)
        o = orc.solve_cyclic(y, j, fused=fused, order=order, max_trace=40000)
        assert o["info"] > 0 and r["info"] == o["info"], name
        assert np.array_equal(bits(r["SY"]), bits(o["SY"])), name
        assert np.array_equal(bits(r["YU"]), bits(o["YU"])), name
        assert np.array_equal(bits(r["ZZ"]), bits(o["ZZ"])), name
        t = r["trace"]
        assert len(t) == len(o["trace"])
        for f in ["step", "oldlen", "newlen", "trig", "hyp"]:
            assert np.array_equal(t[f], o["trace"][f]), (name, f)
        assert np.array_equal(bits(t["tw"]), bits(o["trace"]["tw"])), name


@pytest.mark.parametrize("order", [1, 2, 3])
def test_cyclic_svd_with_convergence_test(ctx, order):
    for n, seed in [(64, 5), (101, 6)]:
        y, j = svd_case(n, seed)
        r = ctx.dsolve_mixed(y, j, n // 2, ordering=order)
        d = ctx.dsolve(y, j, n // 2)
        sv = np.sort(np.linalg.svd(y, compute_uv=False))
        assert np.max(np.abs(np.sort(r["SY"]) - sv) / sv) <= 64 * n * EPS
        assert r["steps"] > d["steps"]  # the dynamic ordering needs fewer steps


def test_cyclic_rejects(ctx):
    from jkogb_amd import jkg
    y, j = svd_case(16, 1)
    for bad in (4, -1):
        with pytest.raises(jkg.JKGError) as e:
            ctx.dsolve_mixed(y, j, 8, ordering=bad)
        assert e.value.code == -10
    for order in (1, 2, 3):
        for kw in ({"threshold": 1e-6}, {"ctl_table": np.zeros(8, np.float32)}):
            with pytest.raises(jkg.JKGError):
                ctx.dsolve_mixed(y, j, 8, ordering=order, **kw)
        with pytest.raises(jkg.JKGError):
            ctx.dsolve_mixed(y.astype(np.complex128), j, 8, ordering=order)
