"""Test matrices with a prescribed hyperbolic spectrum (jkg_dgen) on the GPU.

Pinned by the mathematics, not by the oracle: eig(G J G^T) equals lambda (to
O(N eps) of |G|^2), and every solver run on (G, J) finds sqrt|lambda|; the
solvers' relative error is bounded by 4x the CPU oracle's on the same (G, J)
+ 64 N eps.
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


def by_class(v, J):
    J = np.asarray(J)
    return np.sort(v[J > 0]), np.sort(v[J < 0])


@pytest.mark.parametrize("n,tmax", [(64, 0.5), (97, 0.25), (160, 0.5)])
def test_generated_spectrum(ctx, n, tmax):
    from jkogb_amd import jkdatgen
    lam = jkdatgen.genlam(n, 1, 2, 1e-2, 1.0)
    G, J, lam = jkdatgen.generate(ctx, lam, tmax=tmax, seed=7)
    assert np.array_equal(J, np.where(lam > 0, 1, -1))
    ev = np.linalg.eigvalsh(G @ np.diag(J.astype(float)) @ G.T)
    assert np.max(np.abs(np.sort(ev) - np.sort(lam))) <= 16 * n * EPS * np.linalg.norm(G, 2) ** 2
    again, _, _ = jkdatgen.generate(ctx, lam, tmax=tmax, seed=7)
    assert np.array_equal(G, again)  # deterministic for a seed


@pytest.mark.parametrize("n,tmax", [(64, 0.5), (97, 0.25)])
def test_solvers_find_the_prescribed_values(ctx, orc, n, tmax):
    from jkogb_amd import jkdatgen
    lam = jkdatgen.genlam(n, 2, 3, 1e-1, 1.0)
    G, J, lam = jkdatgen.generate(ctx, lam, tmax=tmax, seed=3)
    sigma = np.sqrt(np.abs(lam))
    exact = by_class(sigma, J)

    def err(sy):
        return max(np.max(np.abs(a - b) / b) for a, b in zip(by_class(sy, J), exact))

    o = orc.solve(G, J, n // 2)
    e_orc = err(o["SY"])
    bound = 4 * e_orc + 64 * n * EPS
    for name, out in [("plain", ctx.dsolve(G, J, n // 2)), ("mixed", ctx.dsolve_mixed(G, J, n // 2)),
                      ("ncp3", ctx.dsolve(G, J, n // 2, id_ncp=3))]:
        assert out["info"] >= 0, name
        assert err(out["SY"]) <= bound, (name, err(out["SY"]), e_orc)
    # the complex solvers on the same (real) data
    Gc = G.astype(np.complex128)
    for name, out in [("zplain", ctx.dsolve(Gc, J, n // 2)), ("zwarm", ctx.dsolve_mixed(Gc, J, n // 2, fp32_warm=1))]:
        assert err(out["SY"]) <= bound, (name, err(out["SY"]), e_orc)


def test_jkdatgen_files(ctx, tmp_path):
    # the command writes the drop-in's input files and the eigenvalues
    from jkogb_amd import inputs, jkdatgen
    fn = str(tmp_path / "g48")
    assert jkdatgen.main(["2", "1", "48", fn, "2", "0.05", "2.0"]) == 0
    Y, J = inputs.read_yj(fn, 48)
    lam = np.fromfile(fn + ".LAM", dtype=np.float64)
    assert lam.size == 48 and np.array_equal(J, np.where(lam > 0, 1, -1))
    assert np.all(J[:int(np.sum(J > 0))] == 1)  # JPART order
    out = ctx.dsolve(Y, J, 24)
    got = by_class(out["SY"], J)
    want = by_class(np.sqrt(np.abs(lam)), J)
    for a, b in zip(got, want):
        assert np.max(np.abs(a - b) / b) <= 1e-12
