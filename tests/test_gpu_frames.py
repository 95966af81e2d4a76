"""Per-step frames of |A| (jkg_set_frames) through the C ABI: the first frame is the
quantised input (checked against a numpy restatement of the map, within one level
for log10 rounding at a level boundary), later frames follow the solve, the count
follows the schedule, and the feature switches off."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from jkogb_amd import jkg
    c = jkg.Context(0)
    yield c
    c.set_frames(0)
    c.close()


def quantise(A, size):
    n = A.shape[0]
    out = np.zeros((size, size), np.int64)
    for u in range(size):
        r0, r1 = u * n // size, (u + 1) * n // size
        for v in range(size):
            c0, c1 = v * n // size, (v + 1) * n // size
            m = np.max(np.abs(A[r0:r1, c0:c1])) if r1 > r0 and c1 > c0 else 0.0
            out[u, v] = 0 if m == 0 else min(255, max(1, 1 + int(np.floor((np.log10(m) + 16.0) * 8.0))))
    return out


def test_frames_of_a_real_solve(ctx):
    from jkogb_amd import inputs
    y, j = inputs.make("c2", n=40, seed=4)
    ctx.set_frames(every=1, size=16, max_frames=4096)
    out = ctx.dsolve(y, j, 20)
    f = ctx.frames()
    assert f.shape == (out["steps"], 16, 16)
    assert np.max(np.abs(f[0].astype(np.int64) - quantise(y, 16))) <= 1
    # at the end the off-diagonal blocks sit far below the diagonal ones
    last = f[-1].astype(np.int64)
    assert np.mean(np.diag(last)) > np.mean(last[~np.eye(16, dtype=bool)]) + 40
    ctx.set_frames(every=3, size=8, max_frames=5)
    out = ctx.dsolve(y, j, 20)
    assert ctx.frames().shape == (min((out["steps"] - 1) // 3 + 1, 5), 8, 8)
    ctx.set_frames(0)
    ctx.dsolve(y, j, 20)
    assert ctx.frames().shape[0] == 0


def test_frames_of_complex_and_warm_start_solves(ctx):
    from jkogb_amd import inputs
    y, j = inputs.make("c3", n=48, seed=5)
    ctx.set_frames(every=2, size=12, max_frames=2048)
    for kw in ({}, {"fp32_warm": 1}):
        out = ctx.dsolve_mixed(y, j, 24, **kw)
        f = ctx.frames()
        assert f.shape[0] == (out["steps"] - 1) // 2 + 1
        assert np.max(np.abs(f[0].astype(np.int64) - quantise(y, 12))) <= 1
    ctx.set_frames(0)


def test_frames_set_twice_with_the_same_arguments(ctx):
    """jkg_set_frames replaces its device buffer on every call.  A step graph captured with the
    old buffer must not be replayed with the new one: same schedule and size, other buffer (the
    old address is taken by another context's frames in between, so the new one differs)."""
    from jkogb_amd import inputs, jkg
    y, j = inputs.make("c2", n=40, seed=4)
    ctx.set_frames(every=1, size=16, max_frames=4096)
    first = ctx.dsolve(y, j, 20)
    f1 = ctx.frames()
    other = jkg.Context(0)
    try:
        ctx.set_frames(every=1, size=8, max_frames=16)
        other.set_frames(every=1, size=16, max_frames=4096)
        ctx.set_frames(every=1, size=16, max_frames=4096)
        second = ctx.dsolve(y, j, 20)
        f2 = ctx.frames()
    finally:
        other.set_frames(0)
        other.close()
    assert second["steps"] == first["steps"] and f2.shape == f1.shape
    assert np.array_equal(f1, f2)
    assert np.max(np.abs(f2[0].astype(np.int64) - quantise(y, 16))) <= 1
    ctx.set_frames(0)
