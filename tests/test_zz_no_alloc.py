"""jkg_dsolve_noalloc through the C ABI: the allocation-free variant of jkg_dsolve
(device buffers in, device buffers out, every working buffer in the library's
fixed scratch array).  Its results are jkg_dsolve's bit for bit (the same kernels
and arithmetic), on two calls in a row, the second with a smaller input (the
array keeps the first call's contents), and it rejects sizes beyond its limit
without touching the outputs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from jkogb_amd import jkg
    c = jkg.Context(0)
    yield c
    c.close()


def device_call(ctx, y, j, n2, **kw):
    import torch
    n = y.shape[0]
    dev = torch.device("cuda", 0)
    d_y = torch.from_numpy(np.asfortranarray(y).reshape(-1, order="F").copy()).to(dev)
    d_yu = torch.full((n * n,), np.nan, dtype=torch.float64, device=dev)
    d_zz = torch.full((n * n,), np.nan, dtype=torch.float64, device=dev)
    d_sy = torch.full((n,), np.nan, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    r = ctx.dsolve_noalloc(n, d_y.data_ptr(), j, d_yu.data_ptr(), d_zz.data_ptr(), d_sy.data_ptr(), n2, **kw)
    torch.cuda.synchronize()
    r["YU"] = d_yu.cpu().numpy().reshape((n, n), order="F")
    r["ZZ"] = d_zz.cpu().numpy().reshape((n, n), order="F")
    r["SY"] = d_sy.cpu().numpy()
    return r


def same(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.uint64), np.ascontiguousarray(b).view(np.uint64))


def test_noalloc_matches_dsolve_on_two_calls(ctx):
    from jkogb_amd import inputs
    for cfg, n, seed, kw in (("c5", 200, 3, {}), ("c1", 96, 4, {"fused": True})):
        y, j = inputs.make(cfg, n=n, seed=seed)
        a = device_call(ctx, y, j, n // 2, **kw)
        b = ctx.dsolve(y, j, n // 2, **kw)
        assert a["info"] == b["info"] > 0 and a["steps"] == b["steps"], (cfg, a["info"], b["info"])
        for key in ("SY", "YU", "ZZ"):
            assert same(a[key], b[key]), (cfg, key)
        for f in ("step", "oldlen", "newlen", "trig", "hyp"):
            assert np.array_equal(a["trace"][f], b["trace"][f]), (cfg, f)
        assert same(a["trace"]["tw"], b["trace"]["tw"])
        assert a["tt"] == b["tt"] and a["pn"] == b["pn"]


def test_noalloc_step_limit(ctx):
    from jkogb_amd import inputs
    y, j = inputs.make("c2", n=64, seed=5)
    a = device_call(ctx, y, j, 32, max_steps=7)
    b = ctx.dsolve(y, j, 32, max_steps=7)
    assert a["info"] == b["info"] == 7 and a["steps"] == b["steps"]
    assert same(a["SY"], b["SY"]) and same(a["YU"], b["YU"])


def test_noalloc_rejects_sizes_beyond_its_limit(ctx):
    import torch
    from jkogb_amd import inputs, jkg
    n = 300
    y, j = inputs.make("c1", n=n, seed=6)
    dev = torch.device("cuda", 0)
    d_y = torch.from_numpy(y.reshape(-1, order="F").copy()).to(dev)
    out = torch.full((n * n,), 7.0, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    with pytest.raises(jkg.JKGError) as e:
        ctx.dsolve_noalloc(n, d_y.data_ptr(), j, out.data_ptr(), out.data_ptr(), out.data_ptr(), n // 2)
    assert e.value.code == jkg.JKG_ENOTSUP
    assert bool(torch.all(out == 7.0))
