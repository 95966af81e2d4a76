"""The device radix sort through the ASRT plug-in (jkg_srt), on lists large enough for thousands of
tiles and on key distributions that exercise every record width and constant digits.

The expected order is AW_CMP's (src/shared/atypes.F90:85-120): weight descending with -0 == +0,
then band q - p descending, then p descending -- computed here independently by numpy's lexsort
on (key(w), pair index), where key(w) maps the IEEE bits to an integer that grows as w decreases
and the pair index enumerates the pairs in the tie-break order.  Small lists are also checked
against the oracle's ASRT restatement.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from jkogb_amd import jkg
    c = jkg.Context(0)
    yield c
    c.close()


def weight_keys(w):
    u = np.ascontiguousarray(w, np.float64).view(np.uint64).copy()
    u[w == 0.0] = 0
    neg = (u >> np.uint64(63)) == 1
    ordv = np.where(neg, ~u, u | np.uint64(1 << 63))
    return ~ordv


def pair_index(n, p, q):
    m = n - (q - p)
    return m * (m - 1) // 2 + (m - p)


def make_list(n, w, rng):
    q, p = np.triu_indices(n, 1)[1] + 1, np.triu_indices(n, 1)[0] + 1
    perm = rng.permutation(p.size)
    from oracle import orc
    dz = np.zeros(p.size, orc.AW_DTYPE)
    dz["w"] = w[perm]
    dz["p"] = p[perm]
    dz["q"] = q[perm]
    dz["b"] = q[perm] - p[perm]
    dz["i"] = np.arange(1, p.size + 1)
    return dz


def expected_order(dz, n):
    k = pair_index(n, dz["p"].astype(np.int64), dz["q"].astype(np.int64))
    return np.lexsort((k, weight_keys(dz["w"])))


DISTS = {
    # every digit varies: all eight passes run, records shrink from 12 to 8 to 4 bytes
    "wide": lambda rng, m: rng.standard_normal(m) * 10.0 ** rng.uniform(-150, 150, m),
    # a handful of values: the low six digits are the same in every key (long equal-digit runs)
    "ties": lambda rng, m: rng.integers(-3, 4, m).astype(np.float64) * 0.5,
    # FP32 values: the low 29 mantissa bits are zero, the lowest three digits constant
    "fp32": lambda rng, m: rng.standard_normal(m).astype(np.float32).astype(np.float64),
    # one value: every pass keeps the input order (ties by pair index only)
    "const": lambda rng, m: np.full(m, 2.5),
    # +-0 and tiny subnormals next to ordinary values
    "zeros": lambda rng, m: np.where(rng.random(m) < 0.5, np.where(rng.random(m) < 0.5, 0.0, -0.0),
                                     rng.standard_normal(m) * 1e-310),
}


@pytest.mark.parametrize("n", [5, 91, 92, 1500, 2900])
@pytest.mark.parametrize("dist", sorted(DISTS))
def test_srt_order(ctx, n, dist):
    if n == 2900 and dist not in ("wide", "const"):
        pytest.skip("the largest list only for the widest and the narrowest keys")
    rng = np.random.default_rng(n * 7 + len(dist))
    m = n * (n - 1) // 2
    dz = make_list(n, DISTS[dist](rng, m), rng)
    print("srt: n=%d (%d pairs, %d tiles of 4096), %s keys" % (n, m, (m + 4095) // 4096, dist), flush=True)
    a = dz.copy()
    ctx.srt(a)
    e = dz[expected_order(dz, n)]
    assert np.array_equal(a["p"], e["p"]) and np.array_equal(a["q"], e["q"])
    assert np.array_equal(a["w"].view(np.uint64), e["w"].view(np.uint64))
    if m <= 5000:
        from oracle import orc
        b = np.zeros(2 * m, orc.AW_DTYPE)
        b[:m] = dz
        orc.lib().orc_srt.restype = C.c_int64
        orc.lib().orc_srt(m, len(b), orc._p(b))
        assert np.array_equal(a["p"], b["p"][:m]) and np.array_equal(a["q"], b["q"][:m])


def test_srt_partial_list(ctx):
    """A list of some of the pairs only (the sort's length is not NN of any order)."""
    rng = np.random.default_rng(11)
    n = 700
    dz = make_list(n, DISTS["wide"](rng, n * (n - 1) // 2), rng)[:123457]
    a = dz.copy()
    ctx.srt(a)
    e = dz[expected_order(dz, int(dz["q"].max()))]
    assert np.array_equal(a["p"], e["p"]) and np.array_equal(a["q"], e["q"])


def test_ncp3_plugin_rejects_a_pair_listed_twice(ctx):
    from oracle import orc
    from jkogb_amd import jkg
    dz = np.zeros(3, orc.AW_DTYPE)
    dz[0] = (3.0, 1, 2, 1, 1)
    dz[1] = (2.0, 3, 4, 1, 2)
    dz[2] = (1.0, 1, 2, 1, 3)
    with pytest.raises(jkg.JKGError):
        ctx.ncp(4, dz, 2, id_ncp=3)
