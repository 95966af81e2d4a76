"""Host side of the test-matrix generator (jkogb_amd/jkdatgen.py): LAPACK's DLARAN
pinned by its closed form, GENLAM's redraw and scaling, JPART's column order."""
import math

import numpy as np


def test_dlaran_is_the_48_bit_congruential_generator():
    # x_{k+1} = a x_k mod 2^48 with the documented multiplier 33952834046453, output x / 2^48
    from jkogb_amd import jkdatgen as g
    m1, m2, m3, m4 = 494, 322, 2508, 2549
    a = ((m1 * 4096 + m2) * 4096 + m3) * 4096 + m4
    assert a == 33952834046453
    for seed in g.ISEEDS.values():
        iseed = list(seed)
        x = ((seed[0] * 4096 + seed[1]) * 4096 + seed[2]) * 4096 + seed[3]
        for _ in range(2000):
            v = g.dlaran(iseed)
            x = (a * x) % (1 << 48)
            assert v == x / float(1 << 48)
            assert iseed == [(x >> 36) & 4095, (x >> 24) & 4095, (x >> 12) & 4095, x & 4095]


def test_dlarnd_distributions():
    from jkogb_amd import jkdatgen as g
    s1, s2 = list(g.ISEEDS[1]), list(g.ISEEDS[1])
    for _ in range(100):
        t = g.dlaran(s2)
        assert g.dlarnd(2, s1) == 2.0 * t - 1.0
    s1, s2 = list(g.ISEEDS[2]), list(g.ISEEDS[2])
    for _ in range(100):
        t1, t2 = g.dlaran(s2), g.dlaran(s2)
        assert g.dlarnd(3, s1) == math.sqrt(-2.0 * math.log(t1)) * math.cos(2.0 * math.pi * t2)


def test_genlam_redraws_small_values_and_scales():
    from jkogb_amd import jkdatgen as g
    lam = g.genlam(500, 1, 2, 0.25, 1.0)
    assert lam.size == 500 and np.all(np.abs(lam) > 0.25)
    # the same stream, filtered by hand
    iseed, ref = list(g.ISEEDS[1]), []
    while len(ref) < 500:
        v = g.dlarnd(2, iseed)
        if abs(v) > 0.25:
            ref.append(v)
    assert np.array_equal(lam, np.array(ref))
    assert np.array_equal(g.genlam(500, 1, 2, 0.25, -3.0), -3.0 * lam)


def test_partition_puts_positive_columns_first():
    from jkogb_amd import jkdatgen as g
    lam = np.array([0.5, -1.0, 2.0, -0.25, 3.0])
    assert np.array_equal(g.partition(lam), [0.5, 2.0, 3.0, -1.0, -0.25])
