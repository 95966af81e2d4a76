"""CPU checks of the C ABI library: it loads, exports every entry point that
include/jkg.h declares, and its host-side (non-GPU) routines agree with the
oracle.  No compute call reaches the GPU here."""
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT


def header_functions():
    txt = open(os.path.join(ROOT, "include", "jkg.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(jkg_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from jkogb_amd import jkg
    out = subprocess.run(["nm", "-D", "--defined-only", jkg.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if " T " in l)
    declared = header_functions()
    assert declared, "no declarations parsed"
    missing = [f for f in declared if f not in exported]
    assert not missing, missing
    assert set(jkg.EXPORTS) <= set(declared)


def test_library_has_gfx950_code_object():
    from jkogb_amd import jkg
    data = open(jkg.LIB_PATH, "rb").read()
    assert b"gfx950" in data


@pytest.mark.parametrize("n", [3, 5, 16, 33])
def test_tru_and_jstep_len_match_oracle(orc, n):
    from jkogb_amd import jkg
    rng = np.random.default_rng(n)
    j = np.where(rng.random(n) < 0.5, 1, -1)
    for id_tru in (1, 2):
        a = jkg.tru(j, id_tru)
        b = orc.tru(j, id_tru)
        assert a[0] == b[0]
        assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])
    assert jkg.tru(np.array([1, 0, -1]))[0] == -2
    for n2 in range(-2, n):
        assert jkg.jstep_len(n, n2) == orc.jstep_len(n, n2)


def test_pair_index_is_the_aw_tie_break_order():
    """Ascending pair index == descending band B, then descending P."""
    from jkogb_amd import jkg
    for n in (3, 4, 9, 64):
        pairs = [(p, q) for q in range(2, n + 1) for p in range(1, q)]
        idx = [jkg.pair_index(n, p, q) for p, q in pairs]
        assert sorted(idx) == list(range(len(pairs)))
        order = sorted(pairs, key=lambda pq: (-(pq[1] - pq[0]), -pq[0]))
        assert [jkg.pair_index(n, p, q) for p, q in order] == list(range(len(pairs)))


def test_sort_key_roundtrip_and_order():
    from jkogb_amd import jkg
    w = np.array([np.inf, 1e300, 3.5, 1.0, 5e-324, 0.0, -0.0, -5e-324, -2.0, -1e300, -np.inf])
    # the device encoding restated on the host: ascending key <=> descending weight
    u = w.view(np.uint64)
    ordv = np.where(u >> np.uint64(63) == 1, ~u, u | np.uint64(1 << 63))
    keys = ~ordv
    keys[6] = keys[5]  # -0 is encoded as +0
    assert np.all(np.diff(keys.astype(object)) >= 0)
    back = jkg.keys_to_weights(keys)
    assert np.array_equal(back[[0, 1, 2, 3, 4, 7, 8, 9, 10]], w[[0, 1, 2, 3, 4, 7, 8, 9, 10]])
    assert jkg.key_weight(int(keys[2])) == 3.5
    assert np.isnan(jkg.key_weight(0xFFFFFFFFFFFFFFFF))
