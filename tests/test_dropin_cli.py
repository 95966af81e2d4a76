"""Command-line checks of the drop-in programs that stop before any device work:
READCL's argument rules (src/shared/readcl.F90:1-85) and the reference's messages
(src/D/djk.F90:22-29, src/Z/zjk.F90)."""
import os
import subprocess

import pytest

from conftest import ROOT

DJK = os.path.join(ROOT, "host", "djk.exe")
ZJK = os.path.join(ROOT, "host", "zjk.exe")

pytestmark = pytest.mark.skipif(not (os.path.exists(DJK) and os.path.exists(ZJK)), reason="drop-ins not built")


def run(exe, *args):
    return subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("args", [(), ("a",), ("a", 4, 2, 0, 0, 9), ("a", -3), ("a", "x"), ("a", 4, -1),
                                  ("a", 4, 2, -1), ("a", 4, 2, 0, -2)])
def test_usage_message(args):
    for exe, name in ((DJK, "djk.exe"), (ZJK, "zjk.exe")):
        p = run(exe, *args)
        assert p.returncode != 0
        assert "%s FN N N_2 [ID_NCP [ID_TRU]]" % name in p.stderr + p.stdout


def test_small_n_and_bad_ids():
    p = run(DJK, "a", 2)
    assert p.returncode != 0 and "N should be at least 3" in p.stderr + p.stdout
    p = run(DJK, "a", 8, 4, 4)
    assert p.returncode != 0 and "DPROC_INIT" in p.stderr + p.stdout


def test_missing_input_files(tmp_path):
    p = run(DJK, str(tmp_path / "none"), 8, 4)
    assert p.returncode != 0 and "OPEN_YJ_RO" in p.stderr + p.stdout
