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


def run_env(exe, env, *args):
    e = dict(os.environ)
    e.update(env)
    return subprocess.run([exe] + [str(a) for a in args], capture_output=True, text=True, timeout=60, env=e)


@pytest.mark.parametrize("env,msg", [
    ({"JKG_CHECKPOINT": "0"}, "JKG_CHECKPOINT"), ({"JKG_CHECKPOINT": "x"}, "JKG_CHECKPOINT"),
    ({"JKG_RESUME": "2"}, "JKG_RESUME"), ({"JKG_FRAMES": "0"}, "JKG_FRAMES"), ({"JKG_FRAMES": "4,-1"}, "JKG_FRAMES"),
    ({"JKG_CHECKPOINT": "10", "JKG_SOLVER": "mixed"}, "plain solver only"),
    ({"JKG_RESUME": "1", "JKG_SOLVER": "warm"}, "plain solver only")])
def test_checkpoint_and_frame_switches(tmp_path, env, msg):
    for exe in (DJK, ZJK):
        p = run_env(exe, env, str(tmp_path / "none"), 8, 4)
        assert p.returncode != 0 and msg in p.stderr + p.stdout, (env, p.stderr)


def test_resume_needs_a_checkpoint(tmp_path):
    import numpy as np
    fn = tmp_path / "a"
    np.zeros(64).tofile(str(fn) + ".CA")  # FN.CU missing
    np.ones(8, dtype=np.int64).tofile(str(fn) + ".J")
    p = run_env(DJK, {"JKG_RESUME": "1"}, str(fn), 8, 4)
    assert p.returncode != 0 and "READ_CHECKPOINT" in p.stderr + p.stdout
    for name in ("CU", "CZ"):
        np.zeros(64).tofile(str(fn) + "." + name)
    np.array([-1], dtype=np.int64).tofile(str(fn) + ".CS")  # a checkpoint being written
    p = run_env(DJK, {"JKG_RESUME": "1"}, str(fn), 8, 4)
    assert p.returncode != 0 and "READ_CHECKPOINT" in p.stderr + p.stdout
