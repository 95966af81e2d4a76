"""CPU checks of how the gfx950 library is built: every object's device code
is free of packed FP32 arithmetic (the build passes
-Xclang -target-feature -Xclang -packed-fp32-ops), and a source that does not
compile fails the build instead of leaving an old object behind."""
import glob
import os
import re
import shutil
import subprocess

import pytest

from conftest import ROOT

LLVM = "/opt/rocm/llvm/bin"
CSRC = os.path.join(ROOT, "jkogb_amd", "csrc")


def device_disassembly(obj, tmp):
    fat = os.path.join(tmp, "x.fatbin")
    co = os.path.join(tmp, "x.co")
    subprocess.run([os.path.join(LLVM, "llvm-objcopy"), "--dump-section", ".hip_fatbin=" + fat, obj, os.devnull],
                   check=True, capture_output=True)
    subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--type=o", "--input=" + fat,
                    "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", "--output=" + co, "--unbundle"],
                   check=True, capture_output=True)
    return subprocess.run([os.path.join(LLVM, "llvm-objdump"), "-d", co], check=True, capture_output=True,
                          text=True).stdout


@pytest.mark.skipif(not os.path.exists(os.path.join(LLVM, "clang-offload-bundler")), reason="no ROCm LLVM tools")
def test_no_packed_fp32_arithmetic_in_device_code(tmp_path):
    objs = sorted(glob.glob(os.path.join(CSRC, "_build", "*.o")))
    if not objs:
        pytest.skip("library objects not built here")
    packed = re.compile(r"\bv_pk_(add|mul|fma)_f32\b")
    seen = 0
    for obj in objs:
        asm = device_disassembly(obj, str(tmp_path))
        assert "s_endpgm" in asm, obj  # the device code was found
        hits = packed.findall(asm)
        assert not hits, (os.path.basename(obj), len(hits))
        seen += 1
    assert seen == len(glob.glob(os.path.join(CSRC, "*.hip")))


def test_failed_compile_fails_the_build(tmp_path):
    """The library's Makefile on a tree with one source that does not compile: make fails and
    leaves no object for it, even where an object of an earlier good build was there."""
    mk = tmp_path / "a" / "b" / "csrc"
    mk.mkdir(parents=True)
    # the Makefile's header dependency is ../../include/jkg.h from its own directory
    (tmp_path / "a" / "include").mkdir()
    (tmp_path / "a" / "include" / "jkg.h").write_text("")
    shutil.copy(os.path.join(CSRC, "Makefile"), mk / "Makefile")
    (mk / "bad.hip").write_text("__global__ void k() { int x = ; }\n")
    (mk / "_build").mkdir()
    stale = mk / "_build" / "bad.o"
    stale.write_bytes(b"old object")
    os.utime(stale, (1, 1))  # older than the source: make rebuilds it
    r = subprocess.run(["make", "-s", "-C", str(mk)], capture_output=True, text=True)
    assert r.returncode != 0
    assert not stale.exists()
    assert not (tmp_path / "a" / "b" / "libjkg.so").exists()
