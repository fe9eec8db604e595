"""Key arithmetic of the octree kernels, checked on the host: the bit-spread
Morton encode/decode and the two-levels-per-read conversion from Morton keys
to curve keys against their per-level definitions (tests/cpp/sfc_bits_test.cpp,
compiled with hipcc; it makes no HIP calls)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "dendro4_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not shutil.which(HIPCC) and not os.path.exists(HIPCC), reason="hipcc not found")
def test_sfc_key_arithmetic(tmp_path):
    exe = tmp_path / "sfc_bits_test"
    cmd = [HIPCC, "-O2", "-std=c++17", "--offload-arch=gfx950", "-I", CSRC, "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "sfc_bits_test.cpp"), os.path.join(CSRC, "dgx_tables.cpp"),
           "-o", str(exe)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
