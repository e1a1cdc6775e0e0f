"""The workgroup -> element mapping of the LDG kernels, checked on the host:
tests/blockmap_check.cpp includes the kernel header, is built with hipcc for
the host only (no GPU needed) and run.  Multiply-shift division equals
integer division; block_elements covers every element exactly once with the
library's tilings; the XCD-aware numbering reaches every logical block once."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC) or shutil.which("g++") is None, reason="needs hipcc")
def test_block_mapping_covers_every_element_once(tmp_path):
    exe = tmp_path / "blockmap_check"
    cmd = [HIPCC, "-x", "hip", "--cuda-host-only", "-std=c++17", "-O2", "-Wno-unused-function",
           "-I" + os.path.join(ROOT, "local-dg_amd", "csrc"), "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "blockmap_check.cpp"), "-o", str(exe)]
    b = subprocess.run(cmd, capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-2000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("ok")
