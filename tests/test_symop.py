"""Even/odd split of the centro-symmetric 1D operators (local-dg_amd/csrc/symop.hpp).

The header is plain C++ on the host, so tests/symop_check.cpp is built with
g++ and run: sym_apply / sym_apply2 against the plain matrix-vector product on
random (anti-)centro-symmetric matrices (N = 2..7) and on the element's own
operators from basis.hpp (p = 1..5), whose asymmetry must be round-off only.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"),
                    reason="needs g++ and the HIP headers")
def test_symop_split_matches_plain_product(tmp_path):
    exe = tmp_path / "symop_check"
    # host build with the address and undefined-behaviour sanitizers: an
    # index past an operator's split arrays aborts the check
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wno-unknown-pragmas", "-D__HIP_PLATFORM_AMD__",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
           "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "local-dg_amd", "csrc"),
           os.path.join(ROOT, "tests", "symop_check.cpp"), "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("ok")
