"""The C++ host (local-dg_amd/host: Mesh, MPIUtil, Solver, dgTimeStep, io)
built with the address and undefined-behaviour sanitizers against a CPU
test double of libldg.so (tests/ldg_abi_stub.cpp), and driven the way the
reference's driver block does (driver.cpp:26-47): the reference's
Mesh{nx, nx, nx, botLeft, topRight, mpi} form, a run to tf, a wall-limit
stop with snapshots off followed by a continued run and a resume, and a
checkpoint refused across precisions (tests/host_solver_check.cpp)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "local-dg_amd", "host")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_solver_under_sanitizers(tmp_path):
    exe = tmp_path / "host_solver_check"
    srcs = [os.path.join(ROOT, "tests", "host_solver_check.cpp"), os.path.join(ROOT, "tests", "ldg_abi_stub.cpp")]
    srcs += [os.path.join(HOST, f) for f in ("mesh.cpp", "solver.cpp", "io.cpp")]
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer", "-I" + HOST, "-I" + os.path.join(ROOT, "include")] + srcs +
                   ["-o", str(exe)], check=True, capture_output=True, text=True)
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("ok")
