"""Binary checkpoint format: the Python reader/writer (ldg_amd.io) and the
C++ one (local-dg_amd/host/io.cpp) write and read the same bytes.  The GPU
resume test is in test_gpu_parity.py."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from ldg_amd import io

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HAVE_GXX = shutil.which("g++") is not None


def _sample(p=1, nx=3, ny=4, nz=5):
    n = (p + 1) ** 3
    u = (0.5 * np.arange(nx * ny * nz * 5 * n) - 1.0).reshape(nx * ny * nz, 5, n)
    return dict(p=p, nx=nx, ny=ny, nz=nz, precision=1, step=7, timesteps=11, dt=0.125, tf=1.375, L=1.0,
                Re=1600.0, M0=0.1, V0=1.0, Pr=0.71, lo=[-3.0, -4.0, -5.0], hi=[3.0, 4.0, 5.0],
                ke=1.0 / np.arange(1, 8)), u


def test_python_round_trip(tmp_path):
    meta, u = _sample()
    f = tmp_path / "a.ckpt"
    io.save_checkpoint(f, u, **meta)
    c = io.load_checkpoint(f)
    assert np.array_equal(c["u"], u) and np.array_equal(c["ke"], meta["ke"])
    for k in ("p", "nx", "ny", "nz", "precision", "step", "timesteps", "dt", "tf", "L", "Re", "M0", "V0", "Pr"):
        assert c[k] == meta[k], k
    assert c["lo"] == meta["lo"] and c["hi"] == meta["hi"]


def test_python_rejects_malformed(tmp_path):
    meta, u = _sample()
    f = tmp_path / "a.ckpt"
    io.save_checkpoint(f, u, **meta)
    raw = f.read_bytes()
    (tmp_path / "short").write_bytes(raw[:-8])
    (tmp_path / "magic").write_bytes(b"X" + raw[1:])
    for name in ("short", "magic"):
        with pytest.raises(ValueError):
            io.load_checkpoint(tmp_path / name)
    with pytest.raises(ValueError):
        io.save_checkpoint(tmp_path / "b", u[:-1], **meta)


@pytest.mark.skipif(not HAVE_GXX, reason="needs g++")
def test_cpp_and_python_agree(tmp_path):
    exe = tmp_path / "checkpoint_check"
    host = os.path.join(ROOT, "local-dg_amd", "host")
    # built with the address and undefined-behaviour sanitizers, so a read or
    # write past a buffer in the reader/writer aborts the check
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                    "-fno-omit-frame-pointer", "-I" + host, os.path.join(ROOT, "tests", "checkpoint_check.cpp"),
                    os.path.join(host, "io.cpp"), "-o", str(exe)], check=True, capture_output=True, text=True)
    meta, u = _sample()
    # C++ writes, Python reads
    subprocess.run([str(exe), "write", str(tmp_path / "c.ckpt")], check=True)
    c = io.load_checkpoint(tmp_path / "c.ckpt")
    assert np.array_equal(c["u"], u) and np.array_equal(c["ke"], meta["ke"])
    assert (c["p"], c["nx"], c["ny"], c["nz"], c["step"], c["timesteps"]) == (1, 3, 4, 5, 7, 11)
    # Python writes the same run: identical bytes
    io.save_checkpoint(tmp_path / "p.ckpt", u, **meta)
    assert (tmp_path / "p.ckpt").read_bytes() == (tmp_path / "c.ckpt").read_bytes()
    # and C++ reads what Python wrote
    r = subprocess.run([str(exe), "read", str(tmp_path / "p.ckpt")], check=True, capture_output=True, text=True)
    f = r.stdout.split()
    assert f[:7] == ["1", "3", "4", "5", "1", "7", "11"]
    assert float(f[7]) == 0.125 and float(f[14]) == -5.0 and float(f[15]) == 4.0
    assert int(f[16]) == u.size and float(f[17]) == float(np.sum(u))
    assert int(f[18]) == 7 and abs(float(f[19]) - float(np.sum(meta["ke"]))) < 1e-15
    # a truncated file is refused
    (tmp_path / "t.ckpt").write_bytes((tmp_path / "p.ckpt").read_bytes()[:-3])
    r = subprocess.run([str(exe), "read", str(tmp_path / "t.ckpt")], capture_output=True, text=True)
    assert r.returncode == 1
