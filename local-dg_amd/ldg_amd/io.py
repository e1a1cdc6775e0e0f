"""Paraview XYZV snapshot files in the reference's text format (io.cpp:8-99).

Header "X, Y, Z, u0, ..., u4", then one line per node "x, y, z, u0, ..., u4"
with every value printed like C++'s default ostream (6 significant digits,
%g style), nodes fastest then elements, and files named <name>.<index>.
"""
import math

import numpy as np


def _g(v):
    return format(float(v), ".6g")


def init_xyzv_file(filename, timeseries, valuename, nstates):
    """initXYZVFile (io.cpp:8-35): truncate and write the header."""
    try:
        with open(f"{filename}.{timeseries}", "w") as f:
            f.write("X, Y, Z, " + "".join(f"{valuename}{i}, " for i in range(nstates - 1))
                    + f"{valuename}{nstates - 1}\n")
        return True
    except OSError:
        return False


def export_to_xyzv_file(filename, timeseries, coords, u):
    """exportToXYZVFile (io.cpp:42-99): append every node of every element.
    coords is (K, dofs, 3) and u is (K, 5, dofs)."""
    try:
        with open(f"{filename}.{timeseries}", "a") as f:
            K, ns, dofs = u.shape
            for k in range(K):
                for i in range(dofs):
                    vals = [_g(c) for c in coords[k, i]] + [_g(u[k, s, i]) for s in range(ns)]
                    f.write(", ".join(vals) + "\n")
        return True
    except OSError:
        return False


def node_coordinates(p, mesh):
    """Node coordinates (K, dofs, 3): element vertices plus scaled
    Chebyshev-Lobatto nodes (mesh.cpp:127-146, 389-408)."""
    n = p + 1
    r = np.array([math.cos((p - i) * math.pi / p) for i in range(n)])
    lo = np.array([mesh.botLeft.x, mesh.botLeft.y, mesh.botLeft.z])
    hi = np.array([mesh.topRight.x, mesh.topRight.y, mesh.topRight.z])
    ns = np.array([mesh.nx, mesh.ny, mesh.nz])
    K = mesh.nx * mesh.ny * mesh.nz
    k = np.arange(K)
    ie = np.stack([k % mesh.nx, (k // mesh.nx) % mesh.ny, k // (mesh.nx * mesh.ny)], axis=1)
    i = np.arange(n ** 3)
    inode = np.stack([i % n, (i // n) % n, i // (n * n)], axis=1)
    a = ie * (hi - lo) / ns + lo
    b = (ie + 1) * (hi - lo) / ns + lo
    scale = 0.5 * (r[inode] + 1.0)
    return a[:, None, :] + scale[None, :, :] * (b - a)[:, None, :]


# ------------------------------------------------------------------ checkpoints
# Binary checkpoint of a run, the layout of local-dg_amd/host/io.hpp
# (writeCheckpoint / readCheckpoint), so that either side can resume the
# other's runs.  No reference counterpart: the reference only dumps an XYZV
# text file at its wall-clock limit (solver.cpp:349-360).
CKPT_MAGIC = b"LDGCKPT1"
_CKPT_HEAD = np.dtype([("magic", "S8"), ("ints", "<i4", 5), ("steps", "<i8", 2), ("reals", "<f8", 13)])


def save_checkpoint(filename, u, *, p, nx, ny, nz, precision, step, timesteps, dt, tf, L, Re, M0, V0, Pr,
                    lo, hi, ke):
    """Write the state u ((K, 5, dofs), i.e. the reference's (dofs, 5, K)
    column-major layout), the step reached and the KE history of steps
    1..step."""
    u = np.ascontiguousarray(u, dtype="<f8")
    ke = np.ascontiguousarray(ke, dtype="<f8")
    n = (p + 1) ** 3 * 5 * nx * ny * nz
    if u.size != n or not 0 <= step <= timesteps or ke.size > timesteps:
        raise ValueError("checkpoint sizes do not match the run")
    head = np.zeros((), dtype=_CKPT_HEAD)
    head["magic"] = CKPT_MAGIC
    head["ints"] = [p, nx, ny, nz, precision]
    head["steps"] = [step, timesteps]
    head["reals"] = [dt, tf, L, Re, M0, V0, Pr, *lo, *hi]
    with open(filename, "wb") as f:
        f.write(head.tobytes())
        f.write(np.int64(u.size).tobytes())
        f.write(u.tobytes())
        f.write(np.int64(ke.size).tobytes())
        f.write(ke.tobytes())


def load_checkpoint(filename):
    """Read a checkpoint written by save_checkpoint or the C++ Solver; raises
    ValueError on a malformed file."""
    with open(filename, "rb") as f:
        raw = f.read()
    if len(raw) < _CKPT_HEAD.itemsize + 16 or raw[:8] != CKPT_MAGIC:
        raise ValueError(f"{filename}: not a checkpoint")
    head = np.frombuffer(raw[:_CKPT_HEAD.itemsize], dtype=_CKPT_HEAD)[0]
    p, nx, ny, nz, precision = (int(v) for v in head["ints"])
    step, timesteps = (int(v) for v in head["steps"])
    r = [float(v) for v in head["reals"]]
    off = _CKPT_HEAD.itemsize
    n = int(np.frombuffer(raw[off:off + 8], dtype="<i8")[0])
    if n != (p + 1) ** 3 * 5 * nx * ny * nz or len(raw) < off + 8 + 8 * n + 8:
        raise ValueError(f"{filename}: state size does not match the header")
    u = np.frombuffer(raw[off + 8:off + 8 + 8 * n], dtype="<f8").reshape(nx * ny * nz, 5, (p + 1) ** 3).copy()
    off += 8 + 8 * n
    m = int(np.frombuffer(raw[off:off + 8], dtype="<i8")[0])
    if m < 0 or m > timesteps or len(raw) != off + 8 + 8 * m:
        raise ValueError(f"{filename}: KE history size does not match the header")
    ke = np.frombuffer(raw[off + 8:], dtype="<f8").copy()
    return {"p": p, "nx": nx, "ny": ny, "nz": nz, "precision": precision, "step": step, "timesteps": timesteps,
            "dt": r[0], "tf": r[1], "L": r[2], "Re": r[3], "M0": r[4], "V0": r[5], "Pr": r[6],
            "lo": r[7:10], "hi": r[10:13], "u": u, "ke": ke}
