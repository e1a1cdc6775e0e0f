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
