"""XYZV snapshot format (io.cpp:8-99) of the Python writer, on the CPU."""
import os

import numpy as np

from ldg_amd import io as lio
import ldg_amd


def test_header_and_rows(tmp_path):
    name = str(tmp_path / "xyzu.txt")
    assert lio.init_xyzv_file(name, 3, "u", 5)
    coords = np.array([[[0.1, 2.0, -3.14159265]]])          # (K=1, dofs=1, 3)
    u = np.array([[[1.0], [0.123456789], [-2e-7], [178.571428], [1e6]]])   # (K, 5, dofs)
    assert lio.export_to_xyzv_file(name, 3, coords, u)
    lines = open(name + ".3").read().splitlines()
    assert lines[0] == "X, Y, Z, u0, u1, u2, u3, u4"
    # C++ ostream default: 6 significant digits, %g style
    assert lines[1] == "0.1, 2, -3.14159, 1, 0.123457, -2e-07, 178.571, 1e+06"


def test_unwritable_directory_reports_failure(tmp_path):
    assert not lio.init_xyzv_file(str(tmp_path / "missing" / "x"), 0, "u", 5)


def test_node_coordinates_match_cgl_layout():
    mesh = ldg_amd.Mesh(3, 4, 5, ldg_amd.Point(-1, -1, -1), ldg_amd.Point(1, 2, 3))
    c = lio.node_coordinates(2, mesh)
    assert c.shape == (60, 27, 3)
    # element (1,2,3), node (ix,iy,iz) = (2,0,1): x at right end, y at left end, z mid
    k = 1 + 3 * (2 + 4 * 3)
    iN = 2 + 3 * (0 + 3 * 1)
    np.testing.assert_allclose(c[k, iN], [-1 + 2 * 2 / 3, -1 + 2 * 3 / 4, -1 + (3 + 0.5) * 4 / 5], atol=1e-14)
