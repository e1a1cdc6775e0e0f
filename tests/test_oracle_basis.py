"""Pins of the oracle's setup-time operators (dgMath.cpp / solver.cpp:74-192)
against closed forms, numpy's Gauss-Legendre rule and exactness on
polynomials."""
import numpy as np
import pytest

import oracle as O
from ldg_numpy import lagrange_ops, cgl_nodes


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_gauss_rule_matches_numpy(p):
    o = O.Oracle(p, 3)
    x, w = o.gauss()
    xr, wr = np.polynomial.legendre.leggauss(p + 1)
    np.testing.assert_allclose(x, xr, rtol=0, atol=2e-15)
    np.testing.assert_allclose(w, wr, rtol=0, atol=2e-15)


@pytest.mark.parametrize("p", [1, 2, 4])
def test_node_coordinates_are_cgl(p):
    """globalCoords: element vertices plus scaled CGL nodes (mesh.cpp:389-408)."""
    o = O.Oracle(p, 3)
    c = o.coords()
    r = cgl_nodes(p)
    n = p + 1
    h = 2 * np.pi / 3
    for k in [0, 5, 26]:
        ex, ey, ez = k % 3, (k // 3) % 3, k // 9
        for iN in range(n ** 3):
            ix, iy, iz = iN % n, (iN // n) % n, iN // (n * n)
            expect = [-np.pi + (e + 0.5 * (r[i] + 1)) * h for e, i in ((ex, ix), (ey, iy), (ez, iz))]
            np.testing.assert_allclose(c[k, iN], expect, atol=1e-14)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_interp_is_tensor_lagrange(p):
    """Interp3D = I (x) I (x) I and Interp2D = I (x) I in (slow..fast) =
    (z,y,x) order, where I is the CGL-to-Gauss Lagrange matrix."""
    o = O.Oracle(p, 3)
    _, _, _, I, _, _ = lagrange_ops(p)
    np.testing.assert_allclose(o.operator("Interp3D"), np.kron(I, np.kron(I, I)), atol=5e-14)
    np.testing.assert_allclose(o.operator("Interp2D"), np.kron(I, I), atol=5e-14)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_interp_exact_on_polynomials(p):
    """Interpolation from CGL nodes reproduces any degree <= p polynomial at
    the Gauss points (independent of the basis construction)."""
    o = O.Oracle(p, 3)
    r = cgl_nodes(p)
    x, _ = o.gauss()
    n = p + 1
    rng = np.random.default_rng(p)
    c = rng.standard_normal((n, n, n))
    f = lambda X, Y, Z: np.polynomial.polynomial.polyval3d(X, Y, Z, c)
    Z, Y, X = np.meshgrid(r, r, r, indexing="ij")
    Zq, Yq, Xq = np.meshgrid(x, x, x, indexing="ij")
    vals = f(X, Y, Z).ravel()
    np.testing.assert_allclose(o.operator("Interp3D") @ vals, f(Xq, Yq, Zq).ravel(), atol=1e-12 * np.abs(c).sum())


@pytest.mark.parametrize("p", [1, 2, 3, 4])
@pytest.mark.parametrize("nx", [3, 4])
def test_mass_stiffness_kron_structure(p, nx):
    """Mel = J M1(x)M1(x)M1 and Sels_l = (J/alpha_l) (... S1 in slot l ...)
    with M1 = I^T W I, S1 = D^T W I (solver.cpp:142-165)."""
    o = O.Oracle(p, nx, 3, 5)     # non-cubic: alpha differs per direction
    _, _, w, I, D, _ = lagrange_ops(p)
    W = np.diag(w)
    M1 = I.T @ W @ I
    S1 = D.T @ W @ I
    h = 2 * np.pi / np.array([nx, 3, 5])
    alpha = h / 2
    J = np.prod(alpha)
    np.testing.assert_allclose(o.operator("Mel"), J * np.kron(M1, np.kron(M1, M1)), atol=1e-14 * J * 50)
    S = o.operator("Sels")
    mats = [np.kron(M1, np.kron(M1, S1)), np.kron(M1, np.kron(S1, M1)), np.kron(S1, np.kron(M1, M1))]
    for l in range(3):
        np.testing.assert_allclose(S[l], J / alpha[l] * mats[l], atol=1e-13 * J / alpha[l] * 50)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_mass_matrix_integrates_exactly(p):
    """1^T Mel 1 = element volume; Mel is SPD."""
    o = O.Oracle(p, 4)
    M = o.operator("Mel")
    vol = (2 * np.pi / 4) ** 3
    assert abs(M.sum() - vol) < 1e-12 * vol
    assert np.all(np.linalg.eigvalsh(0.5 * (M + M.T)) > 0)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_kels_and_kels2d(p):
    """Kels_l(q,i) = d_l phi_i(q) w_q J/alpha_l and Kels2D_l = Interp2D w2 J/alpha_l."""
    o = O.Oracle(p, 3)
    _, _, w, I, D, _ = lagrange_ops(p)
    alpha = np.pi / 3
    J = alpha ** 3
    W3 = np.kron(w, np.kron(w, w))
    W2 = np.kron(w, w)
    K3 = o.operator("Kels")
    dphi = [np.kron(I, np.kron(I, D)), np.kron(I, np.kron(D, I)), np.kron(D, np.kron(I, I))]
    K2 = o.operator("Kels2D")
    for l in range(3):
        np.testing.assert_allclose(K3[l], dphi[l] * W3[:, None] * J / alpha, atol=1e-13)
        np.testing.assert_allclose(K2[l], np.kron(I, I) * W2[:, None] * J / alpha, atol=1e-13)


def test_physics_constants():
    """c0 = V0/M0, p0 = rho0 c0^2/gamma, T0 = p0/(R rho0) (solver.cpp:30-43)."""
    ph = O.Oracle(2, 3).physics()
    assert ph["c0"] == pytest.approx(10.0)
    assert ph["p0"] == pytest.approx(100 / 1.4)
    assert ph["T0"] == pytest.approx(100 / 1.4 / 8.314)


def test_reference_dt_formula():
    """dt = cfl min(dx)/(V0 (2p+1)) (solver.cpp:47)."""
    o = O.Oracle(2, 8)
    assert o.dt(0.01) == pytest.approx(0.01 * (2 * np.pi / 8) / 5, rel=1e-15)
