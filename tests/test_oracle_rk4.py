"""Pins of the oracle's RK4 driver and diagnostics (solver.cpp:295-480,
1007-1034)."""
import numpy as np
import pytest

import oracle as O
from ldg_numpy import lagrange_ops


def test_rk4_is_fourth_order():
    """Halving dt divides the end-time error by ~16 (pins the tableau
    a = (1/2, 1/2, 1), b = (1/6, 1/3, 1/3, 1/6))."""
    o = O.Oracle(1, 3)
    u0 = o.initial_condition()
    T = 0.08
    ref, _ = o.step(u0, T / 256, 256, record_ke=False)
    errs = []
    for m in (8, 16):
        u, _ = o.step(u0, T / m, m, record_ke=False)
        errs.append(np.abs(u - ref).max())
    ratio = errs[0] / errs[1]
    assert 14.0 < ratio < 18.5, (errs, ratio)


def test_ke_uniform_flow_closed_form():
    """KE normalisation: a uniform state (rho, m) gives |m|^2/(2 rho) / rho0."""
    o = O.Oracle(2, 4)
    u = np.zeros(o.shape)
    u[:, 0], u[:, 1], u[:, 2], u[:, 3], u[:, 4] = 2.0, 1.0, -2.0, 0.5, 300.0
    assert o.ke(u) == pytest.approx((1 + 4 + 0.25) / (2 * 2.0), rel=1e-13)


@pytest.mark.parametrize("node", [0, 5, 13, 26])
def test_ke_uses_nodal_values_with_gauss_weights(node):
    """The reference KE weights the NODAL value u(iN) by the Gauss weight of
    the same index (solver.cpp:1021-1024): a single non-zero node gives
    J w_a w_b w_c |m|^2/(2 rho) / (rho0 (2 pi)^3)."""
    p, n = 2, 3
    o = O.Oracle(p, 3)
    _, _, w, _, _, _ = lagrange_ops(p)
    u = np.zeros(o.shape)
    u[:, 0] = 1.0
    u[4, 1, node] = 3.0
    ix, iy, iz = node % n, (node // n) % n, node // (n * n)
    J = (np.pi / 3) ** 3
    expect = J * w[ix] * w[iy] * w[iz] * 9.0 / 2.0 / (2 * np.pi) ** 3
    assert o.ke(u) == pytest.approx(expect, rel=1e-13)


def _tgv_integrals(nq=48):
    """High-order quadrature of the analytic Taylor-Green kinetic energy and
    enstrophy (rho from the isothermal pressure, m = rho0 V0 (...))."""
    x, w = np.polynomial.legendre.leggauss(nq)
    x = np.pi * x
    w = np.pi * w
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    W = np.einsum("i,j,k->ijk", w, w, w)
    p0 = 100 / 1.4
    pr = p0 + (np.cos(2 * X) + np.cos(2 * Y)) * (np.cos(2 * Z) + 2) / 16
    rho = pr / p0
    mx = np.sin(X) * np.cos(Y) * np.cos(Z)
    my = -np.cos(X) * np.sin(Y) * np.cos(Z)
    ke = np.sum(W * (mx ** 2 + my ** 2) / (2 * rho)) / (2 * np.pi) ** 3
    # vorticity of v = m / rho
    h = 1e-5
    def v(Xs, Ys, Zs):
        prr = p0 + (np.cos(2 * Xs) + np.cos(2 * Ys)) * (np.cos(2 * Zs) + 2) / 16
        r = prr / p0
        return (np.sin(Xs) * np.cos(Ys) * np.cos(Zs) / r, -np.cos(Xs) * np.sin(Ys) * np.cos(Zs) / r, 0 * Xs)
    d = lambda f, a: [(fp - fm) / (2 * h) for fp, fm in zip(f(*[c + (h if i == a else 0) for i, c in enumerate((X, Y, Z))]),
                                                         f(*[c - (h if i == a else 0) for i, c in enumerate((X, Y, Z))]))]
    dx, dy, dz = d(v, 0), d(v, 1), d(v, 2)
    wx = dy[2] - dz[1]
    wy = dz[0] - dx[2]
    wz = dx[1] - dy[0]
    ens = np.sum(W * rho * (wx ** 2 + wy ** 2 + wz ** 2) / 2) / (2 * np.pi) ** 3
    return ke, ens


def test_gauss_diagnostics_match_analytic_tgv():
    """Gauss-point KE and enstrophy of the interpolated t=0 state converge to
    the analytic integrals (0.125 and 0.375 up to O(M^2) density effects)."""
    ke_ex, ens_ex = _tgv_integrals()
    o = O.Oracle(4, 8)
    d = o.diagnostics(o.initial_condition())
    assert abs(d[0] - ke_ex) < 2e-6
    assert abs(d[1] - ens_ex) < 2e-3 * ens_ex
    assert abs(ke_ex - 0.125) < 1e-3 and abs(ens_ex - 0.375) < 3e-3


def test_cfl_0p1_diverges_cfl_0p01_is_stable():
    """The reference time-step rule at CFL 0.1 (10x the reference's 0.01) is
    beyond the RK4 stability limit of this scheme at Mach 0.1 (the acoustic
    speed is 11 V0): the 8^3, p=2 Taylor-Green run blows up within 100 steps,
    while the reference's own CFL 0.01 stays bounded."""
    o = O.Oracle(2, 8)
    u0 = o.initial_condition()
    u, ke = o.step(u0, o.dt(0.1), 100)
    assert not np.all(np.isfinite(u)) or np.nanmax(ke) > 1.0
    u, ke = o.step(u0, o.dt(0.01), 20)
    assert np.all(np.isfinite(u)) and abs(ke[-1] - 0.125) < 1e-3


def test_oracle_roundoff_floor_on_c1():
    """The oracle's own reproducibility: the same source built with fused
    multiply-adds differs from the plain build after the 100-step 8^3, p=2
    run by less than 1e-12 in rho, rho u, rho v, rho E but by more than that
    in rho w (which starts at zero), so a 1e-12 bar on rho w is below the
    reference algorithm's FP64 floor."""
    o = O.Oracle(2, 8)
    of = O.Oracle(2, 8, variant="fma")
    u0 = o.initial_condition()
    dt = o.dt(0.01)
    u1, _ = o.step(u0, dt, 100, record_ke=False)
    u2, _ = of.step(u0, dt, 100, record_ke=False)
    spread = np.sqrt(((u2 - u1) ** 2).sum(axis=(0, 2))) / np.sqrt((u1 ** 2).sum(axis=(0, 2)))
    assert np.all(spread[[0, 1, 2, 4]] < 1e-12), spread
    assert 1e-12 < spread[3] < 1e-10, spread
