"""Pins of the oracle's pointwise physics against independent textbook
formulations (solver.cpp:1040-1374)."""
import numpy as np
import pytest

import oracle as O
from ldg_numpy import euler_flux, viscous_flux, roe_flux_eig


def random_state(rng, mach=0.3):
    rho = rng.uniform(0.5, 2.0)
    v = rng.standard_normal(3) * mach * 3
    p = rng.uniform(20, 100)
    E = p / 0.4 + 0.5 * rho * v @ v
    return np.concatenate([[rho], rho * v, [E]])


def test_fluxC_textbook():
    rng = np.random.default_rng(0)
    for _ in range(50):
        u = random_state(rng)
        F = O.fluxC(u).reshape(3, 5)        # F[l, s] = fluxes[s + 5 l]
        np.testing.assert_allclose(F, euler_flux(u), rtol=1e-14, atol=1e-12)


def test_fluxV_textbook():
    rng = np.random.default_rng(1)
    for Re, Pr in [(1600.0, 0.71), (10.0, 0.9)]:
        for _ in range(50):
            u = random_state(rng)
            G = rng.standard_normal((3, 5))
            Fv = O.fluxV(u, G.ravel(), Re, Pr).reshape(3, 5)
            np.testing.assert_allclose(Fv, viscous_flux(u, G, Re, Pr), rtol=1e-12, atol=1e-14)


def test_fluxV_zero_for_uniform_state():
    rng = np.random.default_rng(2)
    u = random_state(rng)
    assert np.all(O.fluxV(u, np.zeros(15)) == 0.0)


@pytest.mark.parametrize("normal", [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, 0, -1), (0.6, 0.0, 0.8)])
def test_roe_matches_eigendecomposition(normal):
    rng = np.random.default_rng(3)
    n = np.array(normal, float)
    for _ in range(40):
        uN, uK = random_state(rng), random_state(rng)
        np.testing.assert_allclose(O.roe(uN, uK, n), roe_flux_eig(uN, uK, n), rtol=1e-11, atol=1e-11)


def test_roe_consistency():
    """F*(u, u, n) = F(u) . n"""
    rng = np.random.default_rng(4)
    for _ in range(20):
        u = random_state(rng)
        n = rng.standard_normal(3)
        n /= np.linalg.norm(n)
        np.testing.assert_allclose(O.roe(u, u, n), n @ euler_flux(u), rtol=1e-13, atol=1e-11)


def test_roe_supersonic_upwind():
    """All Roe wave speeds positive along n: F* = F(uK).n (flow leaves K);
    all negative: F* = F(uN).n.  Holds exactly because of the Roe property."""
    rng = np.random.default_rng(5)
    n = np.array([1.0, 0.0, 0.0])
    for _ in range(20):
        uK = random_state(rng, 0.0)
        uN = random_state(rng, 0.0)
        for u in (uK, uN):      # vx = +4 c-ish: supersonic in +x
            rho = u[0]
            p = 0.4 * (u[4] - 0.5 * (u[1:4] @ u[1:4]) / rho)
            c = np.sqrt(1.4 * p / rho)
            vx = 3.0 * c
            u[4] += 0.5 * rho * vx * vx
            u[1] = rho * vx
        np.testing.assert_allclose(O.roe(uN, uK, n), n @ euler_flux(uK), rtol=1e-11, atol=1e-9)
        np.testing.assert_allclose(O.roe(uN, uK, -n), -n @ euler_flux(uN), rtol=1e-11, atol=1e-9)


def test_roe_conservation_symmetry():
    """F*(uN, uK, n) = -F*(uK, uN, -n): both sides see equal and opposite fluxes."""
    rng = np.random.default_rng(6)
    for _ in range(20):
        uN, uK = random_state(rng), random_state(rng)
        n = np.array([0.0, 0.0, 1.0])
        np.testing.assert_allclose(O.roe(uN, uK, n), -O.roe(uK, uN, -n), rtol=1e-12, atol=1e-10)
