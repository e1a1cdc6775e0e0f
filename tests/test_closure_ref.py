"""Pins the numpy checker of the closure (tests/mlp_numpy.py) on the CPU:
the network gradient against central finite differences, the L2 projection
against what the mathematics fixes (exact on the coarse space, idempotent,
residual orthogonal to the coarse space, identity between equal orders),
and the BF16 rounding against hand-computed values."""
import numpy as np
import pytest

import mlp_numpy as M


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.0e-3], dtype=np.float32)
    r = M.bf16(x)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3 2^-8 is a tie -> even (1 + 2^-6)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -7 and r[3] == 1.0 + 2 ** -6
    assert r[4] == -2.5
    assert abs(r[5] - 3.0e-3) <= 3.0e-3 * 2 ** -8


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    nin, hid, nout, B = 7, 5, 4, 6
    npar = hid * nin + hid + hid * hid + hid + nout * hid + nout
    params = rng.standard_normal(npar) * 0.5
    X = rng.standard_normal((B, nin))
    T = rng.standard_normal((B, nout))
    g = M.gradient(params, X, T, nin, hid, nout)
    h = 1e-6
    fd = np.empty(npar)
    for i in range(npar):
        e = np.zeros(npar)
        e[i] = h
        fd[i] = (M.loss(params + e, X, T, nin, hid, nout) - M.loss(params - e, X, T, nin, hid, nout)) / (2 * h)
    assert np.max(np.abs(g - fd)) <= 1e-7 * max(1.0, np.max(np.abs(fd)))


def test_gradient_detects_a_planted_mistake():
    """A dropped SiLU derivative factor must fail the finite-difference check."""
    rng = np.random.default_rng(1)
    nin, hid, nout, B = 4, 3, 2, 5
    npar = hid * nin + hid + hid * hid + hid + nout * hid + nout
    params = rng.standard_normal(npar)
    X = rng.standard_normal((B, nin))
    T = rng.standard_normal((B, nout))
    orig = M.dsilu
    try:
        M.dsilu = lambda z: np.ones_like(z)
        g = M.gradient(params, X, T, nin, hid, nout)
    finally:
        M.dsilu = orig
    h = 1e-6
    fd = np.array([(M.loss(params + h * np.eye(npar)[i], X, T, nin, hid, nout)
                    - M.loss(params - h * np.eye(npar)[i], X, T, nin, hid, nout)) / (2 * h) for i in range(npar)])
    assert np.max(np.abs(g - fd)) > 1e-3


@pytest.mark.parametrize("pf,pc", [(5, 2), (4, 2), (3, 1), (5, 3)])
def test_projection_exact_on_coarse_space_and_idempotent(pf, pc):
    rng = np.random.default_rng(pf * 10 + pc)
    P = M.projection_1d(pf, pc)
    rf, rc = M.cgl_nodes(pf), M.cgl_nodes(pc)
    # a coarse polynomial sampled at the fine nodes projects to itself
    coef = rng.standard_normal(pc + 1)
    f_fine = np.polyval(coef, rf)
    f_coarse = np.polyval(coef, rc)
    assert np.allclose(P @ f_fine, f_coarse, atol=1e-13)
    # idempotent: embed the projection back (coarse interpolant at fine nodes) and project again
    g = rng.standard_normal(pf + 1)
    pg = P @ g
    emb = M.lagrange(rc, rf) @ pg
    assert np.allclose(P @ emb, pg, atol=1e-13)
    # residual orthogonal to the coarse space (Gauss rule exact for the products)
    x, w = np.polynomial.legendre.leggauss(pf + pc + 3)
    res = M.lagrange(rf, x) @ g - M.lagrange(rc, x) @ pg
    assert np.max(np.abs(M.lagrange(rc, x).T @ (w * res))) <= 1e-13


def test_projection_between_equal_orders_is_identity():
    for p in range(1, 6):
        assert np.allclose(M.projection_1d(p, p), np.eye(p + 1), atol=1e-13)


def test_projection_of_a_tensor_state_is_exact_on_q2():
    """A state whose fields are tensor-product quadratics is reproduced by
    the 3D projection at the coarse nodes."""
    rng = np.random.default_rng(3)
    pf, pc, K = 5, 2, 3
    rf, rc = M.cgl_nodes(pf), M.cgl_nodes(pc)

    def field(r, c):
        fx, fy, fz = (np.polyval(c[i], r) for i in range(3))
        return np.einsum("z,y,x->zyx", fz, fy, fx).ravel()

    u = np.empty((K, 5, (pf + 1) ** 3))
    ref = np.empty((K, 5, (pc + 1) ** 3))
    for e in range(K):
        for s in range(5):
            c = rng.standard_normal((3, 3))
            u[e, s] = field(rf, c)
            ref[e, s] = field(rc, c)
    assert np.allclose(M.project_state(u, pf, pc), ref, atol=1e-12)
