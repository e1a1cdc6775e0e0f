"""Size-independent properties at the full benchmark size: Taylor-Green on
64^3 elements, p = 4 (163,840,000 DOF), the workload the throughput figure is
quoted on.  The oracle cannot run at this size; what is checked instead:
- a periodic shift of the state by one element permutes the FP64 RHS bit for
  bit (every element's arithmetic depends only on its own and its
  neighbours' data, wherever it sits in the mesh and in the launch);
- mass is conserved over RK4 steps (FP64 to round-off, mixed to FP32
  round-off of the face fluxes);
- the mixed path stays within a few 1e-8 of FP64 after a few steps."""
import numpy as np
import pytest

import ldg_amd

pytestmark = pytest.mark.gpu

P, NX = 4, 64


def _mass_weights():
    from ldg_numpy import lagrange_ops
    _, _, w, I, _, _ = lagrange_ops(P)
    m1 = (I.T @ np.diag(w) @ I).sum(axis=0)        # 1^T M1
    return np.einsum("i,j,k->ijk", m1, m1, m1).ravel()


def test_periodic_shift_is_bitwise_at_64cubed_p4():
    print("64^3 p=4 FP64: RHS of a state shifted by one element along x", flush=True)
    g = ldg_amd.LDG(P, NX)
    g.initial_condition()
    u = g.get_state()
    rng = np.random.default_rng(64)
    u[:, 3] += 0.01 * rng.standard_normal(u[:, 3].shape)
    k = g.rhs(u)
    us = np.roll(u.reshape(NX, NX, NX, 5, -1), 1, axis=2).reshape(u.shape)
    del u
    ks = g.rhs(us)
    g.close()
    assert np.array_equal(ks, np.roll(k.reshape(NX, NX, NX, 5, -1), 1, axis=2).reshape(k.shape))


def test_mass_and_mixed_error_at_64cubed_p4():
    wts = _mass_weights()
    states = {}
    for name, prec, tol in (("fp64", ldg_amd.LDG_FP64, 1e-13), ("mixed", ldg_amd.LDG_MIXED, 1e-9)):
        print(f"64^3 p=4 {name}: 3 RK4 steps", flush=True)
        g = ldg_amd.LDG(P, NX, precision=prec)
        g.initial_condition()
        m0 = np.sum(g.get_state()[:, 0, :] @ wts)
        ke = g.step(g.time_step(0.01), 3, record_ke=True)
        u = g.get_state()
        g.close()
        m1 = np.sum(u[:, 0, :] @ wts)
        assert np.all(np.isfinite(ke)) and np.all(np.diff(ke) < 0), ke
        assert abs(m1 - m0) <= tol * m0, (name, m0, m1)
        states[name] = u
    a, b = states["mixed"], states["fp64"]
    rel = lambda x, y: np.linalg.norm(x - y) / np.linalg.norm(y)  # noqa: E731
    e_rho, e_m, e_E = rel(a[:, 0], b[:, 0]), rel(a[:, 1:4], b[:, 1:4]), rel(a[:, 4], b[:, 4])
    print("mixed vs FP64 after 3 steps: rho %.2e, momentum %.2e, rho E %.2e" % (e_rho, e_m, e_E), flush=True)
    assert e_rho <= 1e-7 and e_m <= 1e-6 and e_E <= 1e-7
