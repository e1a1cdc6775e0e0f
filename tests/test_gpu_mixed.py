"""Mixed-precision path (FP32 interpolation / flux / contractions, FP64
state, RK4, inverse mass and reductions) against the FP64 path and the
oracle.  Bar: relative L2 <= 1e-5 per conserved variable of the state
(the bar stated for the 64^3, p=4 run at t=1; here on smaller meshes and
horizons), and single right-hand sides within FP32 round-off amplified by
the inverse mass: the residual carries FP32 rounding (6e-8) and
M^{-1} = M1^{-1} (x) M1^{-1} (x) M1^{-1} has condition number ~1e3 at
p = 3..4, so a single k is only good to ~1e-4 relative."""
import numpy as np
import pytest

import ldg_amd
import oracle as O

pytestmark = pytest.mark.gpu

STATE_TOL = 1e-5


def rel_l2(a, b):
    return np.sqrt(np.sum((a - b) ** 2, axis=(0, 2))) / np.sqrt(np.sum(b ** 2, axis=(0, 2)))


@pytest.mark.parametrize("p,nx", [(2, 4), (4, 4), (3, 5)])
def test_mixed_rhs_close_to_fp64(p, nx):
    o = O.Oracle(p, nx)
    u = o.initial_condition()
    rng = np.random.default_rng(p)
    u[:, 3] += 0.02 * rng.standard_normal(u[:, 3].shape)     # non-zero rho w
    k64 = ldg_amd.LDG(p, nx).rhs(u)
    k32 = ldg_amd.LDG(p, nx, precision=ldg_amd.LDG_MIXED).rhs(u)
    k_ref = o.rhs(u)
    e = rel_l2(k32, k64)
    eo = rel_l2(k32, k_ref)
    print(f"mixed rhs p={p} {nx}^3: rel L2 vs FP64 path {e}; vs oracle {eo}; FP64 path vs oracle {rel_l2(k64, k_ref)}",
          flush=True)
    assert np.all(e <= 5e-4), e
    assert np.all(eo <= 5e-4), eo


@pytest.mark.parametrize("p,nx,ny,nz", [(4, 5, 3, 4), (3, 3, 4, 5), (2, 6, 4, 3)])
def test_mixed_rhs_on_non_cubic_meshes(p, nx, ny, nz):
    """Non-cubic meshes (a different 2/dx per direction in the FP32
    gradient, flux and lift scalings), perturbed state with non-zero rho w:
    the mixed right-hand side against the oracle and the FP64 path (bar as
    for the cubic cases), and 30 mixed steps against 30 oracle steps."""
    o = O.Oracle(p, nx, ny, nz)
    u = o.initial_condition()
    rng = np.random.default_rng(10 * p + nx)
    u[:, 3] += 0.02 * rng.standard_normal(u[:, 3].shape)
    g64 = ldg_amd.LDG(p, nx, ny, nz)
    g32 = ldg_amd.LDG(p, nx, ny, nz, precision=ldg_amd.LDG_MIXED)
    k64, k32, k_ref = g64.rhs(u), g32.rhs(u), o.rhs(u)
    e, eo = rel_l2(k32, k64), rel_l2(k32, k_ref)
    dt = o.dt(0.01)
    g32.set_state(u)
    g32.step(dt, 30, record_ke=False)
    u_ref, _ = o.step(u, dt, 30)
    es = rel_l2(g32.get_state(), u_ref)
    print(f"mixed p={p} {nx}x{ny}x{nz}: rhs vs FP64 {e}, vs oracle {eo}; state after 30 steps vs oracle {es}",
          flush=True)
    assert np.all(e <= 5e-4), e
    assert np.all(eo <= 5e-4), eo
    assert np.all(es <= STATE_TOL), es


def test_mixed_state_error_after_200_steps():
    """Taylor-Green 8^3, p=4, 200 steps at the reference time step from the
    oracle's initial state: the mixed state stays within 1e-5 relative L2 of
    the CPU oracle's state and of the FP64 GPU state, for every conserved
    variable; KE histories agree with both to 1e-7."""
    p, nx, nsteps = 4, 8, 200
    o = O.Oracle(p, nx)
    u0 = o.initial_condition()
    dt = o.dt(0.01)
    g64 = ldg_amd.LDG(p, nx)
    g32 = ldg_amd.LDG(p, nx, precision=ldg_amd.LDG_MIXED)
    for g in (g64, g32):
        g.set_state(u0)
    ke64 = g64.step(dt, nsteps)
    ke32 = g32.step(dt, nsteps)
    u64, u32 = g64.get_state(), g32.get_state()
    u_ref, ke_ref = o.step(u0, dt, nsteps)
    e, eo = rel_l2(u32, u64), rel_l2(u32, u_ref)
    print("mixed after 200 steps, per variable: vs FP64 path", e, "vs oracle", eo, flush=True)
    assert np.all(e <= STATE_TOL), e
    assert np.all(eo <= STATE_TOL), eo
    assert np.max(np.abs(ke32 - ke64) / ke64) <= 1e-7
    assert np.max(np.abs(ke32 - ke_ref) / ke_ref) <= 1e-7


def test_mixed_diagnostics_and_conservation():
    g = ldg_amd.LDG(4, 16, precision=ldg_amd.LDG_MIXED)
    g.initial_condition()
    d0 = g.diagnostics()
    from ldg_numpy import lagrange_ops
    _, _, w, I, _, _ = lagrange_ops(4)
    m1 = (I.T @ np.diag(w) @ I).sum(axis=0)
    wts = np.einsum("i,j,k->ijk", m1, m1, m1).ravel()
    mass0 = np.sum(g.get_state()[:, 0, :] @ wts)
    g.step(g.time_step(0.01), 10, record_ke=False)
    mass1 = np.sum(g.get_state()[:, 0, :] @ wts)
    # FP32 fluxes: mass is conserved to FP32 round-off of the face fluxes
    assert abs(mass1 - mass0) <= 1e-9 * mass0
    assert abs(d0["ke"] - 0.1248) < 2e-3 and abs(d0["enstrophy"] - 0.3745) < 5e-3


def test_mixed_steps_do_not_depend_on_call_split():
    """Steps after the first in one call read their stage-1 input as the FP32
    copy the previous step left; the first step of a call rounds the FP64
    state to the same values.  Five steps in one call, one per call and by
    graph replay (plus an eager remainder) end bitwise equal."""
    print("mixed call-split invariance p=4 6^3", flush=True)
    runs = []
    for mode in ("one call", "single steps", "graph"):
        g = ldg_amd.LDG(4, 6, precision=ldg_amd.LDG_MIXED)
        g.initial_condition()
        dt = g.time_step(0.01)
        if mode == "graph":
            g.set_graph(True)
        if mode == "single steps":
            ke = np.concatenate([g.step(dt, 1, record_ke=True) for _ in range(5)])
        else:
            ke = g.step(dt, 5, record_ke=True)
        runs.append((g.get_state(), ke))
        g.close()
    for u, ke in runs[1:]:
        assert np.array_equal(u, runs[0][0])
        assert np.array_equal(ke, runs[0][1])
