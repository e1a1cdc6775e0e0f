"""Ensembles on the GPU (ldg_ensemble_create / ldg_ensemble_step).

An ensemble advances independent instances of mixed orders, meshes and
physics in shared kernel launches.  Each instance must behave exactly as a
solver of its own: its state and kinetic-energy history after several steps
are bit-identical to a single-instance handle started from the same state
(the element work, the per-block partial sums and their reduction order do
not depend on where the instance sits in the launch).  The per-instance
physics are checked against the oracle, the phase-shifted initial state
against a periodic roll of the unshifted one, and graph replay against eager
launches."""
import math

import numpy as np
import pytest

import ldg_amd
import oracle as O

pytestmark = pytest.mark.gpu

MEMBERS = [
    {"p": 2, "nx": 6, "Re": 800.0, "M0": 0.2, "V0": 1.03, "phase": (0.3, 1.1, 2.0)},
    {"p": 3, "nx": 4, "ny": 5, "nz": 3, "Re": 2500.0, "M0": 0.12, "V0": 0.97, "phase": (0.0, 0.5, 0.0)},
    {"p": 2, "nx": 5, "Re": 400.0, "M0": 0.3, "V0": 1.0, "phase": (2.5, 0.0, 4.0)},
    {"p": 4, "nx": 3, "Re": 1600.0, "M0": 0.1, "V0": 1.0},
    {"p": 3, "nx": 4, "Re": 1000.0, "M0": 0.08, "V0": 1.05, "phase": (1.0, 1.0, 1.0)},
]


def solo(d, precision):
    return ldg_amd.LDG(d["p"], d["nx"], d.get("ny"), d.get("nz"), Re=d["Re"], M0=d["M0"], V0=d["V0"],
                       precision=precision)


@pytest.mark.parametrize("precision", [ldg_amd.LDG_FP64, ldg_amd.LDG_MIXED])
def test_instances_match_single_solvers_bitwise(precision):
    print(f"ensemble vs single solvers, precision={precision}", flush=True)
    ens = ldg_amd.Ensemble(MEMBERS, precision=precision)
    ens.initial_condition()
    dts = ens.time_steps(0.01)
    u0 = [ens.get_instance_state(i) for i in range(ens.count)]
    nsteps = 5
    ke = ens.step(dts, nsteps)
    for i, d in enumerate(MEMBERS):
        g = solo(d, precision)
        assert g.time_step(0.01) == dts[i]
        g.set_state(u0[i])
        ke_i = g.step(dts[i], nsteps)
        np.testing.assert_array_equal(ens.get_instance_state(i), g.get_state(), err_msg=f"instance {i}")
        np.testing.assert_array_equal(ke[:, i], ke_i, err_msg=f"instance {i}")
        di, ds = ens.instance_diagnostics(i), g.diagnostics()
        assert di == ds, (i, di, ds)
        g.close()
    ens.close()


def test_instance_physics_match_oracle():
    """Each instance with its own Re, M0, V0 (and order/mesh) after 10 steps
    against the oracle built with those constants, to the bar of the
    single-solver parity tests: relative L2 1e-12 per conserved variable, or
    4x the oracle's own FMA rounding spread where that is larger."""
    ens = ldg_amd.Ensemble(MEMBERS)
    ens.initial_condition()
    dts = ens.time_steps(0.01)
    u0 = [ens.get_instance_state(i) for i in range(ens.count)]
    ens.step(dts, 10, record_ke=False)
    for i, d in enumerate(MEMBERS):
        phys = [d["Re"], 0.71, 1.4, d["M0"], 1.0, d["V0"], 8.314]
        args = (d["p"], d["nx"], d.get("ny"), d.get("nz"))
        o = O.Oracle(*args, phys=phys)
        assert o.dt(0.01) == dts[i]
        ref, _ = o.step(u0[i], dts[i], 10, record_ke=False)
        fma, _ = O.Oracle(*args, phys=phys, variant="fma").step(u0[i], dts[i], 10, record_ke=False)
        u = ens.get_instance_state(i)
        num = np.sqrt(np.sum((u - ref) ** 2, axis=(0, 2)))
        den = np.sqrt(np.sum(ref ** 2, axis=(0, 2)))
        spread = np.sqrt(np.sum((fma - ref) ** 2, axis=(0, 2)))
        scale = np.maximum(den, 1e-300)
        e = num / scale
        tol = np.maximum(1e-12, 4 * spread / scale)
        print(f"instance {i}: rel L2 {e}, tol {tol}", flush=True)
        assert np.all(e <= tol), (i, e, tol)


def test_phase_shift_is_a_periodic_roll():
    """A phase of exactly three element widths in x equals the unshifted
    Taylor-Green state rolled by three elements (up to the rounding of the
    node coordinates)."""
    n = 8
    dx = 2 * math.pi / n
    ens = ldg_amd.Ensemble([{"p": 3, "nx": n, "Re": 1600.0, "M0": 0.1, "V0": 1.0},
                            {"p": 3, "nx": n, "Re": 1600.0, "M0": 0.1, "V0": 1.0, "phase": (3 * dx, 0.0, 0.0)}])
    ens.initial_condition()
    a = ens.get_instance_state(0).reshape(n, n, n, 5, -1)       # (ez, ey, ex, s, node)
    b = ens.get_instance_state(1).reshape(n, n, n, 5, -1)
    rolled = np.roll(a, -3, axis=2)
    assert np.abs(b - rolled).max() <= 1e-13 * np.abs(a).max()


@pytest.mark.parametrize("precision", [ldg_amd.LDG_FP64, ldg_amd.LDG_MIXED])
def test_graph_replay_matches_eager(precision):
    """Steps replayed from HIP graphs equal eager launches bit for bit:
    odd step counts (an eager remainder that flips the array roles), a
    change of dt between calls, diagnostics between replays, with and without
    the kinetic-energy record.  In the mixed path the first replay of a call
    reads the FP64 state and later ones the FP32 copy the previous replay
    left; the plan needs more graph variants than the cache holds, so
    variants are evicted and rebuilt."""
    print("graph vs eager, precision", precision, flush=True)
    eager = ldg_amd.Ensemble(MEMBERS[:3], precision=precision)
    graph = ldg_amd.Ensemble(MEMBERS[:3], precision=precision)
    graph.set_graph(True)
    for e in (eager, graph):
        e.initial_condition()
    dts = eager.time_steps(0.01)
    plan = [(dts, 7, True), (dts, 4, True), (dts * 0.5, 5, True), (dts * 0.5, 6, False), (dts, 3, True),
            (dts, 2, False), (dts, 9, False), (dts, 8, True), (dts, 1, True), (dts, 6, True)]
    for call, (dt, n, rec) in enumerate(plan):
        ke_a = eager.step(dt, n, record_ke=rec)
        ke_b = graph.step(dt, n, record_ke=rec)
        if rec:
            np.testing.assert_array_equal(ke_a, ke_b, err_msg=f"call {call}")
        for i in range(eager.count):
            np.testing.assert_array_equal(eager.get_instance_state(i), graph.get_instance_state(i),
                                          err_msg=f"call {call} instance {i}")
        assert graph.instance_diagnostics(1) == eager.instance_diagnostics(1)


def test_autotune_keeps_results():
    """The autotuner picks one of the candidate tilings per order; tuning
    leaves the state untouched, and the states after further steps are
    bit-identical to an untuned run (the element work does not depend on the
    tiling).  The kinetic energy sums per-workgroup partials, so a different
    tiling changes its summation order: equal to 1e-15 relative."""
    ens = ldg_amd.Ensemble(MEMBERS)
    ref = ldg_amd.Ensemble(MEMBERS)
    for e in (ens, ref):
        e.initial_condition()
    before = [ens.get_instance_state(i) for i in range(ens.count)]
    chosen = ens.autotune()
    print("autotune:", chosen, flush=True)
    assert sorted(c[0] for c in chosen) == sorted(set(d["p"] for d in MEMBERS))
    for i in range(ens.count):
        np.testing.assert_array_equal(ens.get_instance_state(i), before[i])
    dts = ens.time_steps(0.01)
    ke_a = ens.step(dts, 3)
    ke_b = ref.step(dts, 3)
    for i in range(ens.count):
        np.testing.assert_array_equal(ens.get_instance_state(i), ref.get_instance_state(i))
    assert np.max(np.abs(ke_a - ke_b) / ke_b) <= 1e-15


def test_sampled_ensemble_runs():
    """48 instances drawn by ldg_ensemble_sample (deterministic in the seed),
    advanced with graph replay in mixed precision: finite kinetic energies,
    and three of the instances bit-identical to single solvers run eagerly
    from the same state."""
    inst = ldg_amd.sample_instances(7, 48)
    assert inst == ldg_amd.sample_instances(7, 48)
    ens = ldg_amd.Ensemble(inst, precision=ldg_amd.LDG_MIXED)
    ens.set_graph(True)
    ens.initial_condition()
    picks = [0, 17, 47]
    u0 = {i: ens.get_instance_state(i) for i in picks}
    dts = ens.time_steps(0.01)
    ke = ens.step(dts, 10)
    assert np.all(np.isfinite(ke))
    for i in picks:
        g = solo(inst[i], ldg_amd.LDG_MIXED)
        g.set_state(u0[i])
        ke_i = g.step(dts[i], 10)
        np.testing.assert_array_equal(ens.get_instance_state(i), g.get_state(), err_msg=f"instance {i}")
        np.testing.assert_array_equal(ke[:, i], ke_i, err_msg=f"instance {i}")
        g.close()
