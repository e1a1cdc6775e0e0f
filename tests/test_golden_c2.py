"""Consistency of the committed C2 fixture (tests/golden/c2_tgv32_p4_history.npz,
written by tests/golden/make_c2_history.py from the oracle) with the oracle
and with the physics, on the CPU: the time step follows the reference's rule
(solver.cpp:44-52), the recorded initial diagnostics are the oracle's
diagnostics of the Taylor-Green initial state, and the kinetic energy of the
decaying vortex falls monotonically after the first steps."""
import os

import numpy as np

import oracle as O

FIXTURE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c2_tgv32_p4_history.npz")


def test_c2_fixture_matches_oracle_initial_state_and_physics():
    f = np.load(FIXTURE)
    p, nx = int(f["p"]), int(f["nx"])
    assert (p, nx, int(f["nsteps"])) == (4, 32, 1000)
    o = O.Oracle(p, nx)
    assert float(f["dt"]) == o.dt(float(f["cfl"]))
    u0 = o.initial_condition()
    ke_g, ens = o.diagnostics(u0)
    assert f["ke_gauss"][0] == ke_g and f["enstrophy"][0] == ens
    assert np.array_equal(f["diag_steps"], np.arange(0, 1001, 10))
    ke = f["ke"]
    assert ke.shape == (1000,) and np.all(np.diff(ke[5:]) < 0)
    assert f["sample_state"].shape == (32, 5, 125)
