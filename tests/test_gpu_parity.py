"""GPU parity of libldg.so (through its C ABI) against the CPU oracle.

Two oracles built from the same source (oracle/ldg_oracle.c): the FP64
build, which is the reference's algorithm in the reference's precision, and
the x87 long double build ("ld"), which carries every operation 11 bits
further and is the exact answer to better than 1e-15 here.  The bar of
1e-12 relative L2 per conserved variable is applied against the exact
answer.  The FP64 oracle is itself up to 2.6e-12 (rho w after 100 C1 steps)
and 4e-10 (a single p = 4 right-hand side) away from it: the O(1/M0^2)
background pressure cancels in its dense operators and fluxes.  The GPU
path evaluates fluxes on fluctuations about the background and is tested to
be at least as close to the exact answer as the FP64 oracle is.  Bitwise
where the computation is identical by construction (periodic shifts)."""
import numpy as np
import pytest

import ldg_amd
import oracle as O

pytestmark = pytest.mark.gpu

TOL = 1e-12
FLOOR_FACTOR = 4.0


def rel_l2(a, b, axis=(0, 2)):
    return np.sqrt(np.sum((a - b) ** 2, axis=axis)) / np.sqrt(np.sum(b ** 2, axis=axis))


def perturbed(o, seed, amp=0.05):
    rng = np.random.default_rng(seed)
    u = o.initial_condition()
    pert = rng.standard_normal(u.shape) * amp
    pert[:, 4] *= 20.0
    u = u + pert
    u[:, 0] = np.abs(u[:, 0]) + 0.5
    return u


@pytest.mark.parametrize("p,nx", [(1, 3), (2, 4), (4, 4)])
def test_device_initial_condition_matches_oracle(p, nx):
    o = O.Oracle(p, nx)
    g = ldg_amd.LDG(p, nx)
    g.initial_condition()
    u = g.get_state()
    ref = o.initial_condition()
    assert np.abs(u - ref).max() <= 1e-13 * np.abs(ref).max()


CASES = [(1, 3, 3, 3), (2, 4, 4, 4), (3, 3, 4, 5), (4, 4, 4, 4), (5, 3, 3, 3), (2, 8, 8, 8), (4, 5, 3, 4)]


@pytest.mark.parametrize("p,nx,ny,nz", CASES)
@pytest.mark.parametrize("state", ["tgv", "perturbed"])
def test_rhs_matches_oracle(p, nx, ny, nz, state):
    """One right-hand side and its auxiliary gradient.  Against the exact
    (long double) answer: k within max(1e-12, the FP64 oracle's own error);
    against the FP64 oracle: within max(1e-12, 4 x the spread between its
    plain and fused-multiply-add builds)."""
    print(f"rhs parity p={p} mesh={nx}x{ny}x{nz} state={state}", flush=True)
    o = O.Oracle(p, nx, ny, nz)
    of = O.Oracle(p, nx, ny, nz, variant="fma")
    ol = O.Oracle(p, nx, ny, nz, variant="ld")
    g = ldg_amd.LDG(p, nx, ny, nz)
    u = o.initial_condition() if state == "tgv" else perturbed(o, 100 + p)
    k_ref, du_ref = o.rhs(u, want_du=True)
    k_fma, du_fma = of.rhs(u, want_du=True)
    k_ex, du_ex = ol.rhs(u, want_du=True)
    k, du = g.rhs(u, want_du=True)
    e_gpu, e_orc = rel_l2(k, k_ex), rel_l2(k_ref, k_ex)
    print("  k rel L2 vs exact: GPU", e_gpu, "FP64 oracle", e_orc, flush=True)
    assert np.all(e_gpu <= np.maximum(TOL, e_orc)), (e_gpu, e_orc)
    ek = rel_l2(k, k_ref)
    tol = np.maximum(TOL, FLOOR_FACTOR * rel_l2(k_fma, k_ref))
    assert np.all(ek <= tol), (ek, tol)
    for l in range(3):
        # rho w and its gradients vanish identically for the TGV state
        nz_mask = np.sqrt(np.sum(du_ex[l] ** 2, axis=(0, 2))) > 1e-10
        e = rel_l2(du[l], du_ex[l])[nz_mask]
        eo = rel_l2(du_ref[l], du_ex[l])[nz_mask]
        assert np.all(e <= np.maximum(TOL, eo)), (l, e, eo)
        e = rel_l2(du[l], du_ref[l])[nz_mask]
        tl = np.maximum(TOL, FLOOR_FACTOR * rel_l2(du_fma[l], du_ref[l]))[nz_mask]
        assert np.all(e <= tl), (l, e, tl)


def test_c1_hundred_steps_match_exact_answer():
    """Taylor-Green, 8^3 elements, p=2, 100 RK4 steps at the reference time
    step (CFL 0.01): every conserved variable within relative L2 1e-12 of the
    exact (long double) answer, rho w included, and the kinetic-energy
    history within 1e-12 relative.  The FP64 oracle's own distance from the
    exact answer is printed beside the GPU's (rho w: 2.6e-12), and the GPU
    must also agree with the FP64 oracle to within 4 x that oracle's own
    rounding spread."""
    p, nx, nsteps = 2, 8, 100
    o = O.Oracle(p, nx)
    g = ldg_amd.LDG(p, nx)
    dt = o.dt(0.01)
    assert g.time_step(0.01) == dt
    u0 = o.initial_condition()
    g.set_state(u0)
    ke = g.step(dt, nsteps)
    u = g.get_state()
    u_ref, ke_ref = o.step(u0, dt, nsteps)
    u_ex, ke_ex = O.Oracle(p, nx, variant="ld").step(u0, dt, nsteps)
    u_fma, _ = O.Oracle(p, nx, variant="fma").step(u0, dt, nsteps, record_ke=False)
    e_gpu, e_orc = rel_l2(u, u_ex), rel_l2(u_ref, u_ex)
    print("C1 per-variable relative L2 after 100 steps vs exact: GPU", e_gpu, "FP64 oracle", e_orc, flush=True)
    assert np.all(e_gpu <= TOL), e_gpu
    assert np.max(np.abs(ke - ke_ex) / np.abs(ke_ex)) <= TOL
    e = rel_l2(u, u_ref)
    tol = np.maximum(TOL, FLOOR_FACTOR * rel_l2(u_fma, u_ref))
    assert np.all(e <= tol), (e, tol)
    assert np.max(np.abs(ke - ke_ref) / np.abs(ke_ref)) <= TOL


def test_p4_kinetic_energy_and_enstrophy_history():
    """p=4 on 8^3: reference KE every step and Gauss KE / enstrophy every 10
    steps, against the oracle."""
    p, nx, nsteps = 4, 8, 30
    o = O.Oracle(p, nx)
    g = ldg_amd.LDG(p, nx)
    dt = o.dt(0.01)
    u = o.initial_condition()
    u_fma, _ = O.Oracle(p, nx, variant="fma").step(u, dt, nsteps, record_ke=False)
    g.set_state(u)
    for block in range(3):
        d_ref = o.diagnostics(u)
        d = g.diagnostics()
        assert abs(d["ke_ref"] - o.ke(u)) <= TOL * abs(o.ke(u))
        assert abs(d["ke"] - d_ref[0]) <= TOL * d_ref[0]
        assert abs(d["enstrophy"] - d_ref[1]) <= 1e-10 * d_ref[1]
        ke = g.step(dt, 10)
        u, ke_ref = o.step(u, dt, 10)
        assert np.max(np.abs(ke - ke_ref) / ke_ref) <= TOL
    e = rel_l2(g.get_state(), u)
    tol = np.maximum(TOL, FLOOR_FACTOR * rel_l2(u_fma, u))
    print("p=4 8^3 per-variable relative L2 after 30 steps:", e, "tol", tol, flush=True)
    assert np.all(e <= tol), (e, tol)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_periodic_shift_is_bitwise_at_32cubed_p4(axis):
    """Shifting the state by one element permutes the RHS exactly: every
    element's arithmetic depends only on its own and its neighbours' data."""
    g = ldg_amd.LDG(4, 32)
    g.initial_condition()
    u = g.get_state()
    rng = np.random.default_rng(axis)
    u[:, 3] += 0.01 * rng.standard_normal(u[:, 3].shape)
    k = g.rhs(u)
    G = u.reshape(32, 32, 32, 5, -1)
    us = np.roll(G, 1, axis=2 - axis).reshape(u.shape)
    ks = g.rhs(us)
    assert np.array_equal(ks, np.roll(k.reshape(32, 32, 32, 5, -1), 1, axis=2 - axis).reshape(k.shape))


def test_mass_conserved_over_steps_at_32cubed_p4():
    g = ldg_amd.LDG(4, 32)
    g.initial_condition()
    from ldg_numpy import lagrange_ops
    _, _, w, I, _, _ = lagrange_ops(4)
    m1 = (I.T @ np.diag(w) @ I).sum(axis=0)        # 1^T M1
    wts = np.einsum("i,j,k->ijk", m1, m1, m1).ravel()

    def mass(u):
        return np.sum(u[:, 0, :] @ wts)

    m0 = mass(g.get_state())
    g.step(g.time_step(0.01), 5, record_ke=False)
    assert abs(mass(g.get_state()) - m0) <= 1e-13 * m0


def test_step_rejects_bad_dt_and_reports_nonfinite():
    g = ldg_amd.LDG(2, 3)
    with pytest.raises(ldg_amd.LDGError) as e:
        g.step(-1.0, 1)
    assert e.value.code == ldg_amd.LDG_EINVAL
    u = np.zeros(g.shape)
    u[:, 0] = 1.0
    u[:, 4] = 100.0
    u[0, 0, 0] = np.nan
    g.set_state(u)
    with pytest.raises(ldg_amd.LDGError) as e:
        g.step(0.01, 1)
    assert e.value.code == ldg_amd.LDG_ENONFINITE


def test_cpp_driver_runs_and_writes_snapshots(tmp_path):
    """The C++ host (Mesh / Solver / dgTimeStep over the C ABI) runs the
    Taylor-Green case end to end and writes reference-format snapshots."""
    import subprocess
    import os
    exe = os.path.join(os.path.dirname(ldg_amd.LIB_PATH), "ldg_driver")
    out = tmp_path / "output"
    r = subprocess.run([exe, "-p", "2", "-nx", "4", "-tf", "0.02", "-dtSnaps", "5", "-out", str(out)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Outputting kinetic energy dissipation" in r.stdout
    snap = (out / "xyzu.txt.0").read_text().splitlines()
    assert snap[0] == "X, Y, Z, u0, u1, u2, u3, u4"
    assert len(snap) == 1 + 27 * 64
    o = O.Oracle(2, 4)
    u0 = o.initial_condition()
    vals = np.array([[float(v) for v in line.split(",")] for line in snap[1:]])
    np.testing.assert_allclose(vals[:, 3:].reshape(64, 27, 5), u0.transpose(0, 2, 1), rtol=1e-5, atol=1e-6)
    assert (out / "xyzu.txt.1").exists()


def _ckpt_meta(prec, step, timesteps, dt, ke):
    return dict(p=2, nx=4, ny=4, nz=4, precision=prec, step=step, timesteps=timesteps, dt=dt, tf=timesteps * dt,
                L=1.0, Re=1600.0, M0=0.1, V0=1.0, Pr=0.71, lo=[-np.pi] * 3, hi=[np.pi] * 3, ke=ke)


@pytest.mark.parametrize("prec", [ldg_amd.LDG_FP64, ldg_amd.LDG_MIXED])
def test_checkpoint_resume_is_bitwise(tmp_path, prec):
    """Stop after 2 of 6 steps, write a checkpoint, read it into a new solver
    and continue: state and KE history equal the uninterrupted run bitwise."""
    from ldg_amd import io
    print(f"checkpoint resume precision={prec}", flush=True)
    g = ldg_amd.LDG(2, 4, precision=prec)
    g.initial_condition()
    dt = g.time_step(0.01)
    ke_all = g.step(dt, 6, record_ke=True)
    u_all = g.get_state()
    g.close()
    g = ldg_amd.LDG(2, 4, precision=prec)
    g.initial_condition()
    ke1 = g.step(dt, 2, record_ke=True)
    io.save_checkpoint(tmp_path / "c.ckpt", g.get_state(), **_ckpt_meta(prec, 2, 6, dt, ke1))
    g.close()
    c = io.load_checkpoint(tmp_path / "c.ckpt")
    g = ldg_amd.LDG(2, 4, precision=prec)
    g.set_state(c["u"])
    ke2 = g.step(c["dt"], c["timesteps"] - c["step"], record_ke=True)
    assert np.array_equal(g.get_state(), u_all)
    assert np.array_equal(np.concatenate([c["ke"], ke2]), ke_all)
    g.close()


def test_cpp_driver_resumes_a_checkpoint(tmp_path):
    """The C++ Solver resumes a checkpoint written from Python at step 3 of
    its own run and ends bitwise where the uninterrupted run ends."""
    import os
    import subprocess
    from ldg_amd import io
    tf = 0.02
    g = ldg_amd.LDG(2, 4)
    g.initial_condition()
    dt0 = g.time_step(0.01)
    n = int(np.ceil(tf / dt0))
    dt = tf / n
    ke_all = g.step(dt, n, record_ke=True)
    u_all = g.get_state()
    g.close()
    g = ldg_amd.LDG(2, 4)
    g.initial_condition()
    ke1 = g.step(dt, 3, record_ke=True)
    meta = _ckpt_meta(ldg_amd.LDG_FP64, 3, n, dt, ke1)
    meta["tf"] = tf
    io.save_checkpoint(tmp_path / "mid.ckpt", g.get_state(), **meta)
    g.close()
    exe = os.path.join(os.path.dirname(ldg_amd.LIB_PATH), "ldg_driver")
    r = subprocess.run([exe, "-p", "2", "-nx", "4", "-tf", str(tf), "-nosnap", "-quiet", "-resume", str(tmp_path / "mid.ckpt"),
                        "-checkpoint", str(tmp_path / "end.ckpt")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    c = io.load_checkpoint(tmp_path / "end.ckpt")
    assert c["step"] == n and c["timesteps"] == n and c["dt"] == dt
    assert np.array_equal(c["u"], u_all)
    assert np.array_equal(c["ke"], ke_all)
    # a checkpoint from another run is refused
    r = subprocess.run([exe, "-p", "2", "-nx", "4", "-tf", str(2 * tf), "-nosnap", "-quiet", "-resume", str(tmp_path / "mid.ckpt")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "different run" in r.stderr


def test_cpp_driver_wall_limit_dumps_and_stops_cleanly(tmp_path):
    """A wall limit that has already passed (-wallLimit 0) makes the C++
    Solver dump the state (xyznavier.tmp.0, the reference's name plus the
    snapshot index its XYZV writer appends) and a checkpoint before its first
    step and return normally (the reference's 19.5 h guard, solver.cpp:349-360);
    resuming from that checkpoint ends bitwise where an uninterrupted run
    ends."""
    import os
    import subprocess
    from ldg_amd import io
    print("C++ driver wall limit", flush=True)
    tf = 0.02
    exe = os.path.join(os.path.dirname(ldg_amd.LIB_PATH), "ldg_driver")
    out = tmp_path / "out"
    r = subprocess.run([exe, "-p", "2", "-nx", "4", "-tf", str(tf), "-quiet", "-out", str(out), "-wallLimit", "0",
                        "-checkpoint", str(tmp_path / "wall.ckpt")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert (out / "xyznavier.tmp.0").stat().st_size > 0
    c = io.load_checkpoint(tmp_path / "wall.ckpt")
    assert c["step"] == 0
    r = subprocess.run([exe, "-p", "2", "-nx", "4", "-tf", str(tf), "-nosnap", "-quiet", "-resume", str(tmp_path / "wall.ckpt"),
                        "-checkpoint", str(tmp_path / "end.ckpt")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    end = io.load_checkpoint(tmp_path / "end.ckpt")
    g = ldg_amd.LDG(2, 4)
    g.initial_condition()
    ke_all = g.step(end["dt"], end["timesteps"], record_ke=True)
    u_all = g.get_state()
    g.close()
    assert end["step"] == end["timesteps"]
    assert np.array_equal(end["u"], u_all)
    assert np.array_equal(end["ke"], ke_all)


def test_cpp_driver_wall_limit_without_snapshots_writes_checkpoint(tmp_path):
    """-nosnap with a passed wall limit: the output directory is created for
    the state dump, the checkpoint is written, and the run returns 0."""
    import os
    import subprocess
    from ldg_amd import io
    exe = os.path.join(os.path.dirname(ldg_amd.LIB_PATH), "ldg_driver")
    out = tmp_path / "never_made"
    r = subprocess.run([exe, "-p", "2", "-nx", "4", "-tf", "0.02", "-nosnap", "-quiet", "-out", str(out),
                        "-wallLimit", "0", "-checkpoint", str(tmp_path / "w.ckpt")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert io.load_checkpoint(tmp_path / "w.ckpt")["step"] == 0
    assert (out / "xyznavier.tmp.0").stat().st_size > 0


def test_cpp_host_check_on_the_gpu(tmp_path):
    """tests/host_solver_check.cpp (the reference's driver block, a wall-limit
    stop continued and resumed, checkpoints refused across precisions)
    against the real libldg.so: the continued and the resumed run end with
    the same KE history bitwise."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    host = os.path.join(root, "local-dg_amd", "host")
    libdir = os.path.dirname(ldg_amd.LIB_PATH)
    exe = tmp_path / "host_solver_check"
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-I" + host, "-I" + os.path.join(root, "include"),
                    os.path.join(root, "tests", "host_solver_check.cpp")] +
                   [os.path.join(host, f) for f in ("mesh.cpp", "solver.cpp", "io.cpp")] +
                   ["-L" + libdir, "-lldg", "-Wl,-rpath," + libdir, "-o", str(exe)], check=True, capture_output=True,
                   text=True)
    r = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("ok")


@pytest.mark.parametrize("graph", [False, True])
def test_long_calls_record_ke_in_chunks(graph):
    """The device KE history holds 1024 steps; a longer call copies it out in
    chunks.  2,101 steps in one call and in calls of 700, 1 and 1,400 steps
    end with the same state and KE history, bitwise (eager and graph
    replay)."""
    print(f"KE history chunks, graph={graph}", flush=True)
    runs = []
    for split in ([2101], [700, 1, 1400]):
        g = ldg_amd.LDG(1, 3)
        g.initial_condition()
        g.set_graph(graph)
        dt = g.time_step(0.01)
        ke = np.concatenate([g.step(dt, n, record_ke=True) for n in split])
        runs.append((g.get_state(), ke))
        g.close()
    assert runs[0][1].shape == (2101,)
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
