"""The learned closure on the GPU (libldg.so ldg_closure_*, through the C ABI).

Checked against the numpy restatement in tests/mlp_numpy.py (itself pinned
by finite differences and projection identities in test_closure_ref.py):
- forward pass and parameter gradient, with the BF16 rounding of the
  matrix-core operands emulated in numpy (FP64 accumulation there, FP32 on
  the GPU): 1e-3 relative -- an activation that lands within FP32 rounding
  of a BF16 rounding boundary rounds to the neighbouring BF16 value on one
  side, one BF16 ulp (2^-8) in a few operands;
- the element-wise L2 projection between orders: 1e-13 relative;
- collected training samples against the oracle's right-hand sides of the
  fine state and of its projection;
- inside the RK4 stage: a closed gate (bound 0) is bit-identical to the
  plain solver; an open gate adds exactly the network's correction;
- training: deterministic in the seed and it reduces the loss on a task
  the network can represent.
"""
import numpy as np
import pytest

import ldg_amd
import mlp_numpy as M
import oracle as O
from ldg_amd import closure as C

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def test_forward_matches_numpy():
    c = C.Closure(2, 256, 0, seed=3)
    params, _ = c.get()
    rng = np.random.default_rng(0)
    X = rng.standard_normal((300, c.nin)).astype(np.float32)      # 300: not a multiple of the 64-row tile
    Y = c.forward(X)
    Yr, _ = M.forward(params, X, c.nin, c.hidden, c.nout, bf=True)
    e = rel(Y, Yr)
    print("forward relative error vs numpy (bf16 operands):", e, flush=True)
    assert e <= 1e-3
    # without emulating BF16 the difference is of BF16 size, far above FP32 rounding
    Yf, _ = M.forward(params, X, c.nin, c.hidden, c.nout, bf=False)
    assert 1e-5 < rel(Y, Yf) < 5e-2


def test_gradient_matches_numpy():
    c = C.Closure(2, 256, 0, seed=4)
    params, _ = c.get()
    rng = np.random.default_rng(1)
    B = 200
    X = rng.standard_normal((B, c.nin)).astype(np.float32)
    T = rng.standard_normal((B, c.nout)).astype(np.float32)
    g, loss = c.gradient(X, T)
    gr = M.gradient(params, X, T, c.nin, c.hidden, c.nout, bf=True)
    Yr, _ = M.forward(params, X, c.nin, c.hidden, c.nout, bf=True)
    assert abs(loss - np.mean((Yr - T) ** 2)) <= 1e-4 * loss
    for name, a, b in zip(["W1", "b1", "W2", "b2", "W3", "b3"], c.unpack(g), c.unpack(gr.astype(np.float32))):
        e = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        print(f"gradient {name}: relative error {e:.2e}", flush=True)
        assert e <= 1e-3, name


@pytest.mark.parametrize("pf,pc,mesh", [(4, 2, (5, 4, 3)), (5, 2, (3, 3, 3)), (3, 1, (4, 3, 3))])
def test_projection_matches_numpy(pf, pc, mesh):
    fine = ldg_amd.LDG(pf, *mesh)
    coarse = ldg_amd.LDG(pc, *mesh)
    rng = np.random.default_rng(pf)
    u = rng.standard_normal(fine.shape)
    fine.set_state(u)
    C.project_state(fine, 0, coarse, 0)
    e = rel(coarse.get_state(), M.project_state(u, pf, pc))
    assert e <= 1e-13, e


def collect_reference(p_f, p_c, ne, u_f):
    """Features (background removed) and targets from the oracle."""
    of, oc = O.Oracle(p_f, ne), O.Oracle(p_c, ne)
    k_f = of.rhs(u_f)
    Pu = M.project_state(u_f, p_f, p_c)
    Pk = M.project_state(k_f, p_f, p_c)
    k_c, du_c = oc.rhs(Pu, want_du=True)
    K, dofs = Pu.shape[0], Pu.shape[2]
    gamma, M0, rho0 = 1.4, 0.1, 1.0
    p0 = rho0 * (1.0 / M0) ** 2 / gamma
    E0 = p0 / (gamma - 1)
    w = Pu.copy()
    w[:, 0] -= rho0
    w[:, 4] -= E0
    q = np.transpose(du_c, (1, 0, 2, 3)).reshape(K, 15 * dofs)     # field s + 5 l
    X = np.concatenate([w.reshape(K, 5 * dofs), q], axis=1)
    T = (Pk - k_c).reshape(K, 5 * dofs)
    return X, T


def test_collected_samples_match_oracle():
    ne, pf, pc = 4, 4, 2
    fine = ldg_amd.Ensemble([{"p": pf, "nx": ne}])
    coarse = ldg_amd.Ensemble([{"p": pc, "nx": ne}])
    fine.initial_condition()
    fine.step(fine.time_steps(0.01), 3, record_ke=False)
    # break the Taylor-Green symmetries (symmetric elements would have equal features)
    u_f = fine.get_instance_state(0)
    u_f[:, 1:4] += 1e-3 * np.random.default_rng(9).standard_normal(u_f[:, 1:4].shape)
    fine.set_instance_state(0, u_f)
    c = C.Closure(pc, 64, ne ** 3)
    c.collect(fine, coarse, keep=1.0, seed=0)
    assert c.samples == ne ** 3
    X, T = c.get_samples()
    Xr, Tr = collect_reference(pf, pc, ne, u_f)
    ex = np.max(np.abs(X - Xr), axis=0) / np.maximum(np.max(np.abs(Xr), axis=0), 1e-30)
    print("features max relative error:", ex.max(), " target error:", rel(T, Tr), flush=True)
    assert ex.max() <= 1e-5
    assert rel(T, Tr) <= 1e-5
    # subsampling keeps whole rows of the same set
    c2 = C.Closure(pc, 64, ne ** 3)
    c2.collect(fine, coarse, keep=0.25, seed=7)
    X2, T2 = c2.get_samples()
    assert len(X2) == 16
    rows = [int(np.argmin(np.sum((X - x) ** 2, axis=1))) for x in X2]
    assert np.array_equal(X[rows], X2) and np.array_equal(T[rows], T2)


def random_norm(c, rng):
    nin, nout = c.nin, c.nout
    return np.concatenate([rng.standard_normal(nin) * 0.1, 0.5 + rng.random(nin),
                           rng.standard_normal(nout) * 1e-3, 1e-3 * (0.5 + rng.random(nout))]).astype(np.float32)


@pytest.mark.parametrize("precision", [ldg_amd.LDG_FP64, ldg_amd.LDG_MIXED])
def test_closed_gate_is_bitwise_plain(precision):
    plain = ldg_amd.LDG(2, 4, precision=precision)
    gated = ldg_amd.LDG(2, 4, precision=precision)
    c = C.Closure(2, 256, 0, seed=5)
    c.set(norm=random_norm(c, np.random.default_rng(2)))
    C.attach(gated, c, 0.0)
    for g in (plain, gated):
        g.initial_condition()
    u = plain.get_state()
    assert np.array_equal(plain.rhs(u), gated.rhs(u))
    assert C.gated(gated) == (64, 64)
    dt = plain.time_step(0.01)
    ka, kb = plain.step(dt, 3), gated.step(dt, 3)
    assert np.array_equal(plain.get_state(), gated.get_state())
    assert np.array_equal(ka, kb)


@pytest.mark.parametrize("hidden", [256, 64])
def test_open_gate_adds_the_network_correction(hidden):
    """hidden 256 at p = 2 runs the fused on-chip inference kernel in the
    stage, other widths the layer-by-layer path; both must add what the
    network computes on the same features."""
    plain = ldg_amd.LDG(2, 4)
    corr = ldg_amd.LDG(2, 4)
    c = C.Closure(2, hidden, 0, seed=6)
    rng = np.random.default_rng(3)
    norm = random_norm(c, rng)
    c.set(norm=norm)
    C.attach(corr, c, 1e30)
    plain.initial_condition()
    u = plain.get_state()
    k0, du = plain.rhs(u, want_du=True)
    k1 = corr.rhs(u)
    assert C.gated(corr) == (0, 64)
    # host-side features from the same state and gradient, then the GPU network
    K, dofs = u.shape[0], u.shape[2]
    E0 = (1.0 / 0.1) ** 2 / 1.4 / (1.4 - 1.0)
    w = u.copy()
    w[:, 0] -= 1.0
    w[:, 4] -= E0
    x = np.concatenate([w.reshape(K, -1), np.transpose(du, (1, 0, 2, 3)).reshape(K, -1)], axis=1)
    nin, nout = c.nin, c.nout
    X = ((x - norm[:nin]) * norm[nin:2 * nin]).astype(np.float32)
    Y = c.forward(X).astype(np.float64)
    expect = (Y * norm[2 * nin + nout:] + norm[2 * nin:2 * nin + nout]).reshape(k0.shape)
    e = rel(k1 - k0, expect)
    print("closure correction relative error:", e, flush=True)
    assert e <= 1e-4


def test_training_is_deterministic_and_learns_a_teacher():
    teacher = C.Closure(2, 256, 0, seed=11)
    rng = np.random.default_rng(4)
    n = 8192
    X = rng.standard_normal((n, teacher.nin)).astype(np.float32)
    T = teacher.forward(X)
    runs = []
    for _ in range(2):
        s = C.Closure(2, 256, n, seed=12)
        s.add_samples(X, T)
        s.fit_normalization()
        losses = s.train(12, batch=512, lr=1e-3, val_fraction=0.125, seed=5)
        runs.append((s.get()[0], losses))
    print("losses (train, val) per epoch:", runs[0][1].tolist(), flush=True)
    assert np.array_equal(runs[0][0], runs[1][0])
    losses = runs[0][1]
    assert np.all(np.isfinite(losses))
    assert losses[-1, 0] < 0.5 * losses[0, 0]
    assert losses[-1, 1] < 0.5 * losses[0, 0]


@pytest.mark.parametrize("hidden", [256, 64])
@pytest.mark.parametrize("mode", ["dissipative", "energizing"])
def test_energy_gate_keeps_only_corrections_of_one_sign(hidden, mode):
    """With the dissipative gate an element keeps its correction only if
    the correction does not raise its nodal kinetic energy
    sum_j |m_j|^2 / (2 rho_j), with the energizing gate only if it does not
    lower it; the others keep the plain right-hand side."""
    plain = ldg_amd.LDG(2, 4)
    corr = ldg_amd.LDG(2, 4)
    c = C.Closure(2, hidden, 0, seed=8)
    rng = np.random.default_rng(5)
    norm = random_norm(c, rng)
    c.set(norm=norm)
    C.attach(corr, c, 1e30, dissipative=True if mode == "dissipative" else "energizing")
    plain.initial_condition()
    u = plain.get_state()
    k0, du = plain.rhs(u, want_du=True)
    k1 = corr.rhs(u)
    K, dofs = u.shape[0], u.shape[2]
    E0 = (1.0 / 0.1) ** 2 / 1.4 / (1.4 - 1.0)
    w = u.copy()
    w[:, 0] -= 1.0
    w[:, 4] -= E0
    x = np.concatenate([w.reshape(K, -1), np.transpose(du, (1, 0, 2, 3)).reshape(K, -1)], axis=1)
    nin, nout = c.nin, c.nout
    Y = c.forward(((x - norm[:nin]) * norm[nin:2 * nin]).astype(np.float32)).astype(np.float64)
    full = (Y * norm[2 * nin + nout:] + norm[2 * nin:2 * nin + nout]).reshape(k0.shape)
    v = u[:, 1:4] / u[:, 0:1]
    prod = np.sum(np.sum(v * full[:, 1:4], axis=1) - 0.5 * np.sum(v * v, axis=1) * full[:, 0], axis=1)
    on = prod <= 0 if mode == "dissipative" else prod >= 0
    off, total = C.gated(corr)
    assert total == K and off == int(np.sum(~on)) and 0 < off < K
    d = k1 - k0
    assert np.max(np.abs(d[~on])) == 0.0
    assert rel(d[on], full[on]) <= 1e-4


def test_snapshot_weighting():
    """Mode 1 weighs each collection by the inverse of its mean normalised
    target energy (weights averaging 1); mode 0 leaves every weight 1."""
    rng = np.random.default_rng(6)
    n = 600
    X = rng.standard_normal((n, 540)).astype(np.float32)
    T1 = rng.standard_normal((n, 135)).astype(np.float32)
    T2 = (0.05 * rng.standard_normal((n, 135)) + 0.3).astype(np.float32)
    for mode in (0, 1):
        c = C.Closure(2, 64, 2 * n)
        c.add_samples(X, T1)
        c.add_samples(X, T2)
        c.set_weighting(mode)
        c.fit_normalization()
        w = c.get_weights()
        if mode == 0:
            assert np.all(w == 1.0)
            continue
        _, Tn = c.get_samples()
        m1 = np.mean(Tn[:n].astype(np.float64) ** 2)
        m2 = np.mean(Tn[n:].astype(np.float64) ** 2)
        assert np.allclose(w[:n], w[0]) and np.allclose(w[n:], w[n])
        assert abs(w[0] / w[n] - m2 / m1) <= 1e-5 * (m2 / m1)
        assert abs(np.mean(w.astype(np.float64)) - 1.0) <= 1e-5
        losses = c.train(2, batch=128, lr=1e-3, val_fraction=0.25, seed=1)
        assert np.all(np.isfinite(losses))


def test_corrected_run_is_deterministic():
    """Two runs with the same closure attached are bit-identical (no
    uninitialised memory or order-dependent sums in the corrected stage)."""
    c = C.Closure(2, 256, 0, seed=21)
    c.set(norm=random_norm(c, np.random.default_rng(7)))
    out = []
    for _ in range(2):
        g = ldg_amd.LDG(2, 8)
        C.attach(g, c, 1e30)
        g.initial_condition()
        ke = g.step(g.time_step(0.01), 40)
        out.append((g.get_state(), ke))
        g.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_late_switch_on_matches_plain_until_attached():
    """run_tgv(t_on=...): the run is the plain run bit for bit until t_on,
    and differs after the closure is attached there."""
    c = C.Closure(2, 256, 0, seed=21)
    c.set(norm=random_norm(c, np.random.default_rng(7)))
    plain = C.run_tgv(2, 4, 0.3, dt_diag=0.1)
    late = C.run_tgv(2, 4, 0.3, closure=c, bound=1e30, dt_diag=0.1, t_on=0.2)
    n_on = 2 * plain["steps"] // 3                  # steps before t = 0.2
    np.testing.assert_array_equal(late["ke_ref"][:n_on], plain["ke_ref"][:n_on])
    np.testing.assert_array_equal(late["ke_gauss"][:3], plain["ke_gauss"][:3])
    assert late["ke_gauss"][3] != plain["ke_gauss"][3]


def test_per_step_closure_holds_the_stage_one_correction():
    """LDG_CLOSURE_PER_STEP: the correction c0 of the step's initial state
    u0 is added in all four stages.  One step equals the RK4 step (the
    reference tableau) of k(U) + c0, restated here with the plain
    right-hand side k of the same library, to FP64 round-off."""
    c = C.Closure(2, 256, 0, seed=21)
    c.set(norm=random_norm(c, np.random.default_rng(7)))
    plain, probe, corr = ldg_amd.LDG(2, 4), ldg_amd.LDG(2, 4), ldg_amd.LDG(2, 4)
    C.attach(probe, c, 1e30)
    C.attach(corr, c, 1e30, per_step=True)
    corr.initial_condition()
    u0 = corr.get_state()
    c0 = probe.rhs(u0) - plain.rhs(u0)
    assert np.max(np.abs(c0)) > 0
    dt = corr.time_step(0.01)
    a, b = [0.5, 0.5, 1.0], [1 / 6, 1 / 3, 1 / 3, 1 / 6]
    U, acc = u0, u0.copy()
    for st in range(4):
        k = plain.rhs(U) + c0
        acc = acc + dt * b[st] * k
        if st < 3:
            U = u0 + dt * a[st] * k
    corr.step(dt, 1, record_ke=False)
    e = rel(corr.get_state(), acc)
    print("per-step closure vs restated RK4 step:", e, flush=True)
    assert e <= 1e-13
    # the per-stage closure (correction of every stage's input) differs
    both = ldg_amd.LDG(2, 4)
    C.attach(both, c, 1e30)
    both.initial_condition()
    both.step(dt, 1, record_ke=False)
    assert not np.array_equal(both.get_state(), corr.get_state())


def test_set_closure_refuses_bad_flags():
    """Both energy gates at once, or an unknown bit, are refused."""
    g = ldg_amd.LDG(2, 4)
    c = C.Closure(2, 256, 0, seed=3)
    lib = ldg_amd.load_library()
    both = ldg_amd.LDG_CLOSURE_DISSIPATIVE | ldg_amd.LDG_CLOSURE_ENERGIZING
    assert lib.ldg_set_closure(g._h, c._c, 1.0, both) == ldg_amd.LDG_EINVAL
    assert lib.ldg_set_closure(g._h, c._c, 1.0, 8) == ldg_amd.LDG_EINVAL
    assert lib.ldg_set_closure(g._h, c._c, 1.0, ldg_amd.LDG_CLOSURE_ENERGIZING | ldg_amd.LDG_CLOSURE_PER_STEP) == \
        ldg_amd.LDG_OK
    assert lib.ldg_set_closure(g._h, None, 0.0, 0) == ldg_amd.LDG_OK


@pytest.mark.parametrize("per_step", [False, True])
def test_closure_under_graph_replay_is_bitwise_eager(per_step):
    """A corrected run replayed from HIP graphs equals the eager run bit
    for bit, per stage and once per step (several replays, with eager
    steps and reads of the state in between)."""
    c = C.Closure(2, 256, 0, seed=21)
    c.set(norm=random_norm(c, np.random.default_rng(7)))
    runs = []
    for graph in (False, True):
        g = ldg_amd.LDG(2, 6)
        C.attach(g, c, 1e30, dissipative="energizing", per_step=per_step)
        g.set_graph(graph)
        g.initial_condition()
        dt = g.time_step(0.01)
        ke = [g.step(dt, 6), g.step(dt, 4)]
        mid = g.get_state()
        ke.append(g.step(dt, 8))
        runs.append((mid, g.get_state(), np.concatenate(ke)))
        C.attach(g, None, 0.0)
        g.close()
    for a, b in zip(runs[0], runs[1]):
        np.testing.assert_array_equal(a, b)


def test_adam_matches_numpy_adam():
    """Five full-batch Adam steps of ldg_closure_train against Algorithm 1
    of Kingma & Ba (2015) in float64 numpy: m_t = b1 m + (1-b1) g,
    v_t = b2 v + (1-b2) g^2, theta -= lr (m_t/(1-b1^t)) / (sqrt(v_t/(1-b2^t)) + eps),
    with g the GPU gradient at the previous parameters (ldg_closure_gradient
    on the same normalised batch).  A missing or wrong bias correction moves
    the first step by a factor of about 3 and fails the check."""
    n, steps, lr = 256, 5, 1e-3
    b1, b2, eps = 0.9, 0.999, 1e-8
    rng = np.random.default_rng(21)
    c = C.Closure(2, 256, n, seed=31)
    c.add_samples(rng.standard_normal((n, c.nin)).astype(np.float32),
                  rng.standard_normal((n, c.nout)).astype(np.float32))
    c.fit_normalization()
    Xn, Tn = c.get_samples()
    theta0, norm = c.get()
    probe = C.Closure(2, 256, 0, seed=0)
    theta = theta0.astype(np.float64)
    m = np.zeros_like(theta)
    v = np.zeros_like(theta)
    for t in range(1, steps + 1):
        probe.set(theta.astype(np.float32), norm)
        g, _ = probe.gradient(Xn, Tn)
        g = g.astype(np.float64)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        theta = theta - lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    c.train(steps, batch=n, lr=lr, val_fraction=0.0, seed=3)
    theta_gpu = c.get()[0].astype(np.float64)
    d_ref, d_gpu = theta - theta0, theta_gpu - theta0
    err = np.abs(d_gpu - d_ref)
    print(f"Adam: max |step| {np.abs(d_ref).max():.3e}, max |gpu - numpy| {err.max():.3e}, "
          f"median {np.median(err):.3e}", flush=True)
    # every parameter moves by about lr per step early in Adam
    assert np.median(np.abs(d_ref)) > 0.5 * lr * steps
    assert np.median(err) <= 1e-3 * np.median(np.abs(d_ref))
    assert np.quantile(err, 0.999) <= 5e-2 * np.abs(d_ref).max()
    probe.close()
    c.close()
