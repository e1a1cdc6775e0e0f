"""Learned closure for an under-resolved LDG run (libldg.so's ldg_closure_*).

No reference counterpart: the reference solver has no closure.  The
closure is a per-element MLP (state + LDG gradient of one element -> a
correction of its right-hand side) trained on high-order runs of this same
solver, L2-projected to the coarse order:

    target = P rhs_fine(u_fine) - rhs_coarse(P u_fine)

Everything numeric (projection, feature assembly, forward, backward, Adam,
inference inside the RK4 stages) runs in libldg.so on the GPU; this module
only drives it.

    python -m ldg_amd.closure train --runs 4 --out closure.npz
    python -m ldg_amd.closure evaluate --weights closure.npz
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

from . import (LDG, LDG_CLOSURE_DISSIPATIVE, LDG_CLOSURE_ENERGIZING, LDG_CLOSURE_PER_STEP, LDG_FP64, LDG_MIXED, LDG_OK,
               Ensemble, LDGError, load_library)

_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)


def _f(a):
    return a.ctypes.data_as(_fp)


class Closure:
    """Owner of an ldg_closure (order p, `hidden` units per hidden layer,
    room for `capacity` training samples)."""

    def __init__(self, p=2, hidden=256, capacity=0, seed=0):
        self._lib = load_library()
        self._c = ctypes.c_void_p()
        rc = self._lib.ldg_closure_create(p, hidden, capacity, seed, ctypes.byref(self._c))
        if rc != LDG_OK:
            raise LDGError(rc, self._lib.ldg_closure_last_error(None).decode())
        nin, nout, hid, npar = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_long()
        self._lib.ldg_closure_sizes(self._c, ctypes.byref(nin), ctypes.byref(nout), ctypes.byref(hid),
                                    ctypes.byref(npar))
        self.p, self.nin, self.nout, self.hidden, self.nparam = p, nin.value, nout.value, hid.value, npar.value

    def close(self):
        if getattr(self, "_c", None) and self._c.value:
            self._lib.ldg_closure_destroy(self._c)
            self._c = ctypes.c_void_p()

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != LDG_OK:
            raise LDGError(rc, self._lib.ldg_closure_last_error(self._c).decode())

    # parameters -----------------------------------------------------------
    def get(self):
        params = np.empty(self.nparam, dtype=np.float32)
        norm = np.empty(2 * (self.nin + self.nout), dtype=np.float32)
        self._check(self._lib.ldg_closure_get(self._c, _f(params), _f(norm)))
        return params, norm

    def set(self, params=None, norm=None):
        p = None if params is None else np.ascontiguousarray(params, dtype=np.float32)
        n = None if norm is None else np.ascontiguousarray(norm, dtype=np.float32)
        if p is not None:
            assert p.shape == (self.nparam,)
        if n is not None:
            assert n.shape == (2 * (self.nin + self.nout),)
        self._check(self._lib.ldg_closure_set(self._c, _f(p) if p is not None else None,
                                              _f(n) if n is not None else None))

    def unpack(self, params):
        """(W1, b1, W2, b2, W3, b3) views of a flat parameter vector."""
        h, i, o = self.hidden, self.nin, self.nout
        sizes = [h * i, h, h * h, h, o * h, o]
        shapes = [(h, i), (h,), (h, h), (h,), (o, h), (o,)]
        out, off = [], 0
        for n, s in zip(sizes, shapes):
            out.append(params[off:off + n].reshape(s))
            off += n
        return out

    def save(self, path, **meta):
        params, norm = self.get()
        np.savez(path, params=params, norm=norm, p=self.p, hidden=self.hidden,
                 meta=json.dumps(meta))

    @classmethod
    def load(cls, path):
        d = np.load(path, allow_pickle=False)
        c = cls(int(d["p"]), int(d["hidden"]))
        c.set(d["params"], d["norm"])
        return c

    # evaluation on host arrays ---------------------------------------------
    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.empty((X.shape[0], self.nout), dtype=np.float32)
        self._check(self._lib.ldg_closure_forward(self._c, _f(X), X.shape[0], _f(Y)))
        return Y

    def gradient(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.empty(self.nparam, dtype=np.float32)
        loss = ctypes.c_double()
        self._check(self._lib.ldg_closure_gradient(self._c, _f(X), _f(T), X.shape[0], _f(g), ctypes.byref(loss)))
        return g, loss.value

    # dataset and training ---------------------------------------------------
    def collect(self, fine, coarse, keep=1.0, seed=0):
        self._check(self._lib.ldg_closure_collect(self._c, fine._h, coarse._h, keep, seed))

    def add_samples(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        self._check(self._lib.ldg_closure_add_samples(self._c, _f(X), _f(T), X.shape[0]))

    @property
    def samples(self):
        n = ctypes.c_long()
        self._check(self._lib.ldg_closure_samples(self._c, ctypes.byref(n)))
        return n.value

    def get_samples(self, first=0, n=None):
        n = self.samples - first if n is None else n
        X = np.empty((n, self.nin), dtype=np.float32)
        T = np.empty((n, self.nout), dtype=np.float32)
        self._check(self._lib.ldg_closure_get_samples(self._c, first, n, _f(X), _f(T)))
        return X, T

    def set_weighting(self, mode):
        """0: uniform; 1: every collection (snapshot) weighted by the inverse
        of its mean normalised target energy."""
        self._check(self._lib.ldg_closure_set_weighting(self._c, int(mode)))

    def get_weights(self, first=0, n=None):
        n = self.samples - first if n is None else n
        w = np.empty(n, dtype=np.float32)
        self._check(self._lib.ldg_closure_get_weights(self._c, first, n, _f(w)))
        return w

    def fit_normalization(self):
        self._check(self._lib.ldg_closure_fit_normalization(self._c))

    def train(self, epochs, batch=2048, lr=1e-3, val_fraction=0.1, seed=0):
        losses = np.empty(2 * epochs)
        self._check(self._lib.ldg_closure_train(self._c, epochs, batch, lr, val_fraction, seed,
                                                losses.ctypes.data_as(_dp)))
        return losses.reshape(epochs, 2)


def project_state(fine, i, coarse, j):
    """L2 projection of instance i of `fine` onto instance j of `coarse`."""
    lib = load_library()
    rc = lib.ldg_project_state(fine._h, i, coarse._h, j)
    if rc != LDG_OK:
        raise LDGError(rc, lib.ldg_last_error(coarse._h).decode())


def attach(solver, closure, bound, dissipative=False, per_step=False):
    """ldg_set_closure: closure=None detaches; dissipative=True also gates
    off corrections that would raise an element's kinetic energy,
    dissipative="energizing" those that would lower it; per_step=True
    evaluates the correction once per RK4 step (stage 1) and adds it in
    all four stages."""
    lib = load_library()
    rc = lib.ldg_set_closure(solver._h, closure._c if closure is not None else None, bound,
                             (LDG_CLOSURE_ENERGIZING if dissipative == "energizing" else
                              LDG_CLOSURE_DISSIPATIVE if dissipative else 0) | (LDG_CLOSURE_PER_STEP if per_step else 0))
    if rc != LDG_OK:
        raise LDGError(rc, lib.ldg_last_error(solver._h).decode())


def gated(solver):
    lib = load_library()
    off, tot = ctypes.c_long(), ctypes.c_long()
    rc = lib.ldg_closure_gated(solver._h, ctypes.byref(off), ctypes.byref(tot))
    if rc != LDG_OK:
        raise LDGError(rc, lib.ldg_last_error(solver._h).decode())
    return off.value, tot.value


# ---------------------------------------------------------------------------
# training data from high-order runs, training, and an evaluation run
# ---------------------------------------------------------------------------

def training_instances(runs, seed, ne=24, M0=0.1, re_range=(800.0, 3200.0)):
    """Seeded Taylor-Green variants: Re log-uniform in re_range, V0 =
    1 + 0.05 N(0,1), phases uniform in [0, 2 pi)^3, Mach M0."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(runs):
        Re = math.exp(rng.uniform(math.log(re_range[0]), math.log(re_range[1])))
        out.append({"nx": ne, "Re": Re, "M0": M0, "V0": 1.0 + 0.05 * rng.standard_normal(),
                    "phase": tuple(rng.uniform(0, 2 * math.pi, 3))})
    return out


def steps_to(dt_max, interval):
    """Integer step count and step size covering `interval` with dt <= dt_max."""
    n = max(1, int(math.ceil(interval / dt_max - 1e-9)))
    return n, interval / n


def generate(closure, runs=4, p_fine=5, ne=24, t_end=10.0, dt_snap=0.1, keep=0.25, cfl=0.01,
             precision=LDG_MIXED, seed=0, log=print):
    """Runs `runs` fine Taylor-Green variants to t_end as one ensemble and
    appends (features, target) samples of every snapshot (every dt_snap)."""
    inst = training_instances(runs, seed, ne)
    fine = Ensemble([dict(d, p=p_fine) for d in inst], precision=precision)
    coarse = Ensemble([dict(d, p=closure.p) for d in inst], precision=LDG_FP64)
    fine.initial_condition()
    dts = fine.time_steps(cfl)
    plan = [steps_to(dt, dt_snap) for dt in dts]
    nsteps = [n for n, _ in plan]
    if len(set(nsteps)) != 1:
        # one ensemble step call advances all instances by the same count:
        # use the largest count (smallest dt) for all, each with its own dt
        n = max(nsteps)
        plan = [(n, dt_snap / n) for _ in plan]
    n_per_snap = plan[0][0]
    dt_vec = np.array([d for _, d in plan])
    nsnap = int(round(t_end / dt_snap))
    t0 = time.perf_counter()
    for k in range(nsnap + 1):
        closure.collect(fine, coarse, keep, seed * 100003 + k)
        if k == nsnap:
            break
        fine.step(dt_vec, n_per_snap, record_ke=True)     # raises if a run blows up
        if log and (k % 10 == 9 or k == 0):
            log(f"closure data: snapshot {k + 1}/{nsnap} t={(k + 1) * dt_snap:.2f} "
                f"samples={closure.samples} wall={time.perf_counter() - t0:.1f}s")
    fine.close()
    coarse.close()
    return {"runs": runs, "instances": inst, "p_fine": p_fine, "ne": ne, "t_end": t_end, "dt_snap": dt_snap,
            "keep": keep, "cfl": cfl, "steps_per_snapshot": n_per_snap, "samples": closure.samples,
            "wall_s": time.perf_counter() - t0}


def run_tgv(p, ne, t_end, closure=None, bound=5.0, precision=LDG_FP64, cfl=0.01, dt_diag=0.1, Re=1600.0,
            M0=0.1, log=None, dissipative=False, ic_from=None, t_on=0.0, per_step=False):
    """Taylor-Green at order p on ne^3 to t_end (reference time-step rule);
    returns times and reference KE after every step, Gauss-point KE and
    enstrophy every dt_diag, wall time of the stepping, gated elements.
    ic_from = q > p starts from the L2 projection of the order-q initial
    state (the states the closure is trained on) instead of the nodal one.
    The closure is attached from time t_on on (rounded to a multiple of
    dt_diag); before that the run is the plain order-p run."""
    g = LDG(p, ne, Re=Re, M0=M0, precision=precision)
    k_on = int(round(t_on / dt_diag))
    attached = closure is not None and k_on <= 0
    if attached:
        attach(g, closure, bound, dissipative, per_step)
    if ic_from:
        f = LDG(ic_from, ne, Re=Re, M0=M0)
        f.initial_condition()
        project_state(f, 0, g, 0)
        f.close()
    else:
        g.initial_condition()
    n_per, dt = steps_to(g.time_step(cfl), dt_diag)
    ndiag = int(round(t_end / dt_diag))
    kes, diag = [], [g.diagnostics()]
    wall = 0.0
    off_max = 0
    diverged = None
    for k in range(ndiag):
        if closure is not None and not attached and k >= k_on:
            attach(g, closure, bound, dissipative, per_step)
            attached = True
        t0 = time.perf_counter()
        try:
            kes.append(g.step(dt, n_per, record_ke=True))
        except LDGError:
            diverged = (k + 1) * dt_diag          # non-finite kinetic energy within this interval
            break
        wall += time.perf_counter() - t0
        diag.append(g.diagnostics())
        if attached:
            off_max = max(off_max, gated(g)[0])
        if log and k % 10 == 9:
            log(f"p={p} t={(k + 1) * dt_diag:.1f} KE={diag[-1]['ke']:.6f} wall={wall:.1f}s")
    g.close()
    ke = np.concatenate(kes) if kes else np.zeros(0)
    return {"p": p, "ne": ne, "dt": dt, "steps": int(len(ke)), "ke_ref": ke,
            "t_diag": np.arange(len(diag)) * dt_diag, "ke_gauss": np.array([d["ke"] for d in diag]),
            "enstrophy": np.array([d["enstrophy"] for d in diag]), "wall_s": wall, "gated_max": off_max,
            "diverged_at": diverged}


def dissipation_curve(t, ke):
    """-dEk/dt by centred differences (one-sided at the ends)."""
    return -np.gradient(ke, t)


def compare(run, ref):
    """Relative max error of -dEk/dt (Gauss-point KE every dt_diag) of run vs
    ref on ref's diagnostic times, normalised by max |ref| (inf if the run
    diverged)."""
    if run.get("diverged_at") is not None or len(run["t_diag"]) < len(ref["t_diag"]):
        return float("inf")
    d_run = dissipation_curve(run["t_diag"], run["ke_gauss"])
    d_ref = dissipation_curve(ref["t_diag"], ref["ke_gauss"])
    d_run = np.interp(ref["t_diag"], run["t_diag"], d_run)
    return float(np.max(np.abs(d_run - d_ref)) / np.max(np.abs(d_ref)))


def projected_run(p_fine, p, ne, t_end, Re=1600.0, M0=0.1, cfl=0.01, dt_diag=0.1, log=None):
    """The order-p_fine Taylor-Green run and its L2 projection onto order p
    every dt_diag: the Gauss-point KE of both.  A closure trained on the
    target P rhs_fine(u) - rhs_p(P u) makes the corrected order-p run follow
    the projection P u (exactly, if the closure were exact and the run
    started from P u(0)), so -d/dt of the projection's KE is the
    dissipation curve such a closure can reach at best."""
    f = LDG(p_fine, ne, Re=Re, M0=M0)
    g = LDG(p, ne, Re=Re, M0=M0)
    f.initial_condition()
    n_per, dt = steps_to(f.time_step(cfl), dt_diag)
    ndiag = int(round(t_end / dt_diag))
    ke_f, ke_p = [], []
    for k in range(ndiag + 1):
        project_state(f, 0, g, 0)
        ke_f.append(f.diagnostics()["ke"])
        ke_p.append(g.diagnostics()["ke"])
        if k == ndiag:
            break
        f.step(dt, n_per)
        if log and k % 10 == 9:
            log(f"p={p_fine} t={(k + 1) * dt_diag:.1f} KE={ke_f[-1]:.6f} projected KE={ke_p[-1]:.6f}")
    f.close()
    g.close()
    t = np.arange(ndiag + 1) * dt_diag
    return {"t_diag": t, "ke_gauss": np.array(ke_f)}, {"t_diag": t, "ke_gauss": np.array(ke_p)}


def apriori(closure, times, ne=24, p_fine=5, Re=1600.0, M0=0.1, seed=12345, precision=LDG_MIXED, cfl=0.01,
            log=None):
    """A-priori accuracy of a trained closure on a fresh fine run: at each
    time, the relative error ||c - T|| / ||T|| of the correction c against
    the true target T = P rhs_fine - rhs_coarse(P u_fine), and the
    explained fraction 1 - ||c - T||^2 / ||T - mean T||^2."""
    inst = [{"nx": ne, "Re": Re, "M0": M0, "V0": 1.0, "phase": (0.0, 0.0, 0.0)}]
    fine = Ensemble([dict(inst[0], p=p_fine)], precision=precision)
    coarse = Ensemble([dict(inst[0], p=closure.p)])
    fine.initial_condition()
    _, norm = closure.get()
    nin, nout = closure.nin, closure.nout
    out, t_now = [], 0.0
    for t in times:
        if t > t_now:
            n, dt = steps_to(fine.time_steps(cfl)[0], t - t_now)
            fine.step(np.array([dt]), n, record_ke=True)
            t_now = t
        probe = Closure(closure.p, 16, ne ** 3)
        probe.collect(fine, coarse, keep=1.0, seed=seed)
        X, T = probe.get_samples()
        probe.close()
        Y = closure.forward(((X - norm[:nin]) * norm[nin:2 * nin]).astype(np.float32)).astype(np.float64)
        c = Y * norm[2 * nin + nout:] + norm[2 * nin:2 * nin + nout]
        T = T.astype(np.float64)
        rel_err = float(np.linalg.norm(c - T) / np.linalg.norm(T))
        expl = float(1.0 - np.sum((c - T) ** 2) / np.sum((T - T.mean(0)) ** 2))
        out.append({"t": t, "rel_err": rel_err, "explained": expl, "target_rms": float(np.sqrt(np.mean(T ** 2)))})
        if log:
            log(f"a priori t={t}: relative error {rel_err:.3f}, explained {expl:.3f}")
    fine.close()
    coarse.close()
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m ldg_amd.closure")
    sub = ap.add_subparsers(dest="cmd", required=True)
    tr = sub.add_parser("train")
    tr.add_argument("--runs", type=int, default=32)
    tr.add_argument("--p-fine", type=int, default=5)
    tr.add_argument("--ne", type=int, default=24)
    tr.add_argument("--t-end", type=float, default=10.0)
    tr.add_argument("--dt-snap", type=float, default=0.1)
    tr.add_argument("--keep", type=float, default=0.25)
    tr.add_argument("--cfl", type=float, default=0.01)
    tr.add_argument("--epochs", type=int, default=30)
    tr.add_argument("--batch", type=int, default=2048)
    tr.add_argument("--lr", type=float, default=1e-3)
    tr.add_argument("--seed", type=int, default=0)
    tr.add_argument("--weighting", choices=["uniform", "snapshot"], default="uniform")
    tr.add_argument("--out", default="closure.npz")
    ev = sub.add_parser("evaluate")
    ev.add_argument("--weights", required=True)
    ev.add_argument("--ne", type=int, default=24)
    ev.add_argument("--t-end", type=float, default=10.0)
    ev.add_argument("--bounds", default="5", help="comma-separated gate bounds")
    ev.add_argument("--gates", default="plain", help="comma-separated: plain, dissipative, energizing")
    ev.add_argument("--relax", default="1", help="comma-separated factors multiplying the correction")
    ev.add_argument("--ic-from", type=int, default=0, help="start p=2 runs from the projected order-q state")
    ev.add_argument("--t-on", default="0", help="comma-separated times from which the closure is attached")
    ev.add_argument("--per-step", action="store_true", help="evaluate the correction once per RK4 step")
    ev.add_argument("--re", type=float, default=1600.0, help="Reynolds number of the evaluation case")
    ev.add_argument("--precision", choices=["fp64", "mixed"], default="fp64", help="precision of the p=2 runs")
    ev.add_argument("--ref-p", type=int, default=5)
    ev.add_argument("--ref-json", default=None, help="take the resolved run from an earlier evaluate output")
    ev.add_argument("--no-ref", action="store_true")
    ev.add_argument("--out", default=None)
    pr = sub.add_parser("projected")
    pr.add_argument("--p-fine", type=int, default=5)
    pr.add_argument("--ne", type=int, default=24)
    pr.add_argument("--t-end", type=float, default=10.0)
    pr.add_argument("--out", default=None)
    ap_ = sub.add_parser("apriori")
    ap_.add_argument("--weights", required=True)
    ap_.add_argument("--ne", type=int, default=24)
    ap_.add_argument("--times", default="0,1,2,4,6,8,9.5")
    a = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    if a.cmd == "projected":
        fine, proj = projected_run(a.p_fine, 2, a.ne, a.t_end, log=log)
        d_f = dissipation_curve(fine["t_diag"], fine["ke_gauss"])
        d_p = dissipation_curve(proj["t_diag"], proj["ke_gauss"])
        res = {"dEk_dt_err_projected": compare(proj, fine), "t": fine["t_diag"].tolist(),
               "ke_fine": fine["ke_gauss"].tolist(), "ke_projected": proj["ke_gauss"].tolist(),
               "dissipation_fine": d_f.tolist(), "dissipation_projected": d_p.tolist()}
        if a.out:
            with open(a.out, "w") as fo:
                json.dump(res, fo)
        print(json.dumps({"dEk_dt_err_projected": res["dEk_dt_err_projected"]}), flush=True)
        return
    if a.cmd == "apriori":
        c = Closure.load(a.weights)
        res = apriori(c, [float(t) for t in a.times.split(",")], a.ne, log=log)
        print(json.dumps(res), flush=True)
        return
    if a.cmd == "train":
        ne, t_end = a.ne, a.t_end
        nsnap = int(round(t_end / a.dt_snap)) + 1
        cap = int(a.runs * nsnap * math.ceil(a.keep * ne ** 3)) + 16
        c = Closure(2, 256, cap, a.seed)
        info = generate(c, a.runs, a.p_fine, ne, t_end, a.dt_snap, a.keep, a.cfl, seed=a.seed, log=log)
        c.set_weighting(1 if a.weighting == "snapshot" else 0)
        info["weighting"] = a.weighting
        c.fit_normalization()
        t0 = time.perf_counter()
        losses = c.train(a.epochs, a.batch, a.lr, 0.1, a.seed)
        info["train_wall_s"] = time.perf_counter() - t0
        info["losses"] = losses.tolist()
        c.save(a.out, **{k: v for k, v in info.items() if k != "instances"})
        print(json.dumps({k: v for k, v in info.items() if k != "instances"}), flush=True)
    else:
        c = Closure.load(a.weights)
        prec2 = LDG_MIXED if a.precision == "mixed" else LDG_FP64
        res, runs = {}, {}
        params, norm = c.get()
        nin, nout = c.nin, c.nout
        for relax in [float(r) for r in a.relax.split(",")]:
            n2 = norm.copy()
            n2[2 * nin:] *= relax          # correction = relax * (y * std + mean); the gate sees y
            c.set(norm=n2)
            for bound in [float(b) for b in a.bounds.split(",")]:
                for gate in a.gates.split(","):
                    for t_on in [float(t) for t in a.t_on.split(",")]:
                        tag = f"closure_b{bound:g}_{gate}_r{relax:g}" + (f"_on{t_on:g}" if t_on > 0 else "")
                        runs[tag] = run_tgv(2, a.ne, a.t_end, c, bound, precision=prec2, Re=a.re, log=log,
                                            dissipative=("energizing" if gate == "energizing" else gate == "dissipative"),
                                            ic_from=a.ic_from, t_on=t_on, per_step=a.per_step)
        c.set(norm=norm)
        runs["plain_p2"] = run_tgv(2, a.ne, a.t_end, precision=prec2, Re=a.re, log=log, ic_from=a.ic_from)
        ref = None
        if a.ref_json:
            d = json.load(open(a.ref_json))[f"plain_p{a.ref_p}"]
            ref = {"t_diag": np.array(d["t"]), "ke_gauss": np.array(d["ke_gauss"]), "wall_s": d["wall_s"],
                   "steps": d["steps"], "gated_max": 0, "enstrophy": np.array(d["enstrophy"])}
        elif not a.no_ref:
            ref = run_tgv(a.ref_p, a.ne, a.t_end, Re=a.re, log=log)
        if ref is not None:
            runs[f"plain_p{a.ref_p}"] = ref
            for k, r in runs.items():
                if k != f"plain_p{a.ref_p}":
                    res[f"dEk_dt_err_{k}"] = compare(r, ref)
                    res[f"wall_ratio_ref_over_{k}"] = ref["wall_s"] / r["wall_s"]
        for k, r in runs.items():
            res[k] = {"wall_s": r["wall_s"], "steps": r["steps"], "gated_max": r["gated_max"],
                      "diverged_at": r.get("diverged_at"),
                      "t": r["t_diag"].tolist(), "ke_gauss": r["ke_gauss"].tolist(),
                      "dissipation": dissipation_curve(r["t_diag"], r["ke_gauss"]).tolist(),
                      "enstrophy": r["enstrophy"].tolist()}
        if a.out:
            with open(a.out, "w") as f:
                json.dump(res, f)
        print(json.dumps({k: v for k, v in res.items() if not isinstance(v, dict)}), flush=True)


if __name__ == "__main__":
    main()
