"""Writes tests/golden/c2_tgv32_p4_history.npz: the oracle's kinetic-energy
and enstrophy histories of the Taylor-Green vortex on 32^3 elements at p = 4
over 1000 RK4 steps (dt = 0.01 dx/(V0 (2p+1)), the reference's time-step rule,
solver.cpp:44-52).

Run in the build container (about 2.5 h on 8 cores):
    python tests/golden/make_c2_history.py [--checkpoint FILE]

Recorded:
  ke[s]          reference kinetic energy after step s+1 (Solver::computeKE,
                 solver.cpp:1007-1034, recorded every step as kes(iStep),
                 solver.cpp:437)
  diag_steps     0, 10, 20, ..., 1000
  ke_gauss, enstrophy   Gauss-point kinetic energy and enstrophy of the state
                 after those steps (no reference counterpart)
  final_norms    per-variable L2 norm and sum of the state after step 1000
  sample_elems, sample_state   the state after step 1000 on 32 seeded elements
The computation is the oracle's FP64 build for this host's vector units,
bitwise equal to the plain build (tests/test_oracle_build.py).
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402

P, NX, NSTEPS, EVERY, CFL = 4, 32, 1000, 10, 0.01
OUT = os.path.join(HERE, "c2_tgv32_p4_history.npz")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--checkpoint", default=None, help="file to save progress to and resume from")
    ap.add_argument("--variant", default="native")
    args = ap.parse_args()
    o = O.Oracle(P, NX, variant=args.variant)
    dt = o.dt(CFL)
    if args.checkpoint and os.path.exists(args.checkpoint):
        c = np.load(args.checkpoint)
        u, done = c["u"], int(c["done"])
        ke, ke_g, ens = list(c["ke"]), list(c["ke_gauss"]), list(c["enstrophy"])
        print(f"resuming after step {done}", flush=True)
    else:
        u, done = o.initial_condition(), 0
        d = o.diagnostics(u)
        ke, ke_g, ens = [], [d[0]], [d[1]]
    t0 = time.time()
    while done < NSTEPS:
        u, k = o.step(u, dt, EVERY)
        ke.extend(k)
        done += EVERY
        d = o.diagnostics(u)
        ke_g.append(d[0])
        ens.append(d[1])
        print(f"step {done}: ke {k[-1]:.15e} enstrophy {d[1]:.15e} ({time.time() - t0:.0f} s)", flush=True)
        if args.checkpoint:
            np.savez(args.checkpoint + ".tmp.npz", u=u, done=done, ke=ke, ke_gauss=ke_g, enstrophy=ens)
            os.replace(args.checkpoint + ".tmp.npz", args.checkpoint)
    rng = np.random.default_rng(2024)
    sample = np.sort(rng.choice(o.K, 32, replace=False))
    np.savez_compressed(
        OUT, p=P, nx=NX, nsteps=NSTEPS, cfl=CFL, dt=dt, ke=np.array(ke),
        diag_steps=np.arange(0, NSTEPS + 1, EVERY), ke_gauss=np.array(ke_g), enstrophy=np.array(ens),
        final_norms=np.sqrt(np.sum(u ** 2, axis=(0, 2))), final_sums=np.sum(u, axis=(0, 2)),
        sample_elems=sample, sample_state=u[sample])
    print(f"wrote {OUT}", flush=True)


if __name__ == "__main__":
    main()
