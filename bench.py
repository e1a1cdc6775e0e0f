#!/usr/bin/env python3
"""Throughput of the LDG right-hand side + RK4 stage pipeline on one MI355X.

A step is one RK4 step (4 stages, 8 kernel launches plus the kinetic-energy
reduction) of the compressible Taylor-Green vortex at 64^3 elements, p=4
(163,840,000 DOF), with the state resident in HBM.  The metric is DOF-stage
updates per second: 4 * steps * K * (p+1)^3 * 5 / wall.

Prints one JSON line.  The CPU baseline is the in-repo FP64 oracle (dense
element operators, OpenMP over elements) timed on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "local-dg_amd"))

METRIC = "DOF-stage updates/s: mixed vs FP64 plain at stated error; -dEk/dt err"
HBM_PEAK_GBS = 8000.0


def cpu_baseline(p, ne, steps):
    """Oracle (FP64, dense operators) on a bounded sample: `steps` RK4 steps
    of the Taylor-Green state on an ne^3 mesh at order p."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    o = O.Oracle(p, ne)
    u = o.initial_condition()
    dt = o.dt(0.01)
    o.step(u, dt, 1, record_ke=False)     # warm-up (first-touch, thread pool)
    t = time.perf_counter()
    o.step(u, dt, steps, record_ke=True)
    wall = time.perf_counter() - t
    dof = o.K * o.dofs * 5
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": 4 * steps * dof / wall, "unit": "DOF-stage updates/s", "cores": cores, "kind": "port",
            "sample": f"oracle FP64 dense, Taylor-Green {ne}^3 elements p={p}, {steps} RK4 steps after 1 warm-up, "
                      f"{wall:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--p", type=int, default=4)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-nx", type=int, default=32)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()

    if args.calibrate_copy:
        import torch
        n = args.calibrate_copy * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 8, "copy_bytes_written": n * 8, "dispatches": 3}))
        return

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU halo exchange is not built", file=sys.stderr)
        sys.exit(2)

    import ldg_amd
    prec = ldg_amd.LDG_FP64 if args.precision == "fp64" else ldg_amd.LDG_MIXED
    p, nx = args.p, args.nx
    g = ldg_amd.LDG(p, nx, precision=prec)
    g.initial_condition()
    dt = g.time_step(0.01)
    dof = g.K * g.dofs * 5

    g.step(dt, args.warmup, record_ke=True)
    t0 = time.perf_counter()
    ke = g.step(dt, args.steps, record_ke=True)      # synchronises on return
    wall = time.perf_counter() - t0
    value = 4 * args.steps * dof / wall

    # per-kernel device time (HIP events on the library's stream), 2 steps
    kt = g.kernel_times(dt, 2)
    # algorithmic HBM bytes per launch, FP64 words per DOF:
    #   grad: read U_s, write Q (3)                                   = 4
    #   rhs : read U_s, Q (3), u^n; acc read+write; write U_{s+1};
    #         averaged over the 4 stages (stage 1 does not read acc,
    #         stage 4 writes no next input): (6 + 8 + 8 + 7) / 4       = 7.25
    bytes_grad = 4 * 8 * dof
    bytes_rhs = 7.25 * 8 * dof
    ach_rhs = bytes_rhs / (kt["rhs_ms"] * 1e-3) / 1e9
    ach_grad = bytes_grad / (kt["grad_ms"] * 1e-3) / 1e9
    diag = g.diagnostics()

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f64 state / f32 flux",
        "data": "synthetic: compressible Taylor-Green initial state (Re 1600, Mach 0.1), no files",
        "config": {"workload": f"LDG RHS + RK4, Taylor-Green {nx}^3 elements p={p} ({dof} DOF), "
                               f"dt = 0.01 dx/(2p+1)", "elements": g.K, "p": p, "dof": dof,
                   "precision": args.precision},
        "roofline": {"bound": "hbm", "achieved": ach_rhs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach_rhs / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "ldg_rhs_kernel", "avg_ms": kt["rhs_ms"]},
        "kernels": {"ldg_grad_kernel": {"avg_ms": kt["grad_ms"], "alg_GBs": ach_grad},
                    "ldg_rhs_kernel": {"avg_ms": kt["rhs_ms"], "alg_GBs": ach_rhs}},
        "ke_last": float(ke[-1]) if len(ke) else None,
        "diagnostics": diag,
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, args.cpu_nx, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
