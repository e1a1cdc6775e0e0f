#!/usr/bin/env python3
"""Throughput of the LDG right-hand side + RK4 stage pipeline on one MI355X.

A step is one RK4 step (4 stages: 8 kernel launches plus the kinetic-energy
reduction) of the compressible Taylor-Green vortex at 64^3 elements, p=4
(163,840,000 DOF), with the state resident in HBM.  The metric is DOF-stage
updates per second: 4 * steps * K * (p+1)^3 * 5 / wall.

Both paths are timed on the same workload: the mixed path (FP32
interpolation / flux / contractions, FP64 state, RK4, inverse mass and
reductions) is `value`; the FP64 plain path is reported beside it with the
speed-up, and the mixed state's relative L2 error against the FP64 state
after the timed steps (both start from the same initial state) together with
the error of the -dEk/dt series.  `--validate T` instead runs both paths to
time T and prints the errors only.

The CPU baseline is the in-repo FP64 oracle (dense element operators, OpenMP
over elements) timed on a bounded sample.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "local-dg_amd"))

METRIC = "DOF-stage updates/s: mixed vs FP64 plain at stated error; -dEk/dt err"
HBM_PEAK_GBS = 8000.0


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or what platform reports."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def cpu_baseline(p, ne, steps, ne1, steps1):
    """Oracle (FP64, dense operators) on bounded samples of the workload:
    `steps` RK4 steps of the Taylor-Green state on an ne^3 mesh at order p
    with all OpenMP threads, and `steps1` steps on an ne1^3 mesh with one."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    lib = O.lib()
    threads = lib.orc_max_threads()

    def timed(ne_, steps_):
        o = O.Oracle(p, ne_)
        u = o.initial_condition()
        dt = o.dt(0.01)
        o.step(u, dt, 1, record_ke=False)     # warm-up (first-touch, thread pool)
        t = time.perf_counter()
        o.step(u, dt, steps_, record_ke=True)
        wall = time.perf_counter() - t
        return 4 * steps_ * o.K * o.dofs * 5 / wall, wall

    value, wall = timed(ne, steps)
    lib.orc_set_threads(1)
    try:
        value1, wall1 = timed(ne1, steps1)
    finally:
        lib.orc_set_threads(threads)
    return {"value": value, "unit": "DOF-stage updates/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 dense, Taylor-Green {ne}^3 elements p={p}, {steps} RK4 steps after 1 warm-up, "
                      f"{threads} OpenMP threads, {wall:.1f} s",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "one_thread": {"value": value1, "unit": "DOF-stage updates/s",
                           "sample": f"Taylor-Green {ne1}^3 elements p={p}, {steps1} RK4 step(s), 1 thread, "
                                     f"{wall1:.1f} s"}}


def ff_bytes_per_dof(p, mixed):
    """Face-flux array bytes per DOF: per element 3 directions x 2 sides x
    5 states x (p+1)^2 face nodes."""
    n, w = p + 1, (4 if mixed else 8)
    return 6 * 5 * n * n * w / (5 * n ** 3)


def alg_bytes(dof, mixed, p=4):
    """Algorithmic HBM bytes per launch of this design, averaged over the 4
    RK stages of a step in a long run: every array a kernel must read or
    write, counted once at its storage precision.  u^n and the running sum
    are 8-byte words; a stage input U_s is 8 bytes (FP64) or a 4-byte
    fluctuation (mixed: stage 1 reads the copy of u^n that the previous
    step's stage 4 wrote); FF is the face-flux array (ff_bytes_per_dof).
    The gradient never reaches HBM: the gradient kernel evaluates it on the
    faces, the rhs kernel at the volume Gauss points.
      grad: read U_s, write FF
      rhs : read U_s, FF and u^n; read (stages 2-4) and write the running
            sum; write U_{s+1} (stages 1-3) or, mixed, the FP32 copy of
            u^{n+1} (stage 4).  In the FP64 path stage 1's U_s is u^n
            itself, read once."""
    ff = ff_bytes_per_dof(p, mixed)
    us = 4 if mixed else 8
    grad = (us + ff) * dof
    s1 = (us + 8 if mixed else 8) + ff + 8 + us             # U_s, (u^n), FF, acc write, U_2
    s23 = us + ff + 8 + 16 + us                             # U_s, FF, u^n, acc read+write, U_{s+1}
    s4 = us + ff + 8 + 16 + (4 if mixed else 0)             # U_4, FF, u^n, acc read+write, (copy of u^{n+1})
    rhs = (s1 + 2 * s23 + s4) / 4 * dof
    return grad, rhs


def survey_bytes_per_dof_stage(mixed):
    """The two-kernel model's compulsory bytes per DOF-stage in which the
    nodal gradient Q makes an HBM round trip (read U_s and write Q; read
    U_s, Q, u^n, read+write the running sum, write U_{s+1}): 12 FP64 words
    = 96 B, or 72 B with Q in FP32."""
    return 72.0 if mixed else 96.0


# the device code the committed rocprofv3 figures were measured on
KERNEL_SOURCES = ("local-dg_amd/csrc/ldg_kernels.hpp", "local-dg_amd/csrc/symop.hpp")


def kernel_source_hash():
    """First 12 hex digits of the SHA-256 of the kernel sources."""
    import hashlib
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:12]


PROFILE = os.path.join(ROOT, "profiles", "ldg_kernels_tgv64_p4.md")


def profiled_value(key, kernel, precision, p, nx, path=None):
    """A per-launch figure of `kernel` from the committed rocprofv3 summary
    (machine-readable lines "key kernel precision value src=HASH"):
    "traffic" = HBM bytes (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, the
    copy-kernel calibration), "valu_insts" = wave64 vector instructions
    (SQ_INSTS_VALU); averaged over the 4 RK stages.  None unless the
    workload is the 64^3 p=4 one they were measured on and HASH is the
    current kernel_source_hash() (a figure of other kernel code is stale)."""
    path = path or PROFILE
    if (p, nx) != (4, 64) or not os.path.exists(path):
        return None
    cur = "src=" + kernel_source_hash()
    for line in open(path):
        f = line.split()
        if len(f) == 5 and f[0] == key and f[1] == kernel and f[2] == precision and f[4] == cur:
            return float(f[3])
    return None


def profiled_traffic(kernel, precision, p, nx):
    return profiled_value("traffic", kernel, precision, p, nx)


# MI355X: 256 CUs x 4 SIMDs; a wave64 vector instruction occupies a SIMD for
# 4 cycles (16 lanes per cycle, no packed FP32); 2.4 GHz peak engine clock.
SIMDS, VALU_CYCLES, CLOCK_HZ = 1024, 4, 2.4e9


def issue_bound(kernel, precision, p, nx, avg_ms):
    """Fraction of the vector-issue ceiling the kernel uses: its profiled
    vector instructions per launch x 4 cycles / (1024 SIMDs x peak clock x
    its measured launch time)."""
    v = profiled_value("valu_insts", kernel, precision, p, nx)
    if v is None or not avg_ms:
        return None
    return {"kernel": f"ldg_{kernel}_kernel", "valu_insts_per_launch": v, "cycles_per_inst": VALU_CYCLES,
            "simds": SIMDS, "clock_ghz": CLOCK_HZ / 1e9, "avg_ms": avg_ms,
            "busy_frac": v * VALU_CYCLES / (SIMDS * CLOCK_HZ * avg_ms * 1e-3),
            "source": "profiles/ldg_kernels_tgv64_p4.md (rocprofv3 SQ_INSTS_VALU)"}


def rel_l2(a, b):
    return (np.sqrt(np.sum((a - b) ** 2, axis=(0, 2))) / np.sqrt(np.sum(b ** 2, axis=(0, 2)))).tolist()


def rel_l2_momentum(a, b):
    """Relative L2 error of the momentum vector (rho u, rho v, rho w) as a whole."""
    return float(np.sqrt(np.sum((a[:, 1:4] - b[:, 1:4]) ** 2)) / np.sqrt(np.sum(b[:, 1:4] ** 2)))


def dissipation(ke, dt):
    ke = np.asarray(ke)
    return -(ke[2:] - ke[:-2]) / (2 * dt)


def dissipation_err(ke, ke_ref, dt):
    """max |(-dEk/dt) - (-dEk/dt)_ref| / max |(-dEk/dt)_ref| (None for < 3 steps)."""
    d, d_ref = dissipation(ke, dt), dissipation(ke_ref, dt)
    if d_ref.size == 0:
        return None
    return float(np.max(np.abs(d - d_ref)) / np.max(np.abs(d_ref)))


def timed_run(ldg_amd, p, nx, prec, steps, warmup, autotune=True):
    """Both paths are run alike: tilings autotuned on the initial state
    (each path gets its fastest elements-per-workgroup choices), warm-up
    steps, then `steps` timed RK4 steps in one call."""
    g = ldg_amd.LDG(p, nx, precision=prec)
    g.initial_condition()
    tiling = g.autotune() if autotune else None
    dt = g.time_step(0.01)
    ke_w = g.step(dt, warmup, record_ke=True)
    t0 = time.perf_counter()
    ke = g.step(dt, steps, record_ke=True)          # synchronises on return
    wall = time.perf_counter() - t0
    return g, dt, wall, np.concatenate([ke_w, ke]), tiling


def ensemble_bench(ldg_amd, args):
    """Seeded ensemble of uneven Taylor-Green instances (ldg_ensemble_sample):
    mixed precision, tilings autotuned, timed with eager launches and with
    graph replay over the same steps."""
    inst = ldg_amd.sample_instances(args.seed, args.ensemble)
    out = {"metric": "ensemble DOF-stage updates/s", "unit": "DOF-stage updates/s", "seed": args.seed,
           "instances": len(inst), "steps": args.steps, "warmup": args.warmup}
    ens = ldg_amd.Ensemble(inst, precision=ldg_amd.LDG_MIXED)
    dof = ens.total_dof()
    out["dof"] = dof
    out["orders"] = {str(q): sum(1 for d in inst if d["p"] == q) for q in (2, 3, 4)}
    ens.initial_condition()
    out["autotune"] = ens.autotune()
    dts = ens.time_steps(0.01)
    for mode in ("eager", "graph"):
        ens.set_graph(mode == "graph")
        ens.step(dts, args.warmup + (args.warmup % 2))
        t0 = time.perf_counter()
        ke = ens.step(dts, args.steps)
        wall = time.perf_counter() - t0
        out[mode] = {"value": 4 * args.steps * dof / wall, "ms_per_step": wall / args.steps * 1e3}
        assert np.all(np.isfinite(ke))
    out["value"] = out["graph"]["value"]
    out["graph_speedup"] = out["graph"]["value"] / out["eager"]["value"]
    ens.close()
    return out


SAMPLE_ELEMENTS = 4096   # elements kept in a validation fingerprint (seeded choice)


def validate_part(ldg_amd, p, nx, t_end, name, out):
    """One path from the Taylor-Green state to t_end (reference dt rule,
    rounded to end exactly at t_end); saves the KE history, the Gauss
    diagnostics, per-variable L2 norms of the final state and the final
    state on a seeded sample of elements, so that two runs too long for one
    process can be compared afterwards (validate_compare)."""
    prec = ldg_amd.LDG_FP64 if name == "fp64" else ldg_amd.LDG_MIXED
    g = ldg_amd.LDG(p, nx, precision=prec)
    g.initial_condition()
    dt0 = g.time_step(0.01)
    n = int(np.ceil(t_end / dt0))
    dt = t_end / n
    kes, done, t0 = [], 0, time.perf_counter()
    while done < n:
        c = min(500, n - done)
        kes.append(g.step(dt, c, record_ke=True))
        done += c
        print(f"{name}: {done}/{n} steps", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    u = g.get_state()
    d = g.diagnostics()
    g.close()
    # u is (K, 5, dofs): elements are the first axis
    idx = np.sort(np.random.default_rng(20261203).choice(u.shape[0], size=min(SAMPLE_ELEMENTS, u.shape[0]), replace=False))
    np.savez_compressed(out, ke=np.concatenate(kes), dt=dt, steps=n, t_end=t_end, wall=wall, idx=idx,
                        sample=u[idx], norms=np.sqrt(np.sum(u ** 2, axis=(0, 2))),
                        enstrophy=d["enstrophy"], ke_gauss=d["ke"], p=p, nx=nx)
    print(json.dumps({"validate_part": name, "t_end": t_end, "steps": n, "dt": dt, "wall_s": wall,
                      "ms_per_step": wall / n * 1e3, "out": out}), flush=True)


def validate_compare(mixed_npz, fp64_npz):
    """Errors of a mixed-precision fingerprint against an FP64 one of the
    same run: relative L2 per variable on the sampled elements, the KE
    history and the -dEk/dt series."""
    a, b = np.load(mixed_npz), np.load(fp64_npz)
    assert int(a["steps"]) == int(b["steps"]) and np.array_equal(a["idx"], b["idx"]), "different runs"
    dt = float(b["dt"])
    return {"t_end": float(b["t_end"]), "steps": int(b["steps"]), "dt": dt,
            "sample_elements": int(len(b["idx"])),
            "rel_l2_mixed_vs_fp64_sampled": rel_l2(a["sample"], b["sample"]),
            "norm_rel_diff": (np.abs(a["norms"] - b["norms"]) / b["norms"]).tolist(),
            "ke_rel_err": float(np.max(np.abs(a["ke"] - b["ke"]) / b["ke"])),
            "dEk_dt_err": dissipation_err(a["ke"], b["ke"], dt),
            "enstrophy_fp64": float(b["enstrophy"]), "enstrophy_mixed": float(a["enstrophy"]),
            "wall_s": {"fp64": float(b["wall"]), "mixed": float(a["wall"])}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--p", type=int, default=4)
    ap.add_argument("--precision", choices=["mixed", "fp64"], default="mixed",
                    help="path reported as `value` (both are timed)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-autotune", action="store_true", help="default tilings instead of autotuned ones")
    ap.add_argument("--cpu-nx", type=int, default=32)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--cpu1-nx", type=int, default=12, help="mesh of the one-thread CPU sample")
    ap.add_argument("--cpu1-steps", type=int, default=1)
    ap.add_argument("--validate", type=float, default=0.0, metavar="T",
                    help="run both paths to time T and print their difference only")
    ap.add_argument("--validate-part", choices=["fp64", "mixed"], default=None,
                    help="with --validate T: run only this path to T and save its fingerprint to --out "
                         "(KE history, enstrophy, norms, the state on a seeded sample of 4096 elements)")
    ap.add_argument("--out", default=None, help="fingerprint file (.npz) for --validate-part")
    ap.add_argument("--validate-compare", nargs=2, default=None, metavar=("MIXED_NPZ", "FP64_NPZ"),
                    help="compare two --validate-part fingerprints and print the errors")
    ap.add_argument("--ensemble", type=int, default=0, metavar="N",
                    help="instead: an N-instance seeded ensemble (mixed precision, autotuned, graph vs eager)")
    ap.add_argument("--seed", type=int, default=0, help="ensemble sampler seed")
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

    if args.validate_compare:
        print(json.dumps(validate_compare(*args.validate_compare)), flush=True)
        return

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU halo exchange is not built", file=sys.stderr)
        sys.exit(2)

    import ldg_amd
    p, nx = args.p, args.nx

    if args.ensemble:
        print(json.dumps(ensemble_bench(ldg_amd, args)), flush=True)
        return

    if args.validate > 0 and args.validate_part:
        validate_part(ldg_amd, p, nx, args.validate, args.validate_part, args.out)
        return

    if args.validate > 0:
        out = {"validate_t": args.validate, "workload": f"Taylor-Green {nx}^3 elements p={p}"}
        res = {}
        for name, prec in (("fp64", ldg_amd.LDG_FP64), ("mixed", ldg_amd.LDG_MIXED)):
            g = ldg_amd.LDG(p, nx, precision=prec)
            g.initial_condition()
            dt0 = g.time_step(0.01)
            n = int(np.ceil(args.validate / dt0))
            dt = args.validate / n
            kes, done, t0 = [], 0, time.perf_counter()
            while done < n:
                c = min(500, n - done)
                kes.append(g.step(dt, c, record_ke=True))
                done += c
                print(f"{name}: {done}/{n} steps", file=sys.stderr, flush=True)
            res[name] = (g.get_state(), np.concatenate(kes), dt, time.perf_counter() - t0, g.diagnostics())
            g.close()
        u64, ke64, dt, w64, d64 = res["fp64"]
        u32, ke32, _, w32, d32 = res["mixed"]
        out.update({"steps": int(len(ke64)), "dt": dt, "rel_l2_mixed_vs_fp64": rel_l2(u32, u64),
                    "ke_rel_err": float(np.max(np.abs(ke32 - ke64) / ke64)),
                    "dEk_dt_err": dissipation_err(ke32, ke64, dt),
                    "enstrophy_fp64": d64["enstrophy"], "enstrophy_mixed": d32["enstrophy"],
                    "wall_s": {"fp64": w64, "mixed": w32}})
        print(json.dumps(out), flush=True)
        return

    runs = {}
    for name, prec in (("fp64", ldg_amd.LDG_FP64), ("mixed", ldg_amd.LDG_MIXED)):
        g, dt, wall, ke, tiling = timed_run(ldg_amd, p, nx, prec, args.steps, args.warmup, not args.no_autotune)
        dof = g.K * g.dofs * 5
        u = g.get_state()
        kt = g.kernel_times(dt, 4)          # per-kernel device time, HIP events on the library's stream
        runs[name] = dict(K=g.K, dt=dt, wall=wall, ke=ke, dof=dof, kt=kt, u=u, tiling=tiling)
        g.close()

    main_name = args.precision
    r = runs[main_name]
    dof = r["dof"]
    value = 4 * args.steps * dof / r["wall"]
    fp64_value = 4 * args.steps * dof / runs["fp64"]["wall"]
    mixed = main_name == "mixed"
    b_grad, b_rhs = alg_bytes(dof, mixed, p)
    ach_rhs = b_rhs / (r["kt"]["rhs_ms"] * 1e-3) / 1e9
    stage_ms = r["kt"]["grad_ms"] + r["kt"]["rhs_ms"]
    err = rel_l2(runs["mixed"]["u"], runs["fp64"]["u"])

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "DOF-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 flux/quadrature + f64 state, RK4, inverse mass, reductions" if mixed else "f64",
        "data": "synthetic: compressible Taylor-Green initial state (Re 1600, Mach 0.1), no files",
        "config": {"workload": f"LDG RHS + RK4, Taylor-Green {nx}^3 elements p={p} ({dof} DOF), "
                               f"dt = 0.01 dx/(2p+1)", "elements": r["K"], "p": p, "dof": dof,
                   "precision": main_name},
        "fp64_value": fp64_value,
        "speedup_vs_fp64": value / fp64_value,
        "error_vs_fp64": {"after_steps": args.warmup + args.steps, "rel_l2_per_variable": err,
                          "rel_l2_momentum": rel_l2_momentum(runs["mixed"]["u"], runs["fp64"]["u"]),
                          "dEk_dt_err": dissipation_err(runs["mixed"]["ke"], runs["fp64"]["ke"], r["dt"]),
                          "note": "rho w is zero in the Taylor-Green initial state, so after a few steps its "
                                  "own norm is tiny; rel_l2_momentum is relative to the whole momentum. "
                                  "Error at t = 1 (bar 1e-5): profiles/mixed_vs_fp64_error_tgv64_p4_t1.md"},
        "roofline": {"bound": "hbm", "achieved": ach_rhs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": ach_rhs / HBM_PEAK_GBS, "traffic": profiled_traffic("rhs", main_name, p, nx),
                     "traffic_source": "profiles/ldg_kernels_tgv64_p4.md (rocprofv3 FETCH_SIZE x2.0 + WRITE_SIZE x1.0)",
                     "kernel": "ldg_rhs_kernel", "alg_bytes_per_launch": b_rhs, "avg_ms": r["kt"]["rhs_ms"],
                     "two_kernel_model": {"bytes_per_dof_stage": survey_bytes_per_dof_stage(mixed),
                                          "effective_GBs": survey_bytes_per_dof_stage(mixed) * dof
                                          / (stage_ms * 1e-3) / 1e9,
                                          "note": "the bytes a stage would move with the nodal gradient in "
                                                  "HBM, over the measured grad + rhs time of a stage"}},
        "issue_bound": issue_bound("rhs", main_name, p, nx, r["kt"]["rhs_ms"]),
        "kernels": {
            name: {"ldg_grad_kernel": {"avg_ms": v["kt"]["grad_ms"],
                                       "alg_GBs": alg_bytes(dof, name == "mixed", p)[0] / (v["kt"]["grad_ms"] * 1e-3) / 1e9},
                   "ldg_rhs_kernel": {"avg_ms": v["kt"]["rhs_ms"],
                                      "alg_GBs": alg_bytes(dof, name == "mixed", p)[1] / (v["kt"]["rhs_ms"] * 1e-3) / 1e9},
                   "ms_per_step": v["wall"] / args.steps * 1e3,
                   "tiling": {"p": v["tiling"][0][0], "grad_elements_per_workgroup": v["tiling"][0][1],
                              "rhs_elements_per_workgroup": v["tiling"][0][2]} if v["tiling"] else "default"}
            for name, v in runs.items()},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, args.cpu_nx, args.cpu_steps, args.cpu1_nx, args.cpu1_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
