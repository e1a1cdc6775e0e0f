"""The committed rocprofv3 summary carries per-launch figures for the
kernel sources in the tree: each machine-readable line is tagged with the
SHA-256 prefix of the kernel sources it was measured on, and bench.py uses
only lines whose tag matches the current sources (so a kernel change without
a new profile shows up here, not as silently stale figures)."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_profile_figures_are_for_the_current_kernels():
    cur = bench.kernel_source_hash()
    for key in ("traffic", "valu_insts"):
        for kernel in ("rhs", "grad"):
            for prec in ("mixed", "fp64"):
                v = bench.profiled_value(key, kernel, prec, 4, 64)
                assert v is not None and v > 0, (key, kernel, prec, cur)
    # measured traffic is at least the algorithmic bytes of the design
    dof = 64 ** 3 * 125 * 5
    assert bench.profiled_traffic("rhs", "mixed", 4, 64) >= bench.alg_bytes(dof, True)[1]
    assert bench.profiled_traffic("rhs", "fp64", 4, 64) >= bench.alg_bytes(dof, False)[1]
