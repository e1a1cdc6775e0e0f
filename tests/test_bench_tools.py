"""Host-side helpers of bench.py: the algorithmic byte model and the split
mixed-vs-FP64 validation (two runs saved as fingerprints, compared later)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_alg_bytes_per_dof():
    # p = 4: an element has 625 DOF and 6 face-flux chunks of 5 x 25 values
    ff64, ff32 = 6 * 125 * 8 / 625, 6 * 125 * 4 / 625
    assert abs(bench.ff_bytes_per_dof(4, False) - ff64) < 1e-12 and abs(bench.ff_bytes_per_dof(4, True) - ff32) < 1e-12
    # FP64: grad U 8 + FF; rhs stages (u^n once, acc write, U_2), 2 x (U, u^n,
    # acc r+w, U_next), (U, u^n, acc r+w), each + FF
    g, r = bench.alg_bytes(1, False)
    assert abs(g - (8 + ff64)) < 1e-12 and abs(r - ((24 + 2 * 40 + 32) / 4 + ff64)) < 1e-12
    # mixed: 4-byte stage inputs; stage 1 reads u^n beside its FP32 copy,
    # stage 4 writes the FP32 copy of u^{n+1}
    g, r = bench.alg_bytes(1, True)
    assert abs(g - (4 + ff32)) < 1e-12 and abs(r - ((24 + 2 * 32 + 32) / 4 + ff32)) < 1e-12
    assert bench.survey_bytes_per_dof_stage(False) == 96.0 and bench.survey_bytes_per_dof_stage(True) == 72.0


def _fingerprint(path, sample, ke, norms, dt=1e-3, steps=None, idx=None):
    steps = len(ke) - 1 if steps is None else steps
    idx = np.arange(sample.shape[0]) if idx is None else idx
    np.savez_compressed(path, ke=ke, dt=dt, steps=steps, t_end=dt * steps, wall=1.0, idx=idx, sample=sample,
                        norms=norms, enstrophy=0.5, ke_gauss=0.1, p=4, nx=8)


def test_validate_compare_matches_direct_errors(tmp_path):
    rng = np.random.default_rng(1)
    ref = rng.standard_normal((16, 5, 125)) + 3.0     # (elements, variables, nodes)
    mix = ref.copy()
    mix[:, 1, :] += 1e-6 * rng.standard_normal((16, 125))   # only rho u differs
    ke = 0.125 - 1e-3 * np.arange(11) ** 2 * 1e-2
    ke_m = ke * (1 + 1e-7 * np.sin(np.arange(11)))
    _fingerprint(tmp_path / "f.npz", ref, ke, np.ones(5))
    _fingerprint(tmp_path / "m.npz", mix, ke_m, np.ones(5) * (1 + 1e-9))
    out = bench.validate_compare(str(tmp_path / "m.npz"), str(tmp_path / "f.npz"))
    err = out["rel_l2_mixed_vs_fp64_sampled"]
    direct = np.linalg.norm(mix[:, 1, :] - ref[:, 1, :]) / np.linalg.norm(ref[:, 1, :])
    assert err[0] == 0.0 and err[2] == 0.0 and err[3] == 0.0 and err[4] == 0.0
    assert abs(err[1] - direct) <= 1e-12 * direct
    assert abs(out["ke_rel_err"] - np.max(np.abs(ke_m - ke) / ke)) < 1e-18
    assert out["steps"] == 10 and out["sample_elements"] == 16
    assert max(out["norm_rel_diff"]) < 2e-9


def test_validate_compare_refuses_different_runs(tmp_path):
    a = np.ones((4, 5, 8))
    _fingerprint(tmp_path / "a.npz", a, np.ones(5), np.ones(5))
    _fingerprint(tmp_path / "b.npz", a, np.ones(7), np.ones(5))
    try:
        bench.validate_compare(str(tmp_path / "a.npz"), str(tmp_path / "b.npz"))
    except AssertionError:
        return
    raise AssertionError("runs of different length were compared")


def test_profiled_figures_and_issue_bound(tmp_path, monkeypatch):
    """bench.py reads "key kernel precision value src=HASH" lines of the
    rocprofv3 summary and drops any whose HASH is not that of the current
    kernel sources."""
    cur = bench.kernel_source_hash()
    prof = tmp_path / "prof.md"
    prof.write_text("    traffic rhs mixed 8635138049 src=%s\n"
                    "    traffic rhs fp64 14481330586 src=000000000000\n"
                    "    valu_insts rhs mixed 1083500000 src=%s\n" % (cur, cur))
    monkeypatch.setattr(bench, "PROFILE", str(prof))
    t = bench.profiled_traffic("rhs", "mixed", 4, 64)
    assert t == 8635138049 and t > bench.alg_bytes(163840000, True)[1]     # measured >= algorithmic
    assert bench.profiled_traffic("rhs", "fp64", 4, 64) is None             # measured on other kernel code
    assert bench.profiled_traffic("rhs", "mixed", 4, 32) is None            # other workloads: not measured
    ib = bench.issue_bound("rhs", "mixed", 4, 64, 2.0)
    v = ib["valu_insts_per_launch"]
    assert v == 1083500000
    assert abs(ib["busy_frac"] - v * 4 / (1024 * 2.4e9 * 2.0e-3)) < 1e-12
    assert bench.issue_bound("rhs", "mixed", 4, 32, 2.0) is None


def test_momentum_error_is_joint_over_components():
    # only rho w differs, and rho w is small: its own relative error is large,
    # the error relative to the whole momentum is not
    rng = np.random.default_rng(2)
    ref = rng.standard_normal((8, 5, 27))
    ref[:, 3, :] *= 1e-4
    mix = ref.copy()
    mix[:, 3, :] += 1e-8 * rng.standard_normal((8, 27))
    per = bench.rel_l2(mix, ref)
    joint = bench.rel_l2_momentum(mix, ref)
    direct = np.linalg.norm(mix[:, 3, :] - ref[:, 3, :]) / np.linalg.norm(ref[:, 1:4, :])
    assert abs(joint - direct) <= 1e-12 * direct
    assert per[3] > 1e2 * joint
