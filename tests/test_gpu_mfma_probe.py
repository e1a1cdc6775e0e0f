"""The matrix-core measurement's three forms of the y/z flux contraction
(ldg_amd.mfma_probe, libldg_probe.so) against an FP64 evaluation of the
same contraction on the CPU.  Bar: 5e-7 relative L2 -- FP32 accumulation of
25-50 terms (~1e-7), and for the FP16 hi/lo form the split's 2^-22."""
import pytest

from ldg_amd import mfma_probe as mp

pytestmark = pytest.mark.gpu


def test_probe_variants_match_fp64():
    out = mp.measure(elements=4096, reps=3, samples=256)
    for name, v in out["variants"].items():
        print(f"{name}: rel L2 {v['rel_l2_vs_fp64']:.2e}", flush=True)
        assert v["rel_l2_vs_fp64"] <= 5e-7, (name, v["rel_l2_vs_fp64"])
