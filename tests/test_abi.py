"""CPU-side checks of the C ABI: the library loads, exports every symbol
include/ldg.h declares, and rejects bad configurations before touching a
device (no compute calls here)."""
import ctypes
import os
import re

import pytest

import ldg_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "ldg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ldg_[a-z_]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert declared_symbols() == sorted(ldg_amd.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = ldg_amd.load_library()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_config_default_is_reference_defaults():
    lib = ldg_amd.load_library()
    cfg = ldg_amd.ldg_config()
    assert lib.ldg_config_default(ctypes.byref(cfg)) == 0
    assert (cfg.p, cfg.nx, cfg.ny, cfg.nz) == (2, 10, 10, 10)
    assert (cfg.Re, cfg.Pr, cfg.gamma, cfg.M0, cfg.rho0, cfg.V0, cfg.L, cfg.R) == (1600.0, 0.71, 1.4, 0.1, 1.0, 1.0, 1.0, 8.314)
    assert cfg.xmin[0] == pytest.approx(-3.141592653589793) and cfg.xmax[2] == pytest.approx(3.141592653589793)


@pytest.mark.parametrize("field,value", [("p", 0), ("p", 6), ("nx", 2), ("nz", 1), ("Re", 0.0), ("gamma", 1.0)])
def test_create_rejects_bad_config(field, value):
    lib = ldg_amd.load_library()
    cfg = ldg_amd.ldg_config()
    lib.ldg_config_default(ctypes.byref(cfg))
    setattr(cfg, field, value)
    h = ctypes.c_void_p()
    rc = lib.ldg_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == ldg_amd.LDG_EINVAL
    assert not h.value
    assert len(lib.ldg_last_error(None)) > 0


def test_null_arguments_are_rejected():
    lib = ldg_amd.load_library()
    assert lib.ldg_create(None, None) == ldg_amd.LDG_EINVAL
    assert lib.ldg_sizes(None, None, None) == ldg_amd.LDG_EINVAL
    assert lib.ldg_step(None, 0.1, 1, None) == ldg_amd.LDG_EINVAL


def test_mesh_rejects_thin_mesh():
    with pytest.raises(ValueError):
        ldg_amd.Mesh(2, 8, 8)


def test_ensemble_sampler_is_deterministic_and_in_range():
    """ldg_ensemble_sample is host code: same seed, same instances; the
    documented ranges hold; different seeds differ."""
    a = ldg_amd.sample_instances(11, 200)
    assert a == ldg_amd.sample_instances(11, 200)
    assert a != ldg_amd.sample_instances(12, 200)
    assert {d["p"] for d in a} == {2, 3, 4}
    assert {d["nx"] for d in a} == {8, 12, 16, 24}
    assert all(d["nx"] == d["ny"] == d["nz"] for d in a)
    assert all(400 <= d["Re"] <= 3200 and 0.05 <= d["M0"] <= 0.3 for d in a)
    assert all(0.7 < d["V0"] < 1.3 for d in a)
    assert all(0 <= ph < 2 * 3.141592653589794 for d in a for ph in d["phase"])
    v0 = [d["V0"] for d in a]
    assert abs(sum(v0) / len(v0) - 1) < 0.02


def test_ensemble_create_rejects_bad_instance():
    bad = [{"p": 2, "nx": 4, "Re": 1600.0, "M0": 0.1, "V0": 1.0}, {"p": 7, "nx": 4}]
    with pytest.raises(ldg_amd.LDGError):
        ldg_amd.Ensemble(bad)
    with pytest.raises(ldg_amd.LDGError):
        ldg_amd.Ensemble([{"p": 2, "nx": 2}])


def _hip_devices():
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        return 0
    n = ctypes.c_int(0)
    return n.value if hip.hipGetDeviceCount(ctypes.byref(n)) == 0 else 0


def test_noalloc_rejects_bad_arguments_before_any_device_work():
    """ldg_rhs_noalloc checks its arguments and the scratch limit on the host,
    before it looks for a device (so these run without a GPU; with one, the
    placeholder pointers below are not used)."""
    if _hip_devices() > 0:
        pytest.skip("a HIP device is visible: the GPU test checks these paths with device buffers")
    lib = ldg_amd.load_library()
    cfg = ldg_amd.ldg_config()
    lib.ldg_config_default(ctypes.byref(cfg))
    assert lib.ldg_rhs_noalloc(None, None, None, None) == ldg_amd.LDG_EINVAL
    cfg.p, cfg.nx, cfg.ny, cfg.nz = 4, 23, 23, 23          # beyond LDG_NOALLOC_SCRATCH_BYTES in FP64
    dummy = ctypes.c_void_p(8)                              # never dereferenced: refused first
    assert lib.ldg_rhs_noalloc(ctypes.byref(cfg), dummy, dummy, None) == ldg_amd.LDG_EINVAL
    assert b"LDG_NOALLOC_SCRATCH_BYTES" in lib.ldg_last_error(None)
    cfg.nx = 2                                              # thinner than the reference allows
    assert lib.ldg_rhs_noalloc(ctypes.byref(cfg), dummy, dummy, None) == ldg_amd.LDG_EINVAL
