"""ldg_rhs_noalloc (the entry point that allocates no device memory) against
ldg_rhs on the same inputs, through the C ABI: bitwise equal (the same
kernels at the same tilings), on two calls in a row with the second on a
smaller mesh, in FP64 and in the mixed path; meshes beyond the fixed scratch
are refused.  Device buffers come from hipMalloc in the test itself."""
import ctypes
import os

import numpy as np
import pytest

import ldg_amd
import oracle as O

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    ldg_amd.load_library()          # loads the HIP runtime libldg.so links
    for name in ("libamdhip64.so", "/opt/rocm/lib/libamdhip64.so"):
        try:
            L = ctypes.CDLL(name)
            break
        except OSError:
            continue
    else:
        pytest.fail("HIP runtime library not found")
    L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    L.hipFree.argtypes = [ctypes.c_void_p]
    for f in (L.hipMalloc, L.hipMemcpy, L.hipFree):
        f.restype = ctypes.c_int
    return L


def _state(p, nx, seed):
    o = O.Oracle(p, nx)
    u = o.initial_condition()
    u[:, 3] += 0.01 * np.random.default_rng(seed).standard_normal(u[:, 3].shape)
    return u


def _noalloc_rhs(hip, cfg, u):
    """ldg_rhs_noalloc on device copies of u; returns k on the host."""
    nbytes = u.nbytes
    du, dk = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(du), nbytes) == 0 and hip.hipMalloc(ctypes.byref(dk), nbytes) == 0
    try:
        assert hip.hipMemcpy(du, u.ctypes.data, nbytes, H2D) == 0
        ldg_amd.rhs_noalloc(cfg, du.value, dk.value)
        k = np.empty_like(u)
        assert hip.hipMemcpy(k.ctypes.data, dk, nbytes, D2H) == 0
        return k
    finally:
        hip.hipFree(du)
        hip.hipFree(dk)


@pytest.mark.parametrize("precision", [ldg_amd.LDG_FP64, ldg_amd.LDG_MIXED])
def test_noalloc_rhs_equals_ldg_rhs(precision):
    hip = _hip()
    for p, nx in ((4, 6), (2, 4)):       # the second call is smaller than the first
        print(f"ldg_rhs_noalloc p={p} {nx}^3 precision={precision}", flush=True)
        g = ldg_amd.LDG(p, nx, precision=precision)
        u = _state(p, nx, 10 * p + nx)
        k_ref = g.rhs(u)
        k = _noalloc_rhs(hip, g.cfg, u)
        assert np.array_equal(k, k_ref), np.max(np.abs(k - k_ref))
        g.close()


def test_noalloc_refuses_meshes_beyond_its_scratch():
    hip = _hip()
    g = ldg_amd.LDG(4, 3)
    cfg = g.cfg
    cfg.nx = cfg.ny = cfg.nz = 23            # 8192 + 23^3 x 30 x 25 x 8 bytes > 64 MiB
    nbytes = 23 ** 3 * 5 * 125 * 8            # full-size device arrays, although the call is refused first
    du, dk = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(du), nbytes) == 0 and hip.hipMalloc(ctypes.byref(dk), nbytes) == 0
    try:
        with pytest.raises(ldg_amd.LDGError) as e:
            ldg_amd.rhs_noalloc(cfg, du.value, dk.value)
    finally:
        hip.hipFree(du)
        hip.hipFree(dk)
    assert e.value.code == ldg_amd.LDG_EINVAL
    assert "LDG_NOALLOC_SCRATCH_BYTES" in str(e.value)
    # the largest cubic FP64 mesh at p = 4 that fits: 22^3
    cfg.nx = cfg.ny = cfg.nz = 22
    u = _state(4, 22, 3)
    k = _noalloc_rhs(hip, cfg, u)
    g22 = ldg_amd.LDG(4, 22)
    assert np.array_equal(k, g22.rhs(u))
