"""Builds of the oracle source (oracle/ldg_oracle.c): the tiled dense
kernels against plain column-by-column loops (bitwise), and the x87 long
double build that serves as the exact answer for the FP64 parity bars."""
import ctypes
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "..", "oracle", "ldg_oracle.c")


def rel_l2(a, b):
    return np.sqrt(np.sum((a - b) ** 2, axis=(0, 2))) / np.sqrt(np.sum(b ** 2, axis=(0, 2)))


@pytest.fixture(scope="module")
def untiled():
    """The oracle built with -DORC_UNTILED (every product and triangular
    solve as one column at a time), under the name "untiled"."""
    d = tempfile.mkdtemp()
    path = os.path.join(d, "libldg_oracle_untiled.so")
    subprocess.run(["gcc", "-O2", "-std=c99", "-fPIC", "-fopenmp", "-D_DEFAULT_SOURCE", "-DORC_UNTILED",
                    "-ffp-contract=off", "-shared", "-o", path, SRC, "-lm"], check=True)
    O._LIB_PATHS["untiled"] = path
    yield "untiled"
    O._libs.pop("untiled", None)
    shutil.rmtree(d, ignore_errors=True)


@pytest.mark.parametrize("p,nx,ny,nz", [(1, 3, 3, 3), (3, 3, 4, 5), (4, 4, 4, 4), (5, 3, 3, 3)])
def test_tiled_kernels_bitwise_equal_untiled(untiled, p, nx, ny, nz):
    a, b = O.Oracle(p, nx, ny, nz), O.Oracle(p, nx, ny, nz, variant=untiled)
    for name in ("Interp3D", "Mel", "Sels", "Kels"):
        assert np.array_equal(a.operator(name), b.operator(name)), name
    u = a.initial_condition()
    u[:, 3] += 1e-2 * np.random.default_rng(p).standard_normal(u[:, 3].shape)
    ka, da = a.rhs(u, want_du=True)
    kb, db = b.rhs(u, want_du=True)
    assert np.array_equal(ka, kb) and np.array_equal(da, db)
    ua, kea = a.step(u, a.dt(0.01), 2)
    ub, keb = b.step(u, b.dt(0.01), 2)
    assert np.array_equal(ua, ub) and np.array_equal(kea, keb)
    assert np.array_equal(a.diagnostics(ua), b.diagnostics(ub))


def test_build_widths():
    assert O.lib("plain").orc_real_size() == 8
    assert O.lib("ld").orc_real_size() in (12, 16)   # x87 80-bit, padded


def test_long_double_build_carries_more_precision():
    """A constant state is an exact steady state.  FP64 rounding leaves a
    right-hand side of order eps * E0 * (condition of the dense operators);
    the long double build leaves it 2^11 times smaller, which it can only do
    if every operator and flux is carried in the wider type."""
    o, ol = O.Oracle(3, 3), O.Oracle(3, 3, variant="ld")
    u = np.zeros(o.shape)
    u[:, 0], u[:, 1], u[:, 4] = 1.0, 0.3, 178.6
    k64, kld = np.abs(o.rhs(u)).max(), np.abs(ol.rhs(u)).max()
    print("constant-state residual: FP64", k64, "long double", kld)
    assert k64 > 1e-14
    assert kld < 1e-3 * k64


def test_long_double_single_rhs_is_the_centre_of_fp64_roundings():
    """The plain and fused-multiply-add FP64 builds round differently; the
    exact answer lies within their rounding noise of both, and a wrong long
    double build (a constant left in double, a dropped term) would sit far
    from both.  p = 2: FP64 errors ~1e-12..1e-11 at Mach 0.1."""
    o, of, ol = O.Oracle(2, 4), O.Oracle(2, 4, variant="fma"), O.Oracle(2, 4, variant="ld")
    u = o.initial_condition()
    u[:, 3] += 1e-2 * np.random.default_rng(7).standard_normal(u[:, 3].shape)
    kp, kf, kl = o.rhs(u), of.rhs(u), ol.rhs(u)
    spread = rel_l2(kp, kf)
    ep, ef = rel_l2(kp, kl), rel_l2(kf, kl)
    print("plain vs exact", ep, "fma vs exact", ef, "plain vs fma", spread)
    assert np.all(ep <= 3 * spread + 1e-15) and np.all(ef <= 3 * spread + 1e-15)
    assert np.all(ep <= 1e-10)


def test_fp64_oracle_error_on_c1_start():
    """Taylor-Green 8^3, p = 2, 20 steps at the reference time step: the FP64
    oracle's distance from the long double build, per conserved variable.
    rho w starts at zero, so its own norm is small and its relative error is
    the largest (2.6e-12 after 100 steps); the other variables stay below
    1e-13 relative."""
    p, nx, n = 2, 8, 20
    o, ol = O.Oracle(p, nx), O.Oracle(p, nx, variant="ld")
    u0 = o.initial_condition()
    dt = o.dt(0.01)
    up, kep = o.step(u0, dt, n)
    ul, kel = ol.step(u0, dt, n)
    e = rel_l2(up, ul)
    print("FP64 oracle vs long double after 20 steps:", e)
    assert np.all(e[[0, 1, 2, 4]] <= 1e-13)
    assert e[3] <= 1e-10
    assert np.max(np.abs(kep - kel) / kel) <= 1e-14
