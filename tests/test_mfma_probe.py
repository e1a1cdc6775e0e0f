"""CPU checks of the matrix-core measurement's operators (ldg_amd.mfma_probe):
its 1D basis against the polynomials it must reproduce, the slab index
convention against two separate 1D sweeps, the FP16 hi/lo split, and the
operator block's size against the library that reads it (no GPU calls)."""
import ctypes
import os

import numpy as np
import pytest

from ldg_amd import mfma_probe as mp


def test_basis_reproduces_polynomials():
    I, D = mp.basis(4)
    r = -np.cos(np.pi * np.arange(5) / 4)
    g, _ = np.polynomial.legendre.leggauss(5)
    for c in ([1.0], [0.3, -1.2], [0.5, 0.0, 2.0, -1.0, 0.7]):
        f = np.polynomial.polynomial.Polynomial(c)
        assert np.allclose(I @ f(r), f(g), atol=1e-13)
        assert np.allclose(D @ f(r), f.deriv()(g), atol=1e-12)


def test_symmetrised_operators_are_the_operators():
    I, D = mp.basis(4)
    It, Dt, _, _ = mp.operators()
    assert np.abs(It - I.T).max() < 1e-13
    assert np.abs(Dt - D.T).max() < 1e-12


def test_slab_operators_equal_two_sweeps():
    """Slab index i = y + 5 z: the dense operators equal a y sweep then a z
    sweep of the 1D operators."""
    It, Dt, Ax, Ag = mp.operators()
    rng = np.random.default_rng(1)
    fx, fy, fz = (rng.standard_normal((5, 5)) for _ in range(3))   # [z][y]
    ox = It @ fx @ It.T                       # z: rows, y: columns
    og = It @ fy @ Dt.T + Dt @ fz @ It.T      # D^T along y for F_y, along z for F_z
    assert np.allclose(Ax @ fx.ravel(), ox.ravel(), atol=1e-13)
    assert np.allclose(Ag @ np.concatenate([fy.ravel(), fz.ravel()]), og.ravel(), atol=1e-12)


@pytest.mark.parametrize("mag", [1e-6, 1.0, 3.6e5])
def test_scaled_hilo_split(mag):
    """The probe's FP16 operand: the column scaled by 2^(14 - x) (its maximum
    m = f 2^x, f in [0.5, 1)), then split into hi + lo.  The scaled column
    stays below 2^14 (no FP16 overflow at any magnitude, 3.6e5 being the
    weighted energy flux at Mach 0.01), the split's error is below 2^-22 of
    the column maximum, and values within 2^12 of the maximum keep 22 bits."""
    rng = np.random.default_rng(2)
    a = (rng.standard_normal(4096) * 2.0 ** rng.integers(-20, 1, 4096) * mag).astype(np.float32)
    m = np.abs(a).max()
    sh = 14 - np.frexp(m)[1]
    s = np.ldexp(a, sh).astype(np.float32)
    assert np.abs(s).max() < 2.0 ** 14
    h = s.astype(np.float16)
    lo = (s - h.astype(np.float32)).astype(np.float16)
    err = np.abs(h.astype(np.float64) + lo.astype(np.float64) - s.astype(np.float64))
    assert err.max() <= 2.0 ** -22 * np.abs(s).max()
    big = np.abs(s) >= 2.0 ** 2
    assert (err[big] / np.abs(s[big])).max() <= 2.0 ** -22


def test_operator_block_matches_library():
    path = os.path.join(mp.HERE, "libldg_probe.so")
    if not os.path.exists(path):
        pytest.skip("libldg_probe.so not built")
    lib = ctypes.CDLL(path)
    assert len(mp.pack(*mp.operators())) == lib.ldg_probe_ops_bytes()
