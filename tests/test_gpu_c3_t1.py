"""Mixed precision at the headline size: Taylor-Green on 64^3 elements,
p = 4 (163.84 M DOF), Re 1600, Mach 0.1, advanced to t = 1 (9,168 steps of
the reference time step, CFL 0.01) by the FP64 path and by the mixed path
(FP32 flux / quadrature, FP64 state, RK4, inverse mass and reductions).
Bar: relative L2 <= 1e-5 per conserved variable of the mixed state against
the FP64 state, and the kinetic-energy histories within 1e-6 relative.

At this size the CPU oracle would need about 4 days, so the FP64 GPU path,
pinned to the oracle and the exact answer on the smaller meshes, is the
reference here.  The run is split into chunks of about 2,300 steps, one test
each, so that every test ends within about a minute."""
import numpy as np
import pytest

import ldg_amd

pytestmark = pytest.mark.gpu

P, NX, T_END, CHUNKS = 4, 64, 1.0, 4
_run = {}


def _setup():
    if _run:
        return _run
    g = ldg_amd.LDG(P, NX)
    dt0 = g.time_step(0.01)
    n = int(np.ceil(T_END / dt0))
    g.close()
    _run.update(dt=T_END / n, nsteps=n, state={}, ke={}, done={})
    return _run


def _advance(prec, chunk):
    r = _setup()
    key = "fp64" if prec == ldg_amd.LDG_FP64 else "mixed"
    while r["done"].get(key, -1) < chunk:
        c = r["done"].get(key, -1) + 1
        g = ldg_amd.LDG(P, NX, precision=prec)
        if c == 0:
            g.initial_condition()
        else:
            g.set_state(r["state"][key])
        lo, hi = r["nsteps"] * c // CHUNKS, r["nsteps"] * (c + 1) // CHUNKS
        print(f"t=1 run {key}: steps {lo}..{hi} of {r['nsteps']}", flush=True)
        ke = g.step(r["dt"], hi - lo, record_ke=True)
        r["state"][key] = g.get_state()
        r["ke"].setdefault(key, []).append(ke)
        r["done"][key] = c
        g.close()


@pytest.mark.parametrize("chunk", range(CHUNKS))
def test_fp64_to_t1(chunk):
    _advance(ldg_amd.LDG_FP64, chunk)


@pytest.mark.parametrize("chunk", range(CHUNKS))
def test_mixed_to_t1(chunk):
    _advance(ldg_amd.LDG_MIXED, chunk)


def test_mixed_error_at_t1():
    _advance(ldg_amd.LDG_FP64, CHUNKS - 1)
    _advance(ldg_amd.LDG_MIXED, CHUNKS - 1)
    u64, u32 = _run["state"]["fp64"], _run["state"]["mixed"]
    e = np.sqrt(np.sum((u32 - u64) ** 2, axis=(0, 2))) / np.sqrt(np.sum(u64 ** 2, axis=(0, 2)))
    ke64 = np.concatenate(_run["ke"]["fp64"])
    ke32 = np.concatenate(_run["ke"]["mixed"])
    eke = np.max(np.abs(ke32 - ke64) / ke64)
    print(f"64^3 p=4 t=1 ({_run['nsteps']} steps): mixed vs FP64 relative L2 per variable {e}, "
          f"KE history max relative difference {eke:.3e}", flush=True)
    assert np.all(e <= 1e-5), e
    assert eke <= 1e-6
    _run.clear()
