"""ctypes binding of the CPU oracle (oracle/ldg_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.  The
product (libldg.so and the ldg_amd package) never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATHS = {"plain": os.path.join(_HERE, "_build", "libldg_oracle.so"),
              "fma": os.path.join(_HERE, "_build", "libldg_oracle_fma.so"),
              "ld": os.path.join(_HERE, "_build", "libldg_oracle_ld.so"),
              "native": os.path.join(_HERE, "_build", "libldg_oracle_native.so")}
_libs = {}

_dp = ctypes.POINTER(ctypes.c_double)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(variant="plain"):
    """The oracle library.  Variants of the same source: "plain" (FP64, the
    oracle), "fma" (FP64 with fused multiply-adds: the oracle's own
    round-off spread), "ld" (x87 long double throughout: the exact answer
    the FP64 builds are measured against), "native" (FP64 for the building
    host's vector units, bitwise equal to "plain"; fixture generation)."""
    if variant not in _libs:
        path = _LIB_PATHS[variant]
        if not os.path.exists(path):
            if variant == "native":
                subprocess.run(["make", "-s", "-C", _HERE, "native"], check=True)
            else:
                build()
        L = ctypes.CDLL(path)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int] * 4 + [_dp, _dp, ctypes.c_double, _dp]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        for f in ("orc_dofs", "orc_nelem", "orc_nq"):
            getattr(L, f).argtypes = [ctypes.c_void_p]
            getattr(L, f).restype = ctypes.c_int
        L.orc_initial_condition.argtypes = [ctypes.c_void_p, _dp]
        L.orc_rhs.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        L.orc_step.argtypes = [ctypes.c_void_p, _dp, ctypes.c_double, ctypes.c_int, _dp]
        L.orc_step.restype = ctypes.c_int
        L.orc_ke.argtypes = [ctypes.c_void_p, _dp]
        L.orc_ke.restype = ctypes.c_double
        L.orc_diagnostics.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_dt.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_dt.restype = ctypes.c_double
        L.orc_get_gauss.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_get_coords.argtypes = [ctypes.c_void_p, _dp]
        L.orc_get_operator.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_get_operator.restype = ctypes.c_int
        L.orc_get_physics.argtypes = [ctypes.c_void_p, _dp]
        L.orc_fluxC.argtypes = [_dp, ctypes.c_double, _dp]
        L.orc_fluxV.argtypes = [_dp, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_roe.argtypes = [_dp, _dp, _dp, ctypes.c_double, _dp]
        L.orc_real_size.restype = ctypes.c_int
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_threads.restype = None
        L.orc_max_threads.restype = ctypes.c_int
        _libs[variant] = L
    return _libs[variant]


def _p(a):
    return a.ctypes.data_as(_dp)


def fluxC(u, gamma=1.4):
    u = np.ascontiguousarray(u, dtype=np.float64)
    out = np.zeros(15)
    lib().orc_fluxC(_p(u), gamma, _p(out))
    return out


def fluxV(u, Du, Re=1600.0, Pr=0.71, gamma=1.4):
    u = np.ascontiguousarray(u, dtype=np.float64)
    Du = np.ascontiguousarray(Du, dtype=np.float64)
    out = np.zeros(15)
    lib().orc_fluxV(_p(u), _p(Du), Re, Pr, gamma, _p(out))
    return out


def roe(uN, uK, n, gamma=1.4):
    uN = np.ascontiguousarray(uN, dtype=np.float64)
    uK = np.ascontiguousarray(uK, dtype=np.float64)
    n = np.ascontiguousarray(n, dtype=np.float64)
    out = np.zeros(5)
    lib().orc_roe(_p(uN), _p(uK), _p(n), gamma, _p(out))
    return out


class Oracle:
    """CPU oracle for one periodic hex mesh at order p.

    Arrays follow the reference layout: u is (K, 5, dofs) in C order, i.e.
    u(iN, iS, iK) column-major as in array.h.
    """

    def __init__(self, p, nx, ny=None, nz=None, xmin=None, xmax=None, L=1.0, phys=None, variant="plain"):
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        xmin = np.full(3, -np.pi * L) if xmin is None else np.asarray(xmin, dtype=np.float64)
        xmax = np.full(3, np.pi * L) if xmax is None else np.asarray(xmax, dtype=np.float64)
        xmin = np.ascontiguousarray(xmin, dtype=np.float64)
        xmax = np.ascontiguousarray(xmax, dtype=np.float64)
        ph = None
        if phys is not None:
            ph = np.ascontiguousarray(phys, dtype=np.float64)
        self._lib = lib(variant)
        self._h = self._lib.orc_create(p, nx, ny, nz, _p(xmin), _p(xmax), L, _p(ph) if ph is not None else None)
        if not self._h:
            raise ValueError("oracle setup failed")
        self.p, self.n = p, p + 1
        self.nx, self.ny, self.nz = nx, ny, nz
        self.dofs = self._lib.orc_dofs(self._h)
        self.K = self._lib.orc_nelem(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            self._lib.orc_destroy(self._h)
            self._h = None

    @property
    def shape(self):
        return (self.K, 5, self.dofs)

    def initial_condition(self):
        u = np.zeros(self.shape)
        self._lib.orc_initial_condition(self._h, _p(u))
        return u

    def rhs(self, u, want_du=False):
        u = np.ascontiguousarray(u, dtype=np.float64)
        k = np.zeros(self.shape)
        du = np.zeros((3,) + self.shape) if want_du else None
        self._lib.orc_rhs(self._h, _p(u), _p(k), _p(du) if want_du else None)
        return (k, du) if want_du else k

    def step(self, u, dt, nsteps, record_ke=True):
        u = np.array(u, dtype=np.float64, order="C", copy=True)
        ke = np.zeros(nsteps) if record_ke else None
        rc = self._lib.orc_step(self._h, _p(u), dt, nsteps, _p(ke) if record_ke else None)
        if rc != 0:
            raise MemoryError("oracle step failed")
        return u, ke

    def ke(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        return self._lib.orc_ke(self._h, _p(u))

    def diagnostics(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros(2)
        self._lib.orc_diagnostics(self._h, _p(u), _p(out))
        return out

    def dt(self, cfl):
        return self._lib.orc_dt(self._h, cfl)

    def gauss(self):
        x = np.zeros(self.n)
        w = np.zeros(self.n)
        self._lib.orc_get_gauss(self._h, _p(x), _p(w))
        return x, w

    def coords(self):
        c = np.zeros((self.K, self.dofs, 3))
        self._lib.orc_get_coords(self._h, _p(c))
        return c

    def operator(self, which):
        names = {"Interp2D": 0, "Interp3D": 1, "Mel": 2, "Sels": 3, "Kels": 4, "Kels2D": 5}
        n, d = self.n, self.dofs
        shapes = {0: (n * n, n * n), 1: (d, d), 2: (d, d), 3: (3, d, d), 4: (3, d, d), 5: (3, n * n, n * n)}
        w = names[which]
        out = np.zeros(int(np.prod(shapes[w])))
        self._lib.orc_get_operator(self._h, w, _p(out))
        # column-major (rows fastest) -> numpy [.., col, row] then transpose last two
        sh = shapes[w]
        if len(sh) == 2:
            return out.reshape(sh[1], sh[0]).T.copy()
        return np.stack([out.reshape(sh[0], sh[2], sh[1])[l].T for l in range(3)])

    def physics(self):
        out = np.zeros(11)
        self._lib.orc_get_physics(self._h, _p(out))
        keys = ["Re", "Pr", "gamma", "M0", "rho0", "V0", "R", "c0", "p0", "T0", "J"]
        return dict(zip(keys, out))
