"""Python mirror of the Local-DG solver API over the libldg.so C ABI.

The reference's public surface is the Mesh and Solver classes
(mesh.h:53-57, solver.h:110-113) plus the driver options (driver.cpp:26-39).
This module exposes the same names and argument meanings:

    mesh = Mesh(nx, ny, nz, botLeft, topRight)
    solver = Solver(p, dtSnaps, tf, L, mesh)
    solver.dgTimeStep()

and a thinner `LDG` handle that maps one-to-one onto include/ldg.h.  All
numerics run in libldg.so on the GPU; there is no CPU fallback: if the
library or a HIP device is missing, construction raises.
"""
import ctypes
import math
import os

import numpy as np

from . import io as io  # noqa: F401  (XYZV snapshot writer)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libldg.so")

LDG_OK, LDG_EINVAL, LDG_ENOMEM, LDG_EDEVICE, LDG_ENONFINITE = 0, -1, -2, -3, -4
LDG_FP64, LDG_MIXED = 0, 1
LDG_CLOSURE_DISSIPATIVE = 1
LDG_CLOSURE_ENERGIZING = 2
LDG_CLOSURE_PER_STEP = 4

EXPORTED = [
    "ldg_config_default", "ldg_create", "ldg_destroy", "ldg_last_error", "ldg_sizes",
    "ldg_initial_condition", "ldg_set_state", "ldg_get_state", "ldg_state_device_ptr",
    "ldg_time_step", "ldg_rhs", "ldg_step", "ldg_diagnostics", "ldg_kernel_times",
    "ldg_ensemble_sample", "ldg_ensemble_create", "ldg_instance_count", "ldg_instance_sizes",
    "ldg_instance_time_step", "ldg_instance_set_state", "ldg_instance_get_state",
    "ldg_instance_diagnostics", "ldg_ensemble_step", "ldg_set_graph", "ldg_autotune",
    "ldg_closure_create", "ldg_closure_destroy", "ldg_closure_last_error", "ldg_closure_sizes",
    "ldg_closure_get", "ldg_closure_set", "ldg_closure_forward", "ldg_closure_gradient",
    "ldg_project_state", "ldg_closure_collect", "ldg_closure_add_samples", "ldg_closure_samples",
    "ldg_closure_get_samples", "ldg_closure_fit_normalization", "ldg_closure_train", "ldg_set_closure",
    "ldg_closure_gated", "ldg_closure_set_weighting", "ldg_closure_get_weights", "ldg_rhs_noalloc",
]


class ldg_config(ctypes.Structure):
    _fields_ = [
        ("p", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("xmin", ctypes.c_double * 3), ("xmax", ctypes.c_double * 3), ("L", ctypes.c_double),
        ("Re", ctypes.c_double), ("Pr", ctypes.c_double), ("gamma", ctypes.c_double),
        ("M0", ctypes.c_double), ("rho0", ctypes.c_double), ("V0", ctypes.c_double),
        ("R", ctypes.c_double), ("precision", ctypes.c_int),
    ]


class ldg_instance(ctypes.Structure):
    _fields_ = [
        ("p", ctypes.c_int), ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("Re", ctypes.c_double), ("M0", ctypes.c_double), ("V0", ctypes.c_double),
        ("phase", ctypes.c_double * 3),
    ]


class LDGError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libldg error {code}: {msg}")
        self.code = code


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)


def load_library(path=LIB_PATH):
    """Load libldg.so and declare every symbol of include/ldg.h."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: build it with __graft_entry__.build() or make -C local-dg_amd")
    L = ctypes.CDLL(path)
    h = ctypes.c_void_p
    L.ldg_config_default.argtypes = [ctypes.POINTER(ldg_config)]
    L.ldg_create.argtypes = [ctypes.POINTER(ldg_config), ctypes.POINTER(h)]
    L.ldg_destroy.argtypes = [h]
    L.ldg_destroy.restype = None
    L.ldg_last_error.argtypes = [h]
    L.ldg_last_error.restype = ctypes.c_char_p
    L.ldg_sizes.argtypes = [h, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    L.ldg_initial_condition.argtypes = [h]
    L.ldg_set_state.argtypes = [h, _dp]
    L.ldg_get_state.argtypes = [h, _dp]
    L.ldg_state_device_ptr.argtypes = [h, ctypes.POINTER(_dp)]
    L.ldg_time_step.argtypes = [h, ctypes.c_double]
    L.ldg_time_step.restype = ctypes.c_double
    L.ldg_rhs.argtypes = [h, _dp, _dp, _dp]
    L.ldg_rhs_noalloc.argtypes = [ctypes.POINTER(ldg_config), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    L.ldg_step.argtypes = [h, ctypes.c_double, ctypes.c_int, _dp]
    L.ldg_diagnostics.argtypes = [h, _dp]
    L.ldg_kernel_times.argtypes = [h, ctypes.c_double, ctypes.c_int, _dp]
    ip = ctypes.POINTER(ctypes.c_int)
    L.ldg_ensemble_sample.argtypes = [ctypes.c_ulonglong, ctypes.c_int, ctypes.POINTER(ldg_instance)]
    L.ldg_ensemble_create.argtypes = [ctypes.POINTER(ldg_config), ctypes.POINTER(ldg_instance), ctypes.c_int,
                                      ctypes.POINTER(h)]
    L.ldg_instance_count.argtypes = [h, ip]
    L.ldg_instance_sizes.argtypes = [h, ctypes.c_int, ip, ip, ip]
    L.ldg_instance_time_step.argtypes = [h, ctypes.c_int, ctypes.c_double, _dp]
    L.ldg_instance_set_state.argtypes = [h, ctypes.c_int, _dp]
    L.ldg_instance_get_state.argtypes = [h, ctypes.c_int, _dp]
    L.ldg_instance_diagnostics.argtypes = [h, ctypes.c_int, _dp]
    L.ldg_ensemble_step.argtypes = [h, _dp, ctypes.c_int, _dp]
    L.ldg_set_graph.argtypes = [h, ctypes.c_int]
    L.ldg_autotune.argtypes = [h, ip]
    fp = ctypes.POINTER(ctypes.c_float)
    lp = ctypes.POINTER(ctypes.c_long)
    L.ldg_closure_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_long, ctypes.c_ulonglong, ctypes.POINTER(h)]
    L.ldg_closure_destroy.argtypes = [h]
    L.ldg_closure_destroy.restype = None
    L.ldg_closure_last_error.argtypes = [h]
    L.ldg_closure_last_error.restype = ctypes.c_char_p
    L.ldg_closure_sizes.argtypes = [h, ip, ip, ip, lp]
    L.ldg_closure_get.argtypes = [h, fp, fp]
    L.ldg_closure_set.argtypes = [h, fp, fp]
    L.ldg_closure_forward.argtypes = [h, fp, ctypes.c_int, fp]
    L.ldg_closure_gradient.argtypes = [h, fp, fp, ctypes.c_int, fp, _dp]
    L.ldg_project_state.argtypes = [h, ctypes.c_int, h, ctypes.c_int]
    L.ldg_closure_collect.argtypes = [h, h, h, ctypes.c_double, ctypes.c_ulonglong]
    L.ldg_closure_add_samples.argtypes = [h, fp, fp, ctypes.c_long]
    L.ldg_closure_samples.argtypes = [h, lp]
    L.ldg_closure_get_samples.argtypes = [h, ctypes.c_long, ctypes.c_long, fp, fp]
    L.ldg_closure_fit_normalization.argtypes = [h]
    L.ldg_closure_train.argtypes = [h, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_ulonglong, _dp]
    L.ldg_set_closure.argtypes = [h, h, ctypes.c_double, ctypes.c_int]
    L.ldg_closure_gated.argtypes = [h, lp, lp]
    L.ldg_closure_set_weighting.argtypes = [h, ctypes.c_int]
    L.ldg_closure_get_weights.argtypes = [h, ctypes.c_long, ctypes.c_long, fp]
    for name in EXPORTED:
        fn = getattr(L, name)
        if name not in ("ldg_destroy", "ldg_last_error", "ldg_time_step", "ldg_closure_destroy",
                        "ldg_closure_last_error"):
            fn.restype = ctypes.c_int
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(_dp)


class LDG:
    """One solver instance on the GPU (a thin owner of an ldg_handle)."""

    def __init__(self, p, nx, ny=None, nz=None, xmin=None, xmax=None, L=1.0, Re=1600.0, Pr=0.71,
                 gamma=1.4, M0=0.1, rho0=1.0, V0=1.0, R=8.314, precision=LDG_FP64):
        lib = load_library()
        cfg = ldg_config()
        lib.ldg_config_default(ctypes.byref(cfg))
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        cfg.p, cfg.nx, cfg.ny, cfg.nz = p, nx, ny, nz
        xmin = [-math.pi * L] * 3 if xmin is None else list(xmin)
        xmax = [math.pi * L] * 3 if xmax is None else list(xmax)
        for l in range(3):
            cfg.xmin[l], cfg.xmax[l] = xmin[l], xmax[l]
        cfg.L, cfg.Re, cfg.Pr, cfg.gamma, cfg.M0 = L, Re, Pr, gamma, M0
        cfg.rho0, cfg.V0, cfg.R, cfg.precision = rho0, V0, R, precision
        self._lib = lib
        self._h = ctypes.c_void_p()
        rc = lib.ldg_create(ctypes.byref(cfg), ctypes.byref(self._h))
        if rc != LDG_OK:
            raise LDGError(rc, lib.ldg_last_error(None).decode())
        d, k = ctypes.c_int(), ctypes.c_int()
        lib.ldg_sizes(self._h, ctypes.byref(d), ctypes.byref(k))
        self.p, self.n, self.dofs, self.K = p, p + 1, d.value, k.value
        self.cfg = cfg

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.ldg_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc != LDG_OK:
            raise LDGError(rc, self._lib.ldg_last_error(self._h).decode())

    @property
    def shape(self):
        return (self.K, 5, self.dofs)

    def initial_condition(self):
        self._check(self._lib.ldg_initial_condition(self._h))

    def set_state(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert u.shape == self.shape, (u.shape, self.shape)
        self._check(self._lib.ldg_set_state(self._h, _p(u)))

    def get_state(self):
        u = np.empty(self.shape)
        self._check(self._lib.ldg_get_state(self._h, _p(u)))
        return u

    def device_ptr(self):
        ptr = _dp()
        self._check(self._lib.ldg_state_device_ptr(self._h, ctypes.byref(ptr)))
        return ctypes.cast(ptr, ctypes.c_void_p).value

    def time_step(self, cfl=0.01):
        return self._lib.ldg_time_step(self._h, cfl)

    def rhs(self, u, want_du=False):
        u = np.ascontiguousarray(u, dtype=np.float64)
        k = np.empty(self.shape)
        du = np.empty((3,) + self.shape) if want_du else None
        self._check(self._lib.ldg_rhs(self._h, _p(u), _p(k), _p(du) if want_du else None))
        return (k, du) if want_du else k

    def step(self, dt, nsteps, record_ke=True):
        ke = np.empty(nsteps) if record_ke else None
        self._check(self._lib.ldg_step(self._h, dt, nsteps, _p(ke) if record_ke else None))
        return ke

    def diagnostics(self):
        out = np.empty(3)
        self._check(self._lib.ldg_diagnostics(self._h, _p(out)))
        return {"ke_ref": out[0], "ke": out[1], "enstrophy": out[2]}

    def kernel_times(self, dt, nsteps):
        out = np.empty(2)
        self._check(self._lib.ldg_kernel_times(self._h, dt, nsteps, _p(out)))
        return {"grad_ms": out[0], "rhs_ms": out[1]}

    def set_graph(self, enable=True):
        """Replay steps from a HIP graph of two RK4 steps (ldg_set_graph)."""
        self._check(self._lib.ldg_set_graph(self._h, int(bool(enable))))

    def autotune(self):
        """Pick the fastest tiling per kernel (ldg_autotune); returns
        [(p, elements per workgroup of the gradient, of the rhs), ...]."""
        out = (ctypes.c_int * 15)()
        self._check(self._lib.ldg_autotune(self._h, out))
        ng = len({p for p in getattr(self, "orders", [self.p])})
        return [tuple(out[3 * g:3 * g + 3]) for g in range(ng)]


def rhs_noalloc(cfg, u_dev, k_dev, stream=None):
    """k = rhs(u) on device arrays (addresses u_dev, k_dev: (K, 5, dofs)
    FP64), with no device allocation (ldg_rhs_noalloc); cfg is an
    ldg_config (LDG(...).cfg)."""
    lib = load_library()
    rc = lib.ldg_rhs_noalloc(ctypes.byref(cfg), ctypes.c_void_p(u_dev), ctypes.c_void_p(k_dev),
                             ctypes.c_void_p(stream) if stream else None)
    if rc != LDG_OK:
        raise LDGError(rc, lib.ldg_last_error(None).decode())


def sample_instances(seed, n):
    """n instance descriptions drawn by ldg_ensemble_sample (a list of dicts)."""
    lib = load_library()
    arr = (ldg_instance * max(n, 1))()
    rc = lib.ldg_ensemble_sample(seed, n, arr)
    if rc != LDG_OK:
        raise LDGError(rc, "ldg_ensemble_sample")
    return [{"p": a.p, "nx": a.nx, "ny": a.ny, "nz": a.nz, "Re": a.Re, "M0": a.M0, "V0": a.V0,
             "phase": tuple(a.phase)} for a in arr[:n]]


class Ensemble(LDG):
    """Many independent Taylor-Green solvers advanced by shared kernel
    launches (ldg_ensemble_create).  `instances` is a list of dicts with
    keys p, nx (ny, nz default to nx), Re, M0, V0 and phase; the domain,
    remaining physics and precision are shared."""

    def __init__(self, instances, xmin=None, xmax=None, L=1.0, Pr=0.71, gamma=1.4, rho0=1.0, R=8.314,
                 precision=LDG_FP64):
        lib = load_library()
        cfg = ldg_config()
        lib.ldg_config_default(ctypes.byref(cfg))
        xmin = [-math.pi * L] * 3 if xmin is None else list(xmin)
        xmax = [math.pi * L] * 3 if xmax is None else list(xmax)
        for l in range(3):
            cfg.xmin[l], cfg.xmax[l] = xmin[l], xmax[l]
        cfg.L, cfg.Pr, cfg.gamma, cfg.rho0, cfg.R, cfg.precision = L, Pr, gamma, rho0, R, precision
        n = len(instances)
        arr = (ldg_instance * max(n, 1))()
        for a, d in zip(arr, instances):
            a.p, a.nx = d["p"], d["nx"]
            a.ny, a.nz = d.get("ny", d["nx"]), d.get("nz", d["nx"])
            a.Re, a.M0, a.V0 = d.get("Re", 1600.0), d.get("M0", 0.1), d.get("V0", 1.0)
            for l, v in enumerate(d.get("phase", (0.0, 0.0, 0.0))):
                a.phase[l] = v
        self._lib = lib
        self._h = ctypes.c_void_p()
        rc = lib.ldg_ensemble_create(ctypes.byref(cfg), arr, n, ctypes.byref(self._h))
        if rc != LDG_OK:
            raise LDGError(rc, lib.ldg_last_error(None).decode())
        self.cfg = cfg
        self.count = n
        self.orders, self.dofs_i, self.K_i = [], [], []
        for i in range(n):
            p, d, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            self._check(lib.ldg_instance_sizes(self._h, i, ctypes.byref(p), ctypes.byref(d), ctypes.byref(k)))
            self.orders.append(p.value)
            self.dofs_i.append(d.value)
            self.K_i.append(k.value)
        self.p, self.n, self.dofs, self.K = self.orders[0], self.orders[0] + 1, self.dofs_i[0], self.K_i[0]

    def autotune(self):
        out = (ctypes.c_int * 15)()
        self._check(self._lib.ldg_autotune(self._h, out))
        return [tuple(out[3 * g:3 * g + 3]) for g in range(len(dict.fromkeys(self.orders)))]

    def instance_shape(self, i):
        return (self.K_i[i], 5, self.dofs_i[i])

    def total_dof(self):
        return sum(k * d * 5 for k, d in zip(self.K_i, self.dofs_i))

    def time_steps(self, cfl=0.01):
        out = np.empty(self.count)
        for i in range(self.count):
            v = ctypes.c_double()
            self._check(self._lib.ldg_instance_time_step(self._h, i, cfl, ctypes.byref(v)))
            out[i] = v.value
        return out

    def set_instance_state(self, i, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        assert u.shape == self.instance_shape(i), (u.shape, self.instance_shape(i))
        self._check(self._lib.ldg_instance_set_state(self._h, i, _p(u)))

    def get_instance_state(self, i):
        u = np.empty(self.instance_shape(i))
        self._check(self._lib.ldg_instance_get_state(self._h, i, _p(u)))
        return u

    def instance_diagnostics(self, i):
        out = np.empty(3)
        self._check(self._lib.ldg_instance_diagnostics(self._h, i, _p(out)))
        return {"ke_ref": out[0], "ke": out[1], "enstrophy": out[2]}

    def step(self, dt, nsteps, record_ke=True):
        """nsteps steps, instance i with dt[i]; returns KE (nsteps, count)."""
        dt = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, dtype=np.float64), (self.count,)))
        ke = np.empty((nsteps, self.count)) if record_ke else None
        self._check(self._lib.ldg_ensemble_step(self._h, _p(dt), nsteps, _p(ke) if record_ke and nsteps else None))
        return ke


class Point:
    """Mesh corner (mesh.h:20-38)."""

    def __init__(self, x=0.0, y=0.0, z=0.0):
        self.x, self.y, self.z = float(x), float(y), float(z)


class Mesh:
    """Mesh(nx, ny, nz, botLeft, topRight) of a periodic box (mesh.cpp:51).

    One GPU holds the whole mesh, so there is no MPI argument; the periodic
    neighbours the reference reaches through MPIUtil are gathered on the
    device.  The reference's default box is [0,1]^3 (mesh.cpp:46)."""

    def __init__(self, nx=10, ny=10, nz=10, botLeft=None, topRight=None):
        if min(nx, ny, nz) < 3:
            raise ValueError("need at least 3 elements per direction (mesh.cpp:102-107)")
        self.nx, self.ny, self.nz = nx, ny, nz
        self.botLeft = botLeft or Point(0.0, 0.0, 0.0)
        self.topRight = topRight or Point(1.0, 1.0, 1.0)
        self.nElements = nx * ny * nz
        self.minDX = (self.topRight.x - self.botLeft.x) / nx
        self.minDY = (self.topRight.y - self.botLeft.y) / ny
        self.minDZ = (self.topRight.z - self.botLeft.z) / nz


class Solver:
    """Solver(p, dtSnaps, tf, L, mesh) and dgTimeStep() (solver.h:110-113).

    dt follows the reference rule dt = cfl min(dx)/(V0 (2p+1)), rounded so
    that an integer number of steps ends exactly at tf (solver.cpp:44-52);
    the reference's cfl is 0.01.  dgTimeStep() advances the Taylor-Green
    state to tf on the GPU, records the reference kinetic energy after every
    step and the dissipation -(KE[i] - KE[i-2])/(2 dt) (solver.cpp:437-440),
    and writes a Paraview XYZV snapshot every dtSnaps steps when
    `output_dir` is set (solver.cpp:366-381).
    """

    def __init__(self, p=2, dtSnaps=10, tf=1.0, L=1.0, mesh=None, cfl=0.01, precision=LDG_FP64,
                 output_dir=None, **physics):
        mesh = mesh or Mesh()
        self.mesh, self.order, self.dtSnaps, self.tf, self.L = mesh, p, dtSnaps, tf, L
        self.ldg = LDG(p, mesh.nx, mesh.ny, mesh.nz,
                       [mesh.botLeft.x, mesh.botLeft.y, mesh.botLeft.z],
                       [mesh.topRight.x, mesh.topRight.y, mesh.topRight.z], L=L, precision=precision, **physics)
        self.dofs = self.ldg.dofs
        dt = self.ldg.time_step(cfl)
        self.timesteps = int(math.ceil(tf / dt))
        self.dt = tf / self.timesteps
        self.output_dir = output_dir
        self.ldg.initial_condition()
        self.kes = None
        self.keDissipation = None

    @property
    def u(self):
        return self.ldg.get_state()

    def dgTimeStep(self):
        kes = np.empty(self.timesteps)
        done = 0
        while done < self.timesteps:
            if self.output_dir is not None and done % self.dtSnaps == 0:
                self._snapshot(done // self.dtSnaps)
            chunk = self.timesteps - done
            if self.output_dir is not None:
                chunk = min(chunk, self.dtSnaps - done % self.dtSnaps)
            kes[done:done + chunk] = self.ldg.step(self.dt, chunk, record_ke=True)
            done += chunk
        self.kes = kes
        self.keDissipation = -(kes[2:] - kes[:-2]) / (2 * self.dt)
        return self.keDissipation

    def _snapshot(self, index):
        from .io import init_xyzv_file, export_to_xyzv_file, node_coordinates
        os.makedirs(self.output_dir, exist_ok=True)
        name = os.path.join(self.output_dir, "xyzu.txt")
        if not init_xyzv_file(name, index, "u", 5):
            raise OSError(f"could not open {name}.{index}")
        coords = node_coordinates(self.order, self.mesh)
        if not export_to_xyzv_file(name, index, coords, self.u):
            raise OSError(f"could not write {name}.{index}")
