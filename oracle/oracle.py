"""ctypes bindings of the FP64 CPU oracle (oracle/liboracle.so) — TEST INFRASTRUCTURE.

The oracle is a plain C++17/OpenMP restatement of the reference's dynamics path
(see the headers of ora_core.hpp and ora_capi.cpp for the file:line each part
follows). This module only marshals numpy arrays; it holds no numerics.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

PARAM_ORDER = (
    "rho_ice", "rho_atm", "rho_ocean", "C_atm", "C_ocean", "F_atm", "F_ocean", "fc",
    "ocean_turning_angle", "gravity", "Pstar", "DeltaMin", "alpha", "beta", "nSteps",
)

_dp = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_sz = ctypes.c_size_t
_vp = ctypes.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load():
    """Load liboracle.so (building it first when it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB):
        build()
    lib = ctypes.CDLL(_LIB)
    sig = {
        "ora_last_error": (ctypes.c_char_p, []),
        "ora_mesh_new": (_vp, []),
        "ora_mesh_free": (None, [_vp]),
        "ora_mesh_from_coords": (ctypes.c_int, [_vp, _sz, _sz, _dp, _dp, ctypes.c_int]),
        "ora_mesh_rectangle": (ctypes.c_int, [_vp, ctypes.c_double, ctypes.c_double, _sz, _sz, ctypes.c_double]),
        "ora_mesh_ring": (ctypes.c_int, [_vp, _sz, _sz]),
        "ora_mesh_read_smesh": (ctypes.c_int, [_vp, ctypes.c_char_p]),
        "ora_mesh_size": (None, [_vp, ctypes.POINTER(_sz), ctypes.POINTER(_sz)]),
        "ora_mesh_dirichlet": (_sz, [_vp, ctypes.c_int, _u64p]),
        "ora_mesh_periodic": (_sz, [_vp, _u64p]),
        "ora_mesh_landmask": (None, [_vp, ctypes.POINTER(ctypes.c_uint8)]),
        "ora_mesh_vertices": (None, [_vp, _dp]),
        "ora_function2dg": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
        "ora_l2error_dg": (ctypes.c_double, [_vp, ctypes.c_int, ctypes.c_int, _dp]),
        "ora_mean_value": (ctypes.c_double, [_vp, ctypes.c_int, _dp]),
        "ora_dg2cg": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp]),
        "ora_cg2dg": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _dp, _dp]),
        "ora_limit": (ctypes.c_int, [ctypes.c_int, _sz, _dp, ctypes.c_double, ctypes.c_int]),
        "ora_transport_new": (_vp, [_vp, ctypes.c_int]),
        "ora_transport_free": (None, [_vp]),
        "ora_transport_set_scheme": (ctypes.c_int, [_vp, ctypes.c_int]),
        "ora_transport_set_velocity": (ctypes.c_int, [_vp, _dp, _dp]),
        "ora_transport_prepare_cg": (ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
        "ora_transport_reinit": (ctypes.c_int, [_vp]),
        "ora_transport_step": (ctypes.c_int, [_vp, ctypes.c_double, _dp]),
        "ora_transport_op": (ctypes.c_int, [_vp, ctypes.c_double, _dp, _dp]),
        "ora_transport_normalvel": (ctypes.c_int, [_vp, _dp, _dp]),
        "ora_advection_test": (ctypes.c_double, [ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]),
        "ora_advection_ring_test": (ctypes.c_double, [ctypes.c_int, ctypes.c_int, _dp]),
        "ora_dyn_new": (_vp, [ctypes.c_int, ctypes.c_int, _dp]),
        "ora_dyn_new_bbm": (_vp, [ctypes.c_int, ctypes.c_int, _dp, _dp]),
        "ora_dyn_free": (None, [_vp]),
        "ora_dyn_initialise": (ctypes.c_int, [_vp, _sz, _sz, _dp, _dp, ctypes.c_int]),
        "ora_dyn_initialise_mesh": (ctypes.c_int, [_vp, _vp]),
        "ora_dyn_set_field": (ctypes.c_int, [_vp, ctypes.c_char_p, _dp, ctypes.c_int]),
        "ora_dyn_get_field": (ctypes.c_int, [_vp, ctypes.c_char_p, _dp, ctypes.c_int]),
        "ora_dyn_get_cg": (ctypes.c_int, [_vp, ctypes.c_char_p, _dp]),
        "ora_dyn_set_cg": (ctypes.c_int, [_vp, ctypes.c_char_p, _dp]),
        "ora_dyn_ncg": (_sz, [_vp]),
        "ora_dyn_update": (ctypes.c_int, [_vp, ctypes.c_double]),
        "ora_dyn_advection": (ctypes.c_int, [_vp, ctypes.c_double]),
        "ora_dyn_prepare_iteration": (ctypes.c_int, [_vp]),
        "ora_dyn_begin_iteration": (ctypes.c_int, [_vp, ctypes.c_double]),
        "ora_dyn_pass": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
        "ora_dyn_mass": (ctypes.c_int, [_vp, _dp]),
        "ora_set_threads": (ctypes.c_int, [ctypes.c_int]),
        "ora_dyn_budget": (ctypes.c_int, [_vp, _dp, ctypes.c_int]),
        "ora_benchmark_wind": (None, [_sz, _sz, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, _dp, _dp]),
        "ora_table": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _check(rc):
    if rc != 0:
        raise RuntimeError("oracle: " + (load().ora_last_error() or b"").decode())


class OracleMesh:
    """ParametricMesh restatement (dynamics/src/include/ParametricMesh.hpp)."""

    def __init__(self):
        self.lib = load()
        self.h = self.lib.ora_mesh_new()

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ora_mesh_free(self.h)
            self.h = None

    @classmethod
    def from_coords(cls, nx, ny, coords, mask, spherical=False):
        m = cls()
        c, k = _f64(coords), _f64(mask)
        _check(m.lib.ora_mesh_from_coords(m.h, nx, ny, _p(c), _p(k), int(bool(spherical))))
        return m

    @classmethod
    def rectangle(cls, Lx, Ly, nx, ny, distort=0.0):
        m = cls()
        _check(m.lib.ora_mesh_rectangle(m.h, Lx, Ly, nx, ny, distort))
        return m

    @classmethod
    def ring(cls, nx, ny):
        m = cls()
        _check(m.lib.ora_mesh_ring(m.h, nx, ny))
        return m

    @classmethod
    def from_smesh(cls, text):
        m = cls()
        _check(m.lib.ora_mesh_read_smesh(m.h, text.encode()))
        return m

    @property
    def shape(self):
        nx, ny = _sz(), _sz()
        self.lib.ora_mesh_size(self.h, ctypes.byref(nx), ctypes.byref(ny))
        return nx.value, ny.value

    def dirichlet(self, edge):
        n = self.lib.ora_mesh_dirichlet(self.h, edge, None)
        out = np.zeros(n, dtype=np.uint64)
        self.lib.ora_mesh_dirichlet(self.h, edge, out.ctypes.data_as(_u64p))
        return out

    def periodic(self):
        n = self.lib.ora_mesh_periodic(self.h, None)
        out = np.zeros((n, 5), dtype=np.uint64)
        if n:
            self.lib.ora_mesh_periodic(self.h, out.ctypes.data_as(_u64p))
        return out

    def landmask(self):
        nx, ny = self.shape
        out = np.zeros(nx * ny, dtype=np.uint8)
        self.lib.ora_mesh_landmask(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
        return out

    def vertices(self):
        nx, ny = self.shape
        out = np.zeros(((nx + 1) * (ny + 1), 2))
        self.lib.ora_mesh_vertices(self.h, _p(out))
        return out


class OracleTransport:
    """DGTransport<DG> restatement (dynamics/src/DGTransport.cpp)."""

    def __init__(self, mesh, dg):
        self.lib = load()
        self.mesh, self.dg = mesh, dg
        self.h = self.lib.ora_transport_new(mesh.h, dg)
        if not self.h:
            raise RuntimeError("oracle: " + self.lib.ora_last_error().decode())

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ora_transport_free(self.h)
            self.h = None

    def set_scheme(self, rk):
        _check(self.lib.ora_transport_set_scheme(self.h, rk))

    def set_velocity(self, vx, vy):
        a, b = _f64(vx), _f64(vy)
        _check(self.lib.ora_transport_set_velocity(self.h, _p(a), _p(b)))

    def prepare_cg(self, cg, u, v):
        a, b = _f64(u), _f64(v)
        _check(self.lib.ora_transport_prepare_cg(self.h, cg, _p(a), _p(b)))

    def reinit(self):
        _check(self.lib.ora_transport_reinit(self.h))

    def step(self, dt, phi):
        out = _f64(phi).copy()
        _check(self.lib.ora_transport_step(self.h, dt, _p(out)))
        return out

    def op(self, dt, phi):
        a = _f64(phi)
        out = np.zeros_like(a)
        _check(self.lib.ora_transport_op(self.h, dt, _p(a), _p(out)))
        return out

    def normal_velocity(self):
        nx, ny = self.mesh.shape
        e = {1: 1, 3: 2, 6: 3, 8: 3}[self.dg]
        x = np.zeros(nx * (ny + 1) * e)
        y = np.zeros((nx + 1) * ny * e)
        _check(self.lib.ora_transport_normalvel(self.h, _p(x), _p(y)))
        return x.reshape(-1, e), y.reshape(-1, e)


MEB_ORDER = ("compaction_param", "nu0", "young", "P0", "undamaged_time_relaxation_sigma",
             "exponent_relaxation_sigma", "exponent_compression_factor", "tan_phi", "compr_strength", "C_lab")
MEB_DEFAULTS = dict(compaction_param=-20.0, nu0=1.0 / 3.0, young=5.96e8, P0=10.e3,
                    undamaged_time_relaxation_sigma=1e7, exponent_relaxation_sigma=5.0,
                    exponent_compression_factor=1.5, tan_phi=0.7, compr_strength=1e10, C_lab=2.0e6)


class OracleDyn:
    """MEVPDynamicsKernel restatement (VPCGDynamicsKernel.hpp / CGDynamicsKernel.cpp);
    with bbm=True the BBMDynamicsKernel (BrittleCGDynamicsKernel.hpp, BBMStressUpdateStep.hpp)."""

    PASSES = {"strain": 0, "stress": 1, "divergence": 2, "momentum": 3, "boundaries": 4, "subcycle": 5}

    def __init__(self, cg, dg, params=None, bbm=False, meb=None):
        self.lib = load()
        self.cg, self.dg = cg, dg
        self.S = 3 if cg == 1 else 8
        p = None
        if params is not None:
            arr = np.array([float(params[k]) for k in PARAM_ORDER])
            p = arr
        if bbm:
            m = np.array([float(dict(MEB_DEFAULTS, **(meb or {}))[k]) for k in MEB_ORDER])
            self._meb = m
            self.h = self.lib.ora_dyn_new_bbm(cg, dg, _p(p) if p is not None else None, _p(m))
        else:
            self.h = self.lib.ora_dyn_new(cg, dg, _p(p) if p is not None else None)
        if not self.h:
            raise RuntimeError("oracle: " + self.lib.ora_last_error().decode())
        self.nx = self.ny = None

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ora_dyn_free(self.h)
            self.h = None

    def initialise(self, nx, ny, coords, mask, spherical=False):
        """coords: vertices (x, y), or (longitude, latitude) in radians when spherical (the
        mesh is then rotated, DynamicsKernel.hpp:47-51)."""
        c, m = _f64(coords), _f64(mask)
        _check(self.lib.ora_dyn_initialise(self.h, nx, ny, _p(c), _p(m), int(spherical)))
        self.nx, self.ny = nx, ny

    def initialise_mesh(self, mesh):
        _check(self.lib.ora_dyn_initialise_mesh(self.h, mesh.h))
        self.nx, self.ny = mesh.shape

    @property
    def ne(self):
        return self.nx * self.ny

    @property
    def ncg(self):
        return self.lib.ora_dyn_ncg(self.h)

    def set_field(self, name, data, ncomp=None):
        a = _f64(data)
        if ncomp is None:
            ncomp = a.size // self.ne
        _check(self.lib.ora_dyn_set_field(self.h, name.encode(), _p(a), ncomp))

    def get_field(self, name, ncomp):
        out = np.zeros(self.ne * ncomp)
        _check(self.lib.ora_dyn_get_field(self.h, name.encode(), _p(out), ncomp))
        return out.reshape(self.ne, ncomp) if ncomp > 1 else out

    def get_cg(self, name):
        out = np.zeros(self.ncg)
        if self.lib.ora_dyn_get_cg(self.h, name.encode(), _p(out)) != 0:
            raise KeyError(name)
        return out

    def set_cg(self, name, data):
        a = _f64(data)
        if self.lib.ora_dyn_set_cg(self.h, name.encode(), _p(a)) != 0:
            raise KeyError(name)

    def update(self, dt):
        _check(self.lib.ora_dyn_update(self.h, dt))

    def advection(self, dt):
        _check(self.lib.ora_dyn_advection(self.h, dt))

    def prepare_iteration(self):
        _check(self.lib.ora_dyn_prepare_iteration(self.h))

    def begin_iteration(self, dt):
        _check(self.lib.ora_dyn_begin_iteration(self.h, dt))

    def run_pass(self, name, repeat=1):
        _check(self.lib.ora_dyn_pass(self.h, self.PASSES[name], repeat))

    def mass(self):
        """Tools::MeanValue of hice and cice: (ice volume, ice area)."""
        out = np.zeros(2)
        _check(self.lib.ora_dyn_mass(self.h, _p(out)))
        return float(out[0]), float(out[1])

    def budget(self, reset=False):
        """Advection mass budget since the last reset, as a dict with the Dirichlet outflow
        and the limiters' change of the MeanValue of h and A."""
        out = np.zeros(4)
        _check(self.lib.ora_dyn_budget(self.h, _p(out), int(bool(reset))))
        return dict(outflow_h=out[0], outflow_A=out[1], limiter_h=out[2], limiter_A=out[3])


def set_threads(n):
    """OpenMP threads of the oracle (n <= 0 leaves it); returns the count in use."""
    return load().ora_set_threads(int(n))


class Oracle:
    """Free functions of the oracle (interpolations, limiters, tools, test cases)."""

    def __init__(self):
        self.lib = load()

    def table(self, name, a, b=0, c=0):
        out = np.zeros(512)
        n = self.lib.ora_table(name.encode(), a, b, c, _p(out))
        if n < 0:
            raise KeyError(name)
        return out[:n]

    def advection_test(self, dg, distort, it):
        ml = ctypes.c_double()
        err = self.lib.ora_advection_test(dg, distort, it, ctypes.byref(ml))
        return err, ml.value

    def advection_ring_test(self, dg, it):
        ml = ctypes.c_double()
        err = self.lib.ora_advection_ring_test(dg, it, ctypes.byref(ml))
        return err, ml.value

    def function2dg(self, mesh, dg, fn):
        nx, ny = mesh.shape
        out = np.zeros(nx * ny * dg)
        _check(self.lib.ora_function2dg(mesh.h, dg, fn, _p(out)))
        return out

    def l2error_dg(self, mesh, dg, fn, phi):
        a = _f64(phi)
        return self.lib.ora_l2error_dg(mesh.h, dg, fn, _p(a))

    def mean_value(self, mesh, dg, phi):
        a = _f64(phi)
        return self.lib.ora_mean_value(mesh.h, dg, _p(a))

    def dg2cg(self, mesh, cg, dg, src):
        nx, ny = mesh.shape
        a = _f64(src)
        out = np.zeros((cg * nx + 1) * (cg * ny + 1))
        _check(self.lib.ora_dg2cg(mesh.h, cg, dg, _p(a), _p(out)))
        return out

    def cg2dg(self, mesh, cg, dg, src):
        nx, ny = mesh.shape
        a = _f64(src)
        out = np.zeros(nx * ny * dg)
        _check(self.lib.ora_cg2dg(mesh.h, cg, dg, _p(a), _p(out)))
        return out

    def limit(self, dg, phi, bound, is_max):
        out = _f64(phi).copy()
        _check(self.lib.ora_limit(dg, out.size // dg, _p(out), bound, 1 if is_max else 0))
        return out

    def benchmark_wind(self, nx, ny, L, t, xoff=0.0, yoff=0.0, vmax=30.0):
        u = np.zeros(nx * ny)
        v = np.zeros(nx * ny)
        self.lib.ora_benchmark_wind(nx, ny, L, t, xoff, yoff, vmax, _p(u), _p(v))
        return u, v
