"""Python mirror of the reference's dynamics interfaces over the gfx950 C ABI.

Class and method names follow the reference so that code and tests read like the
reference's own:
  ParametricMesh        dynamics/src/include/ParametricMesh.hpp:42-442
  DGTransport           dynamics/src/include/DGTransport.hpp:27-152
  MEVPDynamicsKernel    dynamics/src/include/MEVPDynamicsKernel.hpp:18-37 and its bases
                        DynamicsKernel (DynamicsKernel.hpp:37-229), CGDynamicsKernel
                        (CGDynamicsKernel.hpp:26-86), VPCGDynamicsKernel
                        (VPCGDynamicsKernel.hpp:22-176)
  BBMDynamicsKernel     dynamics/src/include/BBMDynamicsKernel.hpp:18-45 over
                        BrittleCGDynamicsKernel (BrittleCGDynamicsKernel.hpp:23-256)
Arrays are numpy float64 in the ModelArray layout (element ix + nx*iy, components
contiguous). Errors raise NsdgError (the reference aborts or throws).
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import MEB_FIELDS, NsdgMEBParams, NsdgParams, check, dptr, u64ptr

PREC_FP64 = 0
PREC_FP32 = 1


def default_params(**overrides):
    """DynamicsParameters + VPParameters defaults (DynamicsParameters.hpp:34-49,
    VPParameters.hpp:22-23) and alpha = beta = 1500, 100 subcycles."""
    p = NsdgParams()
    check(_lib.load().nsdg_params_default(ctypes.byref(p)))
    for k, v in overrides.items():
        if k == "nSteps":
            k = "nsteps"
        setattr(p, k, v)
    return p


def default_meb_params(**overrides):
    """MEBParameters defaults (MEBParameters.hpp:38-70)."""
    p = NsdgMEBParams()
    check(_lib.load().nsdg_meb_params_default(ctypes.byref(p)))
    for k, v in overrides.items():
        if k not in MEB_FIELDS:
            raise KeyError(f"unknown MEB parameter {k}")
        setattr(p, k, v)
    return p


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class ParametricMesh:
    """Structured parametric quadrilateral mesh (host side; copied at use)."""

    BOTTOM, RIGHT, TOP, LEFT = 0, 1, 2, 3

    def __init__(self, nx, ny, coords, spherical=False):
        self.lib = _lib.load()
        self.nx, self.ny = int(nx), int(ny)
        self._coords = _f64(coords).reshape(-1)
        h = ctypes.c_void_p()
        check(self.lib.nsdg_mesh_create(self.nx, self.ny, dptr(self._coords), int(spherical), ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.nsdg_mesh_destroy(self.h)
            self.h = None

    def landmaskFromModelArray(self, mask):
        m = _f64(mask)
        check(self.lib.nsdg_mesh_set_landmask(self.h, dptr(m)))

    def dirichletFromMask(self):
        check(self.lib.nsdg_mesh_dirichlet_from_mask(self.h))

    def dirichletFromEdge(self, edge):
        check(self.lib.nsdg_mesh_dirichlet_from_edge(self.h, int(edge)))

    def set_dirichlet(self, edge, ids):
        a = np.ascontiguousarray(ids, dtype=np.uint64)
        check(self.lib.nsdg_mesh_set_dirichlet(self.h, int(edge), u64ptr(a), a.size))

    def add_periodic(self, entries):
        a = np.ascontiguousarray(entries, dtype=np.uint64).reshape(-1, 4)
        check(self.lib.nsdg_mesh_add_periodic(self.h, u64ptr(a), a.shape[0]))

    def dirichlet(self, edge):
        n = ctypes.c_size_t()
        check(self.lib.nsdg_mesh_get_dirichlet(self.h, int(edge), None, ctypes.byref(n)))
        out = np.zeros(n.value, dtype=np.uint64)
        if n.value:
            check(self.lib.nsdg_mesh_get_dirichlet(self.h, int(edge), u64ptr(out), ctypes.byref(n)))
        return out


class DGTransport:
    """DG upwind transport of one field (DGTransport<DG>)."""

    def __init__(self, mesh, dg):
        self.lib = _lib.load()
        self.mesh, self.dg = mesh, int(dg)
        h = ctypes.c_void_p()
        check(self.lib.nsdg_transport_create(mesh.h, self.dg, ctypes.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.nsdg_transport_destroy(self.h)
            self.h = None

    def settimesteppingscheme(self, scheme):
        check(self.lib.nsdg_transport_set_scheme(self.h, scheme.encode()))

    def set_velocity(self, vx, vy):
        """GetVx() = vx; GetVy() = vy with N_e x DG coefficient arrays."""
        a, b = _f64(vx), _f64(vy)
        check(self.lib.nsdg_transport_set_velocity(self.h, dptr(a), dptr(b)))

    def prepareAdvection(self, cg, u, v):
        a, b = _f64(u), _f64(v)
        check(self.lib.nsdg_transport_prepare_advection(self.h, int(cg), dptr(a), dptr(b)))

    def reinitnormalvelocity(self):
        check(self.lib.nsdg_transport_reinit_normal_velocity(self.h))

    def step(self, dt, phi):
        out = _f64(phi).copy()
        check(self.lib.nsdg_transport_step(self.h, float(dt), dptr(out)))
        return out

    def run(self, dt, phi, nsteps, limit_max=False, limit_min=False):
        """nsteps x {reinitnormalvelocity; step; [LimitMax(phi, 1)]; [LimitMin(phi, 0)]}."""
        out = _f64(phi).copy()
        flags = (1 if limit_max else 0) | (2 if limit_min else 0)
        check(self.lib.nsdg_transport_run(self.h, float(dt), dptr(out), int(nsteps), flags))
        return out


def LimitMax(dg, phi, value):
    out = _f64(phi).copy()
    check(_lib.load().nsdg_limit(int(dg), out.size // dg, dptr(out), float(value), 1))
    return out


def LimitMin(dg, phi, value):
    out = _f64(phi).copy()
    check(_lib.load().nsdg_limit(int(dg), out.size // dg, dptr(out), float(value), 0))
    return out


class MEVPDynamicsKernel:
    """mEVP dynamics kernel (MEVPDynamicsKernel<DG>) with device-resident state."""

    PASSES = ("strain", "stress", "divergence", "momentum", "boundaries")

    def __init__(self, cg_degree=2, dg_adv=6, params=None, **overrides):
        self.lib = _lib.load()
        self.cg, self.dg = int(cg_degree), int(dg_adv)
        self.S = 3 if self.cg == 1 else 8
        p = params if params is not None else default_params(**overrides)
        h = ctypes.c_void_p()
        self._create(p, h)
        self.h = h
        self.params = p
        self.nx = self.ny = None

    def _create(self, p, h):
        check(self.lib.nsdg_dyn_create(self.cg, self.dg, ctypes.byref(p), ctypes.byref(h)))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.nsdg_dyn_destroy(self.h)
            self.h = None

    # ---- DynamicsKernel interface
    def initialise(self, coords, isSpherical, mask, nx=None, ny=None):
        c, m = _f64(coords).reshape(-1), _f64(mask).reshape(-1)
        if nx is None:
            raise ValueError("nx, ny are needed (the reference reads them from ModelArray dimensions)")
        self.nx, self.ny = int(nx), int(ny)
        check(self.lib.nsdg_dyn_initialise(self.h, self.nx, self.ny, dptr(c), int(isSpherical), dptr(m)))

    def initialise_mesh(self, mesh):
        check(self.lib.nsdg_dyn_initialise_mesh(self.h, mesh.h))
        self.nx, self.ny = mesh.nx, mesh.ny

    @property
    def ne(self):
        return self.nx * self.ny

    @property
    def ncg(self):
        return self.lib.nsdg_dyn_cg_nodes(self.h)

    def setData(self, name, data, ncomp=None):
        a = _f64(data)
        if ncomp is None:
            ncomp = a.size // self.ne
        check(self.lib.nsdg_dyn_set_field(self.h, name.encode(), dptr(a), int(ncomp)))

    def getDG0Data(self, name):
        out = np.zeros(self.ne)
        check(self.lib.nsdg_dyn_get_field(self.h, name.encode(), dptr(out), 1))
        return out

    def getDGData(self, name, ncomp=None):
        if ncomp is None:
            ncomp = self.S if name[0] in "se" and name[1:] in ("11", "12", "22") else self.dg
        out = np.zeros(self.ne * ncomp)
        check(self.lib.nsdg_dyn_get_field(self.h, name.encode(), dptr(out), int(ncomp)))
        return out.reshape(self.ne, ncomp)

    def update(self, dt):
        """VPCGDynamicsKernel::update with TimestepTime.step.seconds() == dt."""
        check(self.lib.nsdg_dyn_update(self.h, float(dt)))

    # ---- per-pass and support entries
    def get_cg(self, name):
        out = np.zeros(self.ncg)
        check(self.lib.nsdg_dyn_get_cg(self.h, name.encode(), dptr(out)))
        return out

    def set_cg(self, name, data):
        a = _f64(data)
        check(self.lib.nsdg_dyn_set_cg(self.h, name.encode(), dptr(a)))

    def advectionAndLimits(self, dt):
        check(self.lib.nsdg_dyn_advect(self.h, float(dt)))

    def prepareIteration(self):
        check(self.lib.nsdg_dyn_prepare_iteration(self.h))

    def begin_iteration(self, dt):
        check(self.lib.nsdg_dyn_begin_iteration(self.h, float(dt)))

    def projectVelocityToStrain(self):
        check(self.lib.nsdg_dyn_strain(self.h))

    def stressUpdateHighOrder(self, repeat=1):
        check(self.lib.nsdg_dyn_stress_update(self.h, int(repeat)))

    def stressDivergence(self):
        check(self.lib.nsdg_dyn_divergence(self.h))

    def updateMomentum(self):
        check(self.lib.nsdg_dyn_momentum(self.h))

    def applyBoundaries(self):
        check(self.lib.nsdg_dyn_apply_boundaries(self.h))

    def subcycles(self, n):
        check(self.lib.nsdg_dyn_subcycles(self.h, int(n)))

    def set_precision(self, prec):
        check(self.lib.nsdg_dyn_set_precision(self.h, int(prec)))

    def set_pairing(self, mode):
        """0 (auto, by mesh size), 1 (two subcycles per pass), 2 (one per pass)."""
        check(self.lib.nsdg_dyn_set_pairing(self.h, int(mode)))

    def benchmark_wind(self, L, t, xoff=0.0, yoff=0.0, vmax=30.0):
        check(self.lib.nsdg_dyn_benchmark_wind(self.h, float(L), float(t), float(xoff), float(yoff), float(vmax)))

    def storm_wind(self, L, t, storms):
        """The sum of benchmark cyclones, storms = [(xoff, yoff, vmax), ...] (at most 8)."""
        a = np.ascontiguousarray(np.asarray(storms, dtype=np.float64).reshape(-1, 3).T)
        check(self.lib.nsdg_dyn_storm_wind(self.h, float(L), float(t), dptr(a[0]), dptr(a[1]), dptr(a[2]),
                                           a.shape[1]))

    def reset_dg0(self, name):
        check(self.lib.nsdg_dyn_reset_dg0(self.h, name.encode()))

    def mass(self):
        """Tools::MeanValue of hice and cice (Tools.hpp:24-45): (ice volume, ice area)."""
        v, a = ctypes.c_double(), ctypes.c_double()
        check(self.lib.nsdg_dyn_mass(self.h, ctypes.byref(v), ctypes.byref(a)))
        return v.value, a.value

    def mass_noalloc(self):
        """mass() with no allocation in the call (partial sums in the library's fixed
        device array; at most 8,386,560 elements)."""
        v, a = ctypes.c_double(), ctypes.c_double()
        check(self.lib.nsdg_dyn_mass_noalloc(self.h, ctypes.byref(v), ctypes.byref(a)))
        return v.value, a.value

    def write_restart(self, path, time_seconds=0.0):
        """Write the state as a restart file in the layout of ParaGridIO::dumpModelState
        (netCDF classic, 64-bit offsets; see include/nsdg_dyn.h)."""
        check(self.lib.nsdg_dyn_write_restart(self.h, os.fsencode(path), float(time_seconds)))

    def read_restart(self, path):
        """Load a restart file written for a kernel like this one; returns its time."""
        t = ctypes.c_double()
        check(self.lib.nsdg_dyn_read_restart(self.h, os.fsencode(path), ctypes.byref(t)))
        return t.value

    def set_budget(self, on=True):
        """Switch the advection mass budget on (zeroed) or off."""
        check(self.lib.nsdg_dyn_set_budget(self.h, int(bool(on))))

    def budget(self, reset=False):
        """Advection mass budget since it was switched on or reset: the Dirichlet outflow
        and the limiters' change of the MeanValue of h and A, such that
        mass(after) - mass(before) = limiter - outflow for each."""
        out = np.zeros(4)
        check(self.lib.nsdg_dyn_budget(self.h, dptr(out), int(bool(reset))))
        return dict(outflow_h=out[0], outflow_A=out[1], limiter_h=out[2], limiter_A=out[3])

    def momentum_residual(self):
        """(r, |R|, |M_L A tau_a|): the fixed-point residual of the momentum update with the
        converged VP stress of the current velocity, over the non-Dirichlet nodes."""
        r, nr, nd = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        check(self.lib.nsdg_dyn_residual(self.h, ctypes.byref(r), ctypes.byref(nr), ctypes.byref(nd)))
        return r.value, nr.value, nd.value

    def stress_from_velocity(self):
        """stress <- Pi sigma_VP(eps(u)), the converged mEVP stress of the current velocity."""
        check(self.lib.nsdg_dyn_stress_from_velocity(self.h))

    def synchronize(self):
        check(self.lib.nsdg_dyn_synchronize(self.h))

    def save_state(self):
        n = ctypes.c_size_t()
        check(self.lib.nsdg_dyn_save_state(self.h, None, ctypes.byref(n)))
        buf = ctypes.create_string_buffer(n.value)
        check(self.lib.nsdg_dyn_save_state(self.h, buf, ctypes.byref(n)))
        return buf.raw

    def load_state(self, blob):
        buf = ctypes.create_string_buffer(blob, len(blob))
        check(self.lib.nsdg_dyn_load_state(self.h, buf, len(blob)))

    def set_profiling(self, on):
        check(self.lib.nsdg_dyn_set_profiling(self.h, int(bool(on))))

    def kernel_stats(self, name):
        c, t = ctypes.c_int(), ctypes.c_double()
        check(self.lib.nsdg_dyn_kernel_stats(self.h, name.encode(), ctypes.byref(c), ctypes.byref(t)))
        return c.value, t.value


class BBMDynamicsKernel(MEVPDynamicsKernel):
    """Brittle (BBM) dynamics kernel (BBMDynamicsKernel<DG>) with device-resident state.

    update(dt) advects h and A with the velocity averaged over the previous step's
    substeps, transports the stress and the damage, and runs nSteps brittle substeps of
    dt / nSteps (BrittleCGDynamicsKernel.hpp:92-134). The extra DG field "damage" is set
    and read like hice; the averaged velocity is the CG vector "avgU" / "avgV". The
    mEVP-only facilities (momentum residual, FP32 variant, learned warm start) raise.
    Uniform meshes only.
    """

    def __init__(self, cg_degree=2, dg_adv=6, params=None, meb=None, **overrides):
        self.meb = meb if meb is not None else default_meb_params()
        super().__init__(cg_degree, dg_adv, params, **overrides)

    def _create(self, p, h):
        check(self.lib.nsdg_bbm_create(self.cg, self.dg, ctypes.byref(p), ctypes.byref(self.meb), ctypes.byref(h)))
