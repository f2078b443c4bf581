"""Module-level mirror of the reference's dynamics modules over the gfx950 kernels.

  IDynamics            core/src/modules/include/IDynamics.hpp:16-135
  MEVPDynamics         core/src/modules/DynamicsModule/MEVPDynamics.cpp:28-87
  BBMDynamics          core/src/modules/DynamicsModule/BBMDynamics.cpp:16-98
  BenchmarkAtmosphere  physics/src/modules/AtmosphereBoundaryModule/BenchmarkAtmosphere.cpp:29-76
  BenchmarkOcean       physics/src/modules/OceanBoundaryModule/BenchmarkOcean.cpp:15-37

The reference's ModelArrayRef store is a plain dict here (ModelStore) keyed by the
store's field names; H-type fields are float64 arrays of nx * ny values, element
ix + nx * iy (x fastest), as ModelArray holds them. Every update sets h, A (and the
damage) from the store as DG0 fields, which zeroes their higher DG moments on the
device (the round trip MEVPDynamics::update makes through getDG0Data / setData), runs
the kernel's update, and writes the cell means back.
"""
import math

import numpy as np

from . import inputs
from .dynamics import BBMDynamicsKernel, MEVPDynamicsKernel

# Store names (core/src/include/ModelArrayRef.hpp, Shared:: and Protected:: fields)
H_ICE, C_ICE, H_SNOW = "H_ICE", "C_ICE", "H_SNOW"
DAMAGE, DAMAGE_IN = "DAMAGE", "DAMAGE0"  # Shared::DAMAGE (written), Protected::DAMAGE (read)
WIND_U, WIND_V, OCEAN_U, OCEAN_V, SSH = "WIND_U", "WIND_V", "OCEAN_U", "OCEAN_V", "SSH"

RADIANS = float.fromhex("0x1.1df46a2529d39p-6")  # degrees to radians (MEVPDynamics.cpp:21)
MISSING_VALUE = 1.7e38  # MissingData::defaultValue (MissingData.hpp:18)


class ModelStore(dict):
    """The shared and protected fields the modules exchange (ModelArrayReferenceStore)."""


class TimestepTime:
    """Start (seconds since the run's reference time) and length of a timestep."""

    def __init__(self, start, step):
        self.start, self.step = float(start), float(step)


def _check_spherical(ms):
    """IDynamics::checkSpherical (IDynamics.hpp:113-126)."""
    if "longitude" in ms and "latitude" in ms:
        return True
    if "x" in ms and "y" in ms:
        return False
    raise RuntimeError("Input data must contain either Cartesian (x, y) or spherical (longitude, latitude) "
                       "coordinates.")


class IDynamics:
    """Shared part of the dynamics modules: ice velocity, ice-ocean stress and damage
    outputs (H fields), the store references, and the ocean mask of ModelComponent::mask."""

    uses_damage = False
    kernel_class = MEVPDynamicsKernel

    def __init__(self, store, cg_degree=2, dg_adv=6, params=None, **kernel_args):
        self.store = store
        self.kernel = self.kernel_class(cg_degree, dg_adv, params=params, **kernel_args)
        self.uice = self.vice = self.taux = self.tauy = self.damage = None
        self.ocean = None

    def mask(self, data):
        """ModelComponent::mask (ModelComponent.cpp:95-107): land cells get the missing value."""
        return np.where(self.ocean, data, MISSING_VALUE)

    def setData(self, ms):
        """MEVPDynamics::setData (MEVPDynamics.cpp:37-57): initialise the kernel from the
        coordinates and mask, then set h, A, u, v. Spherical coordinates (longitude,
        latitude in degrees) are converted to radians for the kernel."""
        spherical = _check_spherical(ms)
        nx, ny = ms["nx"], ms["ny"]
        coords = np.asarray(ms["coords"], dtype=np.float64)
        if spherical:
            coords = coords * RADIANS
        mask = np.asarray(ms["mask"], dtype=np.float64).ravel()
        self.ocean = mask == 1.0
        self.kernel.initialise(coords.reshape(-1, 2), spherical, mask, nx, ny)
        ne = nx * ny
        self.uice = np.asarray(ms["u"], dtype=np.float64).ravel().copy()
        self.vice = np.asarray(ms["v"], dtype=np.float64).ravel().copy()
        self.damage = np.zeros(ne) if not self.uses_damage else np.ones(ne)
        for name in ("hice", "cice", "u", "v"):
            self.kernel.setData(name, np.asarray(ms[name], dtype=np.float64).ravel(), 1)

    def _set_inputs(self):
        s, k = self.store, self.kernel
        k.setData("hice", s[H_ICE], 1)
        k.setData("cice", s[C_ICE], 1)
        for name, key in (("uwind", WIND_U), ("vwind", WIND_V), ("uocean", OCEAN_U), ("vocean", OCEAN_V),
                          ("ssh", SSH)):
            k.setData(name, s[key], 1)

    def _get_outputs(self):
        s, k = self.store, self.kernel
        s[H_ICE] = k.getDG0Data("hice")
        s[C_ICE] = k.getDG0Data("cice")
        self.uice = k.getDG0Data("u")
        self.vice = k.getDG0Data("v")
        self.taux = k.getDG0Data("uiostress")
        self.tauy = k.getDG0Data("viostress")

    def update(self, tst):
        """MEVPDynamics::update (MEVPDynamics.cpp:59-87)."""
        self._set_inputs()
        self.kernel.update(tst.step)
        self._get_outputs()

    def getState(self):
        """IDynamics::getState (IDynamics.hpp:40-47): the masked ice velocity."""
        return {"u": self.mask(self.uice), "v": self.mask(self.vice)}

    def getStateRecursive(self, all_components=True):
        """MEVPDynamics::getStateRecursive (MEVPDynamics.cpp:89-101)."""
        st = self.getState()
        if all_components:
            st.update({"hice": self.kernel.getDG0Data("hice"), "cice": self.kernel.getDG0Data("cice")})
        return st


class MEVPDynamics(IDynamics):
    """The mEVP dynamics module (MEVPDynamics.cpp); no damage (IDynamics::setData sets it to 0)."""


class BBMDynamics(IDynamics):
    """The brittle dynamics module (BBMDynamics.cpp): damage read from the store as a DG0
    field every step and written back as the cell mean; an absent initial damage defaults
    to 1 (undamaged) on ocean cells and the missing value on land."""

    uses_damage = True
    kernel_class = BBMDynamicsKernel

    def setData(self, ms):
        """BBMDynamics::setData (BBMDynamics.cpp:20-56)."""
        super().setData(ms)
        if "damage" in ms:
            d = np.asarray(ms["damage"], dtype=np.float64).ravel()
        else:
            d = self.mask(np.ones(self.kernel.ne))
        self.kernel.setData("damage", d, 1)
        self.damage = d.copy()

    def update(self, tst):
        """BBMDynamics::update (BBMDynamics.cpp:58-95)."""
        self.damage = np.asarray(self.store[DAMAGE_IN], dtype=np.float64).copy()
        self._set_inputs()
        self.kernel.setData("damage", self.damage, 1)
        self.kernel.update(tst.step)
        self._get_outputs()
        self.damage = self.kernel.getDG0Data("damage")
        self.store[DAMAGE] = self.damage

    def getStateRecursive(self, all_components=True):
        """BBMDynamics::getStateRecursive (BBMDynamics.cpp:114-127): the full DG fields."""
        st = self.getState()
        if all_components:
            st.update({n: self.kernel.getDGData(n) for n in ("hice", "cice", "damage")})
        return st


class BenchmarkAtmosphere:
    """The moving cyclone wind (BenchmarkAtmosphere.cpp:29-76), at the cell lower-left
    corners, timed from the first update's start."""

    def __init__(self, store, nx, ny, L=inputs.L_BOX, vmax=30.0, offset=(0.0, 0.0)):
        self.store, self.nx, self.ny, self.L, self.vmax, self.offset = store, nx, ny, L, vmax, offset
        self.t0 = None

    def update(self, tst):
        if self.t0 is None:
            self.t0 = tst.start
        frac = (tst.start - self.t0) / 86400.0
        dx, dy = self.L / self.nx, self.L / self.ny
        x0 = self.nx * dx * 0.5 * (1 + frac / 5.0) + self.offset[0]
        y0 = self.ny * dy * 0.5 * (1 + frac / 5.0) + self.offset[1]
        x, y = inputs.cell_corners(self.nx, self.ny, self.L)
        xp, yp = x - x0, y - y0
        ca, sa = math.cos(72.0 / 180.0 * math.pi), math.sin(72.0 / 180.0 * math.pi)
        scale = 1e-5 * np.exp(-1e-5 * np.hypot(xp, yp))
        self.store[WIND_U] = -scale * self.vmax * (ca * xp + sa * yp)
        self.store[WIND_V] = -scale * self.vmax * (-sa * xp + ca * yp)


class BenchmarkOcean:
    """The steady ocean gyre and a flat sea surface (BenchmarkOcean.cpp:15-37)."""

    def __init__(self, store, nx, ny, L=inputs.L_BOX):
        self.store, self.nx, self.ny, self.L = store, nx, ny, L

    def update(self, tst):
        u, v = inputs.ocean_gyre(self.nx, self.ny, self.L)
        self.store[OCEAN_U], self.store[OCEAN_V] = u, v
        self.store[SSH] = np.zeros(self.nx * self.ny)


def benchmark_run(dynamics, nx, ny, steps, dt=120.0, h=None, A=None, seed=0, offset=(0.0, 0.0)):
    """A model-like time loop of the benchmark box: forcing modules, then the dynamics,
    each step; the damage the dynamics writes is its next step's input (no healing).
    Returns the store."""
    store = dynamics.store
    setup = inputs.benchmark_setup(nx, ny, seed)
    h = setup["h"] if h is None else h
    A = setup["A"] if A is None else A
    ms = {"nx": nx, "ny": ny, "x": None, "y": None, "coords": setup["coords"], "mask": setup["mask"],
          "hice": h, "cice": A, "u": np.zeros(nx * ny), "v": np.zeros(nx * ny)}
    dynamics.setData(ms)
    store[H_ICE], store[C_ICE] = h.copy(), A.copy()
    store[H_SNOW] = np.zeros(nx * ny)
    store[DAMAGE_IN] = dynamics.damage.copy()
    atm = BenchmarkAtmosphere(store, nx, ny, offset=offset)
    ocn = BenchmarkOcean(store, nx, ny)
    for k in range(steps):
        tst = TimestepTime(k * dt, dt)
        atm.update(tst)
        ocn.update(tst)
        dynamics.update(tst)
        if dynamics.uses_damage:
            store[DAMAGE_IN] = store[DAMAGE]
    return store
