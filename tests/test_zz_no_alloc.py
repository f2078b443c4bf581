"""nsdg_dyn_mass_noalloc, the allocation-free variant of nsdg_dyn_mass (its block partials
live in a fixed 1 MiB device array of the library instead of the kernel's own buffers),
called through the C ABI: two calls in a row, the second on a smaller mesh, each equal to
nsdg_dyn_mass on the same state (the same kernels in the same order: bit for bit), and a
mesh above the array's capacity refused with NSDG_ERR_ARG."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(k):
    from nextsimdg_amd import _lib

    v, a = ctypes.c_double(), ctypes.c_double()
    rc = _lib.load().nsdg_dyn_mass_noalloc(k.h, ctypes.byref(v), ctypes.byref(a))
    return rc, v.value, a.value


def _kernel(nx, ny, seed, spherical=False):
    import nextsimdg_amd as nd
    from nextsimdg_amd import inputs

    coords = inputs.sphere_patch_coords(nx, ny) if spherical else inputs.box_coords(nx, ny)
    k = nd.MEVPDynamicsKernel(2, 6)
    k.initialise(coords, spherical, inputs.box_mask(nx, ny), nx, ny)
    rng = np.random.default_rng(seed)
    k.setData("hice", rng.normal(1.0, 0.2, (nx * ny, 6)), 6)
    k.setData("cice", rng.uniform(0.0, 1.0, (nx * ny, 6)), 6)
    return k


def test_mass_noalloc_equals_mass_on_two_calls():
    for nx, ny, seed, sph in ((1024, 768, 1, False), (200, 130, 2, True)):
        print(f"mass_noalloc {nx}x{ny} spherical={sph}", flush=True)
        k = _kernel(nx, ny, seed, sph)
        rc, v, a = _noalloc(k)
        assert rc == 0
        vr, ar = k.mass()
        assert (v, a) == (vr, ar), (v, a, vr, ar)
        assert v > 0 and a > 0


def test_mass_noalloc_refuses_meshes_above_its_capacity():
    import nextsimdg_amd as nd

    print("mass_noalloc capacity", flush=True)
    k = _kernel(2900, 2900, 3)  # 8,410,000 elements > 8,386,560
    rc, _, _ = _noalloc(k)
    assert rc == -1  # NSDG_ERR_ARG
    from nextsimdg_amd import _lib

    assert b"at most" in _lib.load().nsdg_last_error()
    # the kernel is still usable afterwards
    assert np.isfinite(k.mass()[0])
    del nd
