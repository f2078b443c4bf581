"""CPU-side checks of the C ABI: the gfx950 library loads, exports every symbol the
header declares, and reports errors through status codes (no GPU compute here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from nextsimdg_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", h) for h in ("nsdg_dyn.h", "nsdg_cnn.h")]


def declared_symbols():
    text = "".join(open(h).read() for h in HEADERS)
    return sorted(set(re.findall(r"^\s*(?:int|size_t|const char\*|nsdg_cnn\*)\s+(nsdg_\w+)\s*\(", text, re.M)))


def test_header_declares_the_abi():
    syms = declared_symbols()
    assert len(syms) >= 40
    assert set(syms) == set(_lib.SIGNATURES), set(syms) ^ set(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for s in declared_symbols():
        assert hasattr(lib, s), s


def test_library_is_built_for_gfx950_only():
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data
    for other in (b"gfx942", b"gfx90a", b"gfx1100"):
        assert other not in data


def test_params_default_matches_reference_constants():
    from nextsimdg_amd import default_params

    p = default_params()
    # DynamicsParameters.hpp:34-49, VPParameters.hpp:22-23, VPCGDynamicsKernel.hpp:125-126
    assert p.rho_ice == 900.0 and p.rho_atm == 1.3 and p.rho_ocean == 1026.0
    assert p.F_atm == 1.2e-3 * 1.3 and p.F_ocean == 5.5e-3 * 1026.0
    assert p.fc == 1.45842e-4 and p.gravity == 9.81 and p.ocean_turning_angle == 0.0
    assert p.Pstar == 27500.0 and p.DeltaMin == 2e-9
    assert p.alpha == 1500.0 and p.beta == 1500.0 and p.nsteps == 100


def test_mesh_boundary_lists_on_host_match_oracle():
    """Mesh set-up is host integer work: compare the library's lists with the oracle's
    on the 25 km NH mesh (ParametricMesh_test.cpp:84-139)."""
    from helpers import read_nh_smesh
    from nextsimdg_amd import ParametricMesh
    from oracle import OracleMesh

    f = OracleMesh.from_smesh(read_nh_smesh())
    nx, ny = f.shape
    m = ParametricMesh(nx, ny, f.vertices())
    m.landmaskFromModelArray(f.landmask().astype(np.float64))
    m.dirichletFromMask()
    for e in range(4):
        m.dirichletFromEdge(e)
    for e in range(4):
        np.testing.assert_array_equal(m.dirichlet(e), f.dirichlet(e))


def test_errors_come_back_as_status_codes():
    lib = _lib.load()
    h = ctypes.c_void_p()
    coords = np.zeros(8)
    # a 1 x 1 mesh is too small; spherical and Cartesian 2 x 2 meshes are fine (host only)
    assert lib.nsdg_mesh_create(1, 1, _lib.dptr(coords), 0, ctypes.byref(h)) == -1
    assert b"2 x 2" in lib.nsdg_last_error()
    c = np.zeros(18)
    assert lib.nsdg_mesh_create(2, 2, _lib.dptr(c), 1, ctypes.byref(h)) == 0
    lib.nsdg_mesh_destroy(h)
    assert lib.nsdg_mesh_create(2, 2, _lib.dptr(c), 0, ctypes.byref(h)) == 0
    bad = np.array([99], dtype=np.uint64)
    assert lib.nsdg_mesh_set_dirichlet(h, 0, _lib.u64ptr(bad), 1) == -1
    assert lib.nsdg_mesh_set_dirichlet(h, 7, _lib.u64ptr(bad), 0) == -1
    lib.nsdg_mesh_destroy(h)
    assert lib.nsdg_dyn_get_field(None, b"u", None, 1) == -1


@pytest.mark.skipif(_lib.load().nsdg_device_count() > 0, reason="GPU present")
def test_no_device_means_error_not_fallback():
    from nextsimdg_amd import MEVPDynamicsKernel
    from nextsimdg_amd._lib import NsdgError

    with pytest.raises(NsdgError):
        MEVPDynamicsKernel(2, 6)
