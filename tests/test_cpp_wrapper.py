"""The header-only C++ wrapper (include/nsdg/nsdg_dyn.hpp) builds against the library
and, on the GPU, drives a module-level run whose output fields match the FP64 oracle's
cell by cell (1e-10 relative L2 per field)."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "tests", "cpp")


def build_demo(name="wrapper_demo"):
    exe = os.path.join(CPP, name)
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-I" + os.path.join(ROOT, "include"),
                    os.path.join(CPP, name + ".cpp"), "-L" + os.path.join(ROOT, "nextsimdg_amd"), "-lnsdg_hip",
                    "-Wl,-rpath," + os.path.join(ROOT, "nextsimdg_amd"), "-o", exe], check=True)
    return exe


def _parse_fields(text):
    """Lines 'name v0 v1 ...' of a demo's output, as arrays."""
    out = {}
    for line in text.strip().splitlines():
        name, *vals = line.split()
        out[name] = np.array(list(map(float, vals)))
    return out


def _compare_fields(got, ref):
    """Every field whole against the oracle: 1e-10 relative L2 over the cells. Cells
    holding the missing value (1.7e38, the brittle damage on land) must hold it on both
    sides and are left out of the norm."""
    from helpers import rel_l2

    assert set(got) == set(ref), (set(got), set(ref))
    for n, want in ref.items():
        g = got[n]
        assert g.shape == want.shape, n
        miss = want == 1.7e38
        assert np.array_equal(g == 1.7e38, miss), n
        err = rel_l2(g[~miss], want[~miss])
        assert err < 1e-10, (n, err)


@pytest.mark.parametrize("name", ["wrapper_demo", "module_demo"])
def test_cpp_wrapper_compiles_and_links(name):
    assert os.path.exists(build_demo(name))


@pytest.mark.gpu
@pytest.mark.parametrize("rheology", ["mevp", "bbm"])
def test_cpp_wrapper_run_matches_oracle(rheology):
    from oracle import Oracle, OracleDyn

    print(f"C++ wrapper run, {rheology}", flush=True)
    out = subprocess.run([build_demo(), "2", rheology], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = _parse_fields(out.stdout)
    nx = ny = 24
    L, dx = 512e3, 512e3 / 24
    j, i = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    coords = np.stack([(i * dx).ravel(), (j * dx).ravel()], 1)
    jc, ic = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    ic, jc = ic.ravel(), jc.ravel()
    mask = np.where((ic == 0) | (jc == 0) | (ic == nx - 1) | (jc == ny - 1), 0.0, 1.0)
    h = 0.3 + 0.005 * (np.sin(6e-5 * ic * dx) + np.sin(3e-5 * jc * dx))
    a = np.ones(nx * ny)
    o = OracleDyn(2, 6, bbm=rheology == "bbm")
    o.initialise(nx, ny, coords, mask)
    o.set_field("uocean", 0.01 * (2.0 * jc / ny - 1.0), 1)
    o.set_field("vocean", 0.01 * (1.0 - 2.0 * ic / nx), 1)
    ora = Oracle()
    for s in range(2):
        o.set_field("hice", h, 1)
        o.set_field("cice", a, 1)
        o.set_field("ssh", 0 * h, 1)
        uw, vw = ora.benchmark_wind(nx, ny, L, s * 120.0, 0.0, 0.0, 30.0)
        o.set_field("uwind", uw, 1)
        o.set_field("vwind", vw, 1)
        o.update(120.0)
        h, a = o.get_field("hice", 1), o.get_field("cice", 1)
    ref = {"u": o.get_field("u", 1), "v": o.get_field("v", 1), "hice": h, "cice": a}
    if rheology == "bbm":
        ref["damage"] = o.get_field("damage", 1)
    _compare_fields(got, ref)


@pytest.mark.gpu
@pytest.mark.parametrize("rheology", ["mevp", "bbm"])
def test_cpp_module_run_matches_oracle(rheology):
    """nsdg_module.hpp: MEVPDynamics / BBMDynamics with the benchmark forcing modules."""
    from oracle import MEB_DEFAULTS, Oracle, OracleDyn
    from nextsimdg_amd import inputs

    print(f"C++ module run, {rheology}", flush=True)
    out = subprocess.run([build_demo("module_demo"), "3", rheology], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    got = _parse_fields(out.stdout)
    nx = ny = 24
    bbm = rheology == "bbm"
    coords, mask = inputs.box_coords(nx, ny), inputs.box_mask(nx, ny)
    x, y = inputs.cell_corners(nx, ny)
    h = 0.3 + 0.005 * (np.sin(60e-3 * x / 1e3) + np.sin(30e-3 * y / 1e3))
    a = np.ones(nx * ny)
    o = OracleDyn(2, 6, bbm=bbm, meb=MEB_DEFAULTS if bbm else None)
    o.initialise(nx, ny, coords, mask)
    zero = np.zeros(nx * ny)
    d = np.where(mask == 1.0, 1.0, 1.7e38)
    for name, v in (("hice", h), ("cice", a), ("u", zero), ("v", zero)):
        o.set_field(name, v, 1)
    uo, vo = inputs.ocean_gyre(nx, ny)
    ora = Oracle()
    for s in range(3):
        uw, vw = ora.benchmark_wind(nx, ny, inputs.L_BOX, s * 120.0, 0.0, 0.0, 30.0)
        for name, v in (("hice", h), ("cice", a), ("uwind", uw), ("vwind", vw), ("uocean", uo), ("vocean", vo),
                        ("ssh", zero)) + ((("damage", d),) if bbm else ()):
            o.set_field(name, v, 1)
        o.update(120.0)
        h, a = o.get_field("hice", 1), o.get_field("cice", 1)
        if bbm:
            d = o.get_field("damage", 1)
    ref = {"hice": h, "cice": a, "u": o.get_field("u", 1), "v": o.get_field("v", 1),
           "uiostress": o.get_field("uiostress", 1), "viostress": o.get_field("viostress", 1)}
    if bbm:
        ref["damage"] = d
    _compare_fields(got, ref)
