"""Module level: MEVPDynamics / BBMDynamics (nextsimdg_amd/module.py) in a model-like time
loop with the benchmark forcing modules, against the oracle driven through the same
sequence of setData / update / getDG0Data calls the reference modules make
(MEVPDynamics.cpp:59-87, BBMDynamics.cpp:58-95).

Tolerance: 1e-10 relative L2 per output field (h, A, u, v, tau, damage) after 5 steps.
"""
import numpy as np
import pytest

from helpers import DEFAULT_PARAMS, rel_l2
from oracle import MEB_DEFAULTS, Oracle, OracleDyn

ORA = Oracle()
NX = NY = 24
DT = 120.0


def test_benchmark_atmosphere_matches_oracle_forcing():
    from nextsimdg_amd import inputs
    from nextsimdg_amd.module import BenchmarkAtmosphere, ModelStore, TimestepTime, WIND_U, WIND_V

    store = ModelStore()
    atm = BenchmarkAtmosphere(store, NX, NY, offset=(3000.0, -2000.0))
    for k in (0, 7, 360):
        atm.update(TimestepTime(1000.0 + k * DT, DT))
        uw, vw = ORA.benchmark_wind(NX, NY, inputs.L_BOX, k * DT, 3000.0, -2000.0, 30.0)
        np.testing.assert_allclose(store[WIND_U], uw, rtol=1e-13, atol=1e-15)
        np.testing.assert_allclose(store[WIND_V], vw, rtol=1e-13, atol=1e-15)


def test_dynamics_module_needs_coordinates():
    from nextsimdg_amd.module import _check_spherical

    assert _check_spherical({"x": 0, "y": 0}) is False
    assert _check_spherical({"longitude": 0, "latitude": 0, "x": 0}) is True
    with pytest.raises(RuntimeError):
        _check_spherical({"coords": 0})


def _oracle_loop(steps, bbm):
    """The reference module's per-step calls, on the oracle."""
    from nextsimdg_amd import inputs
    from nextsimdg_amd.module import MISSING_VALUE

    s = inputs.benchmark_setup(NX, NY, 0)
    o = OracleDyn(2, 6, DEFAULT_PARAMS, bbm=bbm, meb=MEB_DEFAULTS if bbm else None)
    o.initialise(NX, NY, s["coords"], s["mask"])
    h, A = s["h"], s["A"]
    zero = np.zeros(NX * NY)
    for name, v in (("hice", h), ("cice", A), ("u", zero), ("v", zero)):
        o.set_field(name, v, 1)
    d = np.where(s["mask"] == 1.0, 1.0, MISSING_VALUE)
    if bbm:
        o.set_field("damage", d, 1)
    uo, vo = inputs.ocean_gyre(NX, NY)
    out = {}
    for k in range(steps):
        uw, vw = ORA.benchmark_wind(NX, NY, inputs.L_BOX, k * DT, 0.0, 0.0, 30.0)
        for name, v in (("hice", h), ("cice", A), ("uwind", uw), ("vwind", vw), ("uocean", uo), ("vocean", vo),
                        ("ssh", zero)):
            o.set_field(name, v, 1)
        if bbm:
            o.set_field("damage", d, 1)
        o.update(DT)
        h, A = o.get_field("hice", 1), o.get_field("cice", 1)
        if bbm:
            d = o.get_field("damage", 1)
    out = {"hice": h, "cice": A, "u": o.get_field("u", 1), "v": o.get_field("v", 1),
           "taux": o.get_field("uiostress", 1), "tauy": o.get_field("viostress", 1)}
    if bbm:
        out["damage"] = d
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("bbm", [False, True])
def test_dynamics_module_benchmark_run_matches_oracle(bbm):
    from nextsimdg_amd.module import BBMDynamics, C_ICE, H_ICE, MEVPDynamics, ModelStore, benchmark_run

    print(f"module run, {'BBM' if bbm else 'mEVP'}", flush=True)
    store = ModelStore()
    dyn = (BBMDynamics if bbm else MEVPDynamics)(store)
    benchmark_run(dyn, NX, NY, 5, DT)
    ref = _oracle_loop(5, bbm)
    got = {"hice": store[H_ICE], "cice": store[C_ICE], "u": dyn.uice, "v": dyn.vice, "taux": dyn.taux,
           "tauy": dyn.tauy}
    if bbm:
        got["damage"] = dyn.damage
    errs = {n: rel_l2(got[n], ref[n]) for n in ref}
    print("  " + " ".join(f"{n}={e:.1e}" for n, e in errs.items()), flush=True)
    assert np.abs(dyn.uice).max() > 1e-3
    for n, e in errs.items():
        assert e < 1e-10, (n, e)
    st = dyn.getState()
    ocean = dyn.ocean
    assert np.all(st["u"][~ocean] == 1.7e38) and np.array_equal(st["u"][ocean], dyn.uice[ocean])
    full = dyn.getStateRecursive()
    assert full["hice"].shape[0] == NX * NY
