"""Pins of the oracle's BBM (brittle) rheology on the CPU. The reference ships no test for
it, so each part is checked against a closed form or an invariant:

  * elastic increment and viscous factor (BBMStressUpdateStep.hpp:67-118) on a uniform
    strain from a linear velocity, where every Gauss point sees the same values;
  * the compressive branch tildeP = min(Pmax / -sigma_n, 1) (Eqn. 7b);
  * the Mohr-Coulomb damage step, d <- d - d (1 - dcrit) dt / td (Eqns. 22, 29, 30);
  * the brittle momentum update (BrittleCGDynamicsKernel.hpp:200-253) against an
    independent 2x2 solve of its semi-implicit system, with the reference's sign of the
    beta (gradX + tauX) term of v as the one deviation;
  * the DG8 stress transport keeps a constant field exactly in the interior.
"""
import numpy as np
import pytest

from helpers import DEFAULT_PARAMS, box_coords, cg_node_coords
from oracle import MEB_DEFAULTS, OracleDyn, OracleMesh, OracleTransport

DX = 2000.0
NX = NY = 6


def _bbm(meb=None, nsteps=10):
    p = dict(DEFAULT_PARAMS, nSteps=nsteps)
    o = OracleDyn(2, 6, p, bbm=True, meb=meb)
    o.initialise(NX, NY, box_coords(NX, NY, DX), np.ones(NX * NY))
    o.set_field("hice", np.ones(NX * NY), 1)
    o.set_field("cice", np.ones(NX * NY), 1)
    return o, p, dict(MEB_DEFAULTS, **(meb or {}))


def _linear_velocity(o, a, d):
    x, y = cg_node_coords(NX, NY, DX)
    o.set_cg("u", a * x)
    o.set_cg("v", d * y)


def _elastic(m, dt, e11, e22, s0=0.0, h=1.0, A=1.0, dmg=1.0):
    expC = np.exp(m["compaction_param"] * (1 - A))
    tv = m["undamaged_time_relaxation_sigma"] * (dmg * expC) ** (m["exponent_relaxation_sigma"] - 1)
    pmax = m["P0"] * h ** (m["exponent_compression_factor"] + 1) * expC
    tildeP = min(pmax / -s0, 1.0) if s0 < 0 else 0.0
    mult = tv / (tv + (1 - tildeP) * dt)
    E = h * m["young"] * dmg * expC
    nu = m["nu0"]
    du = dt * E / (1 - nu * nu)
    return mult * (s0 + du * (e11 + nu * e22)), mult * (s0 + du * (nu * e11 + e22)), E


@pytest.mark.parametrize("s0", [0.0, -2.0e4, -5.0e3])
def test_bbm_elastic_step_on_uniform_strain(s0):
    meb = {"C_lab": 1e15, "compr_strength": 1e20}  # no failure: the elastic-viscous part alone
    o, p, m = _bbm(meb)
    a, d = 1e-7, 2.5e-7
    _linear_velocity(o, a, d)
    for n in ("s11", "s22"):
        o.set_field(n, np.full(NX * NY, s0), 1)
    o.begin_iteration(10.0)
    o.run_pass("strain")
    o.run_pass("stress")
    s11, s22, _ = _elastic(m, 10.0, a, d, s0)
    got11, got22 = o.get_field("s11", 8), o.get_field("s22", 8)
    np.testing.assert_allclose(got11[:, 0], s11, rtol=1e-12)
    np.testing.assert_allclose(got22[:, 0], s22, rtol=1e-12)
    assert np.abs(got11[:, 1:]).max() < 1e-9 * abs(s11)
    np.testing.assert_allclose(o.get_field("s12", 8), 0.0, atol=1e-9 * abs(s11))
    dmg = o.get_field("damage", 6)
    np.testing.assert_allclose(dmg[:, 0], 1.0, rtol=1e-14)


def test_bbm_damage_step_closed_form():
    meb = {"C_lab": 2.0e4}  # weak enough that the elastic stress leaves the envelope
    o, p, m = _bbm(meb)
    a, d = 4e-6, -1e-6
    _linear_velocity(o, a, d)
    o.begin_iteration(10.0)
    o.run_pass("strain")
    o.run_pass("stress")
    s11, s22, E = _elastic(m, 10.0, a, d)
    sn, tau = 0.5 * (s11 + s22), 0.5 * abs(s11 - s22)
    hE = DX  # sqrt of the area of a square element
    scale = np.sqrt(0.1 / hE)
    dcrit = min(m["C_lab"] * scale / (tau + m["tan_phi"] * sn), 1.0)
    assert dcrit < 1.0
    td = hE * np.sqrt(2 * (1 + m["nu0"]) * p["rho_ice"]) / np.sqrt(E)
    f = 1 - (1 - dcrit) * 10.0 / td
    np.testing.assert_allclose(o.get_field("damage", 6)[:, 0], f, rtol=1e-12)
    np.testing.assert_allclose(o.get_field("s11", 8)[:, 0], s11 * f, rtol=1e-12)
    np.testing.assert_allclose(o.get_field("s22", 8)[:, 0], s22 * f, rtol=1e-12)


def test_brittle_momentum_against_a_2x2_solve():
    o, p, _ = _bbm()
    rng = np.random.default_rng(0)
    n = o.ncg
    for name, sc in (("uwind", 10.0), ("vwind", 10.0), ("uocean", 0.05), ("vocean", 0.05)):
        o.set_field(name, rng.normal(0, sc, NX * NY), 1)
    o.set_field("ssh", rng.normal(0, 0.1, NX * NY), 1)
    o.set_field("hice", rng.uniform(0.5, 2.0, NX * NY), 1)
    o.set_field("cice", rng.uniform(0.6, 1.0, NX * NY), 1)
    o.prepare_iteration()
    dt = 12.0
    o.begin_iteration(dt)
    u0, v0 = rng.normal(0, 0.1, n), rng.normal(0, 0.1, n)
    o.set_cg("u", u0)
    o.set_cg("v", v0)
    for name in ("s11", "s12", "s22"):
        o.set_field(name, rng.normal(0, 1e3, (NX * NY, 8)), 8)
    o.run_pass("divergence")
    o.run_pass("momentum")
    g = {k: o.get_cg(k) for k in ("uocean", "vocean", "uwind", "vwind", "cgH", "cgA", "gradX", "gradY",
                                  "dStressX", "dStressY", "lumpedcgmass", "avgU", "u", "v")}
    m = p["rho_ice"] * g["cgH"]
    cP = g["cgA"] * p["F_ocean"] * np.hypot(g["uocean"] - u0, g["vocean"] - v0)
    Fx = g["cgA"] * p["F_atm"] * np.hypot(g["uwind"], g["vwind"]) * g["uwind"] + cP * g["uocean"] \
        + g["dStressX"] / g["lumpedcgmass"] - m * p["gravity"] * g["gradX"]
    Fy = g["cgA"] * p["F_atm"] * np.hypot(g["uwind"], g["vwind"]) * g["vwind"] + cP * g["vocean"] \
        + g["dStressY"] / g["lumpedcgmass"] - m * p["gravity"] * g["gradY"]
    # semi-implicit step: (1 + dt cP/m) u - dt fc v ... written as M [u v] = [u0 v0] + dt/m [Fx Fy]
    al = 1 + dt / m * cP
    be = dt * p["fc"] * np.ones_like(al)
    for i in range(0, n, 7):
        M = np.array([[al[i], -be[i]], [be[i], al[i]]])
        uu, vv = np.linalg.solve(M, [u0[i] + dt / m[i] * Fx[i], v0[i] + dt / m[i] * Fy[i]])
        assert g["u"][i] == pytest.approx(uu, rel=1e-12, abs=1e-16)
        # the reference's v carries +beta dt/m Fx where the solve has -beta dt/m Fx
        quirk = 2 * be[i] * dt / m[i] * Fx[i] / (al[i] ** 2 + be[i] ** 2)
        assert g["v"][i] == pytest.approx(vv + quirk, rel=1e-12, abs=1e-16)
    np.testing.assert_allclose(g["avgU"], g["u"] / p["nSteps"], rtol=1e-14)


def test_dg8_transport_keeps_a_constant_field_in_the_interior():
    nx = ny = 8
    m = OracleMesh.from_coords(nx, ny, box_coords(nx, ny, 1000.0), np.ones(nx * ny))
    t = OracleTransport(m, 8)
    t.set_scheme(2)
    vx = np.zeros((nx * ny, 8))
    vy = np.zeros((nx * ny, 8))
    vx[:, 0], vy[:, 0] = 0.3, -0.2
    t.set_velocity(vx, vy)
    t.reinit()
    phi = np.zeros((nx * ny, 8))
    phi[:, 0] = 2.5
    out = t.step(100.0, phi)
    ix, iy = np.arange(nx * ny) % nx, np.arange(nx * ny) // nx
    inner = (ix > 1) & (ix < nx - 2) & (iy > 1) & (iy < ny - 2)
    np.testing.assert_allclose(out[inner, 0], 2.5, rtol=1e-14)
    assert np.abs(out[inner, 1:]).max() < 1e-13
