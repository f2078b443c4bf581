"""The momentum residual on the GPU against the oracle's pinned passes.

The residual R = M_L [c1 (u - u0) - A (tau_a + F_oc |du| (u_o - u)) + rho H f_c (u | -v)
+ rho H g grad(eta)] - dS(Pi sigma_VP(eps(u))) is composed here from the oracle's own
strain, stress (alpha = 1, i.e. the VP stress outright) and divergence passes and the
node fields, and checked two ways:
  * against the fixed-point identity of the oracle's momentum update,
    R = M_L (c1 (1 + beta) + A F_oc |du|) (u - update(u)), which pins the formula;
  * the GPU norms |R| and |M_L A tau_a| against the composed ones at 1e-10.
"""
import numpy as np
import pytest

from helpers import DEFAULT_PARAMS, box_coords
from oracle import OracleDyn
from test_gpu_parity import _box_mask, _set_random_state

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _oracle_residual(o, dt, p, dirichlet):
    """R at every node from the oracle's passes (o must run with alpha = 1)."""
    g = {n: o.get_cg(n) for n in ("u", "v", "u0", "v0", "uocean", "vocean", "uwind", "vwind", "cgH", "cgA",
                                   "gradX", "gradY", "lumpedcgmass")}
    o.run_pass("strain")
    o.run_pass("stress")
    o.run_pass("divergence")
    dSx, dSy = o.get_cg("dStressX"), o.get_cg("dStressY")
    u, v, uo, vo = g["u"], g["v"], g["uocean"], g["vocean"]
    absatm = np.hypot(g["uwind"], g["vwind"])
    absocn = np.hypot(uo - u, v - vo)
    c1 = p["rho_ice"] * g["cgH"] / dt
    rhoH = p["rho_ice"] * g["cgH"]
    A, ML = g["cgA"], g["lumpedcgmass"]
    tax, tay = p["F_atm"] * absatm * g["uwind"], p["F_atm"] * absatm * g["vwind"]
    rx = ML * (c1 * (u - g["u0"]) - A * (tax + p["F_ocean"] * absocn * (uo - u)) + rhoH * p["fc"] * u
               + rhoH * p["gravity"] * g["gradX"]) - dSx
    ry = ML * (c1 * (v - g["v0"]) - A * (tay + p["F_ocean"] * absocn * (vo - v)) - rhoH * p["fc"] * v
               + rhoH * p["gravity"] * g["gradY"]) - dSy
    keep = ~dirichlet
    # fixed-point identity with the oracle's momentum update (same VP divergence)
    o.run_pass("momentum")
    un, vn = o.get_cg("u"), o.get_cg("v")
    den = ML * (c1 * (1 + p["beta"]) + A * p["F_ocean"] * absocn)
    np.testing.assert_allclose(rx[keep], (den * (u - un))[keep], rtol=1e-9, atol=1e-12 * np.abs(rx).max())
    np.testing.assert_allclose(ry[keep], (den * (v - vn))[keep], rtol=1e-9, atol=1e-12 * np.abs(ry).max())
    nR = np.sqrt(np.sum(rx[keep] ** 2 + ry[keep] ** 2))
    nD = np.sqrt(np.sum((ML * A * tax)[keep] ** 2 + (ML * A * tay)[keep] ** 2))
    return nR, nD


def _dirichlet_nodes(cg, dg, nx, ny, coords, mask):
    o = OracleDyn(cg, dg, DEFAULT_PARAMS)
    o.initialise(nx, ny, coords, mask)
    o.set_cg("u", np.ones(o.ncg))
    o.run_pass("boundaries")
    return o.get_cg("u") == 0.0


@pytest.mark.parametrize("cg,dg", [(2, 6), (1, 3)])
@pytest.mark.parametrize("uniform,nx,ny", [(True, 18, 14), (True, 63, 15), (False, 18, 14)])
def test_residual_matches_oracle(cg, dg, uniform, nx, ny):
    """Uniform meshes take the one-pass residual kernel (tiles of 31 x 7 elements: one
    tile, and one past two tile multiples), general meshes the split passes."""
    import nextsimdg_amd as nd

    _say(f"residual cg={cg} dg={dg} uniform={uniform} {nx}x{ny}")
    dt = 120.0
    coords = box_coords(nx, ny, 4000.0)
    if not uniform:
        coords[:, 0] += 150.0 * np.sin(coords[:, 1] / 9000.0)
    mask = _box_mask(nx, ny)
    mask[4 * nx + 6] = 0.0
    p1 = dict(DEFAULT_PARAMS, alpha=1.0)
    o = OracleDyn(cg, dg, p1)
    o.initialise(nx, ny, coords, mask)
    g = nd.MEVPDynamicsKernel(cg, dg, params=nd.default_params(**DEFAULT_PARAMS))
    g.initialise(coords, False, mask, nx, ny)
    _set_random_state(o, g, np.random.default_rng(7 + cg), o.S, dg)
    for k in (o,):
        k.prepare_iteration()
        k.begin_iteration(dt)
    g.prepareIteration()
    g.begin_iteration(dt)
    # a velocity away from u0
    rng = np.random.default_rng(1)
    for name in ("u", "v"):
        w = g.get_cg(name) + rng.normal(0, 0.02, o.ncg)
        o.set_cg(name, w)
        g.set_cg(name, w)
    r, nR, nD = g.momentum_residual()
    oR, oD = _oracle_residual(o, dt, p1, _dirichlet_nodes(cg, dg, nx, ny, coords, mask))
    _say(f"  gpu |R|={nR:.6e} |D|={nD:.6e}  oracle |R|={oR:.6e} |D|={oD:.6e}")
    assert nR == pytest.approx(oR, rel=1e-10)
    assert nD == pytest.approx(oD, rel=1e-10)
    assert r == pytest.approx(oR / oD, rel=1e-10)
    # the residual leaves the state alone
    r2, _, _ = g.momentum_residual()
    assert r2 == r


def test_stress_from_velocity_is_the_alpha1_stress_update():
    import nextsimdg_amd as nd

    nx, ny, dt = 16, 12, 120.0
    coords, mask = box_coords(nx, ny, 3000.0), _box_mask(nx, ny)
    p1 = dict(DEFAULT_PARAMS, alpha=1.0)
    o = OracleDyn(2, 6, p1)
    o.initialise(nx, ny, coords, mask)
    g = nd.MEVPDynamicsKernel(2, 6, params=nd.default_params(**DEFAULT_PARAMS))
    g.initialise(coords, False, mask, nx, ny)
    _set_random_state(o, g, np.random.default_rng(3), 8, 6)
    o.run_pass("strain")
    o.run_pass("stress")
    g.stress_from_velocity()
    for n in ("s11", "s12", "s22"):
        ref = o.get_field(n, 8).reshape(nx * ny, 8)[mask.astype(bool)]
        got = g.getDGData(n, 8).reshape(nx * ny, 8)[mask.astype(bool)]
        np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-9 * np.abs(ref).max())


def test_mevp_converges_slowly_in_the_residual():
    """mEVP with alpha = beta = 1500 lowers the residual over its 100 subcycles and keeps
    converging over ~10^4 (64 x 64 heterogeneous box, first step)."""
    import nextsimdg_amd as nd
    from nextsimdg_amd import inputs

    nx = 64
    s = inputs.benchmark_setup(nx, nx, 3, kind="heterogeneous")
    k = nd.MEVPDynamicsKernel(2, 6)
    k.initialise(s["coords"], False, s["mask"], nx, nx)
    for name, key in (("hice", "h"), ("cice", "A"), ("uocean", "uocean"), ("vocean", "vocean")):
        k.setData(name, s[key], 1)
    k.setData("ssh", 0 * s["h"], 1)
    k.benchmark_wind(inputs.L_BOX, 0.0, *inputs.cyclone_offset(3), 30.0)
    k.advectionAndLimits(120.0)
    k.prepareIteration()
    k.begin_iteration(120.0)
    r0 = k.momentum_residual()[0]
    k.subcycles(100)
    r100 = k.momentum_residual()[0]
    k.subcycles(9900)
    r10k = k.momentum_residual()[0]
    _say(f"  r0={r0:.4e} r100={r100:.4e} r10000={r10k:.4e}")
    assert r100 < r0
    assert r10k < 0.2 * r100
