"""GPU parity on spherical meshes (longitude, latitude in radians, rotated by
RotatePoleToGreenland; metric terms in every operator) against the oracle, whose spherical
parts are pinned in test_oracle_spherical.py.

Two meshes: a longitude-latitude patch at mid latitudes, and one whose rotated longitudes
cross the +-pi jump inside elements (correctlongitude, edgevector). Tolerances: single
passes within 1e-12 relative L2 per field; 10 model steps within 1e-10.
"""
import numpy as np
import pytest

from helpers import DEFAULT_PARAMS, rel_l2
from oracle import OracleDyn

pytestmark = pytest.mark.gpu

PATCHES = {"midlat": (-0.30, 0.30, 0.55, 0.95), "dateline": (np.pi - 0.25, np.pi + 0.25, -0.35, 0.15)}


def _say(msg):
    print(msg, flush=True)


def _from_greenland(lon, lat):
    x, y, z = np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)
    aw, bw = np.radians(-40.0), np.radians(-15.0)
    x1, y1, z1 = np.cos(bw) * x - np.sin(bw) * z, y, np.sin(bw) * x + np.cos(bw) * z
    x2, y2, z2 = np.cos(aw) * x1 - np.sin(aw) * y1, np.sin(aw) * x1 + np.cos(aw) * y1, z1
    return np.arctan2(y2, x2), np.arcsin(z2)


def _coords(nx, ny, patch):
    l0, l1, p0, p1 = PATCHES[patch]
    j, i = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    lon = l0 + (l1 - l0) * i.ravel() / nx
    lat = p0 + (p1 - p0) * j.ravel() / ny
    return np.stack(_from_greenland(lon, lat), axis=1)


def _mask(nx, ny):
    m = np.ones((ny, nx))
    m[0, :] = m[-1, :] = m[:, 0] = m[:, -1] = 0
    m[ny // 2, nx // 3] = 0
    return m.ravel()


def _pair(cg, dg, nx, ny, patch):
    import nextsimdg_amd as nd

    coords, mask = _coords(nx, ny, patch), _mask(nx, ny)
    o = OracleDyn(cg, dg, dict(DEFAULT_PARAMS))
    o.initialise(nx, ny, coords, mask, spherical=True)
    g = nd.MEVPDynamicsKernel(cg, dg, params=nd.default_params(**DEFAULT_PARAMS))
    g.initialise(coords, True, mask, nx, ny)
    return o, g


def _state(o, g, rng, S, dg):
    ne = o.ne
    h = np.abs(rng.normal(1.0, 0.3, (ne, dg))) * np.r_[1.0, [0.05] * (dg - 1)]
    A = np.clip(rng.normal(0.9, 0.05, (ne, dg)), 0, 1) * np.r_[1.0, [0.02] * (dg - 1)]
    for name, v in (("hice", h), ("cice", A)):
        o.set_field(name, v, dg)
        g.setData(name, v, dg)
    for name in ("s11", "s12", "s22", "e11", "e12", "e22"):
        v = rng.normal(0, 1e3 if name[0] == "s" else 1e-7, (ne, S))
        o.set_field(name, v, S)
        g.setData(name, v, S)
    for name, scale in (("uwind", 10), ("vwind", 10), ("uocean", 0.05), ("vocean", 0.05)):
        v = rng.normal(0, scale, ne)
        o.set_field(name, v, 1)
        g.setData(name, v, 1)
    ssh = rng.normal(0, 0.1, ne)
    o.set_field("ssh", ssh, 1)
    g.setData("ssh", ssh, 1)
    for name in ("u", "v"):
        v = rng.normal(0, 0.1, o.ncg)
        o.set_cg(name, v)
        g.set_cg(name, v)


@pytest.mark.parametrize("patch", list(PATCHES))
@pytest.mark.parametrize("cg,dg", [(2, 6), (1, 3)])
def test_spherical_single_passes_match_oracle(cg, dg, patch):
    _say(f"spherical single passes cg={cg} dg={dg} {patch}")
    nx, ny = 18, 14
    o, g = _pair(cg, dg, nx, ny, patch)
    S = o.S
    _state(o, g, np.random.default_rng(21 + cg), S, dg)
    for n in ("lumpedcgmass", "uwind", "vwind"):
        assert rel_l2(g.get_cg(n), o.get_cg(n)) < 1e-13, n
    o.prepare_iteration()
    g.prepareIteration()
    for n in ("cgH", "cgA", "gradX", "gradY"):
        assert rel_l2(g.get_cg(n), o.get_cg(n)) < 1e-12, n
    o.begin_iteration(120.0)
    g.begin_iteration(120.0)
    o.run_pass("strain")
    g.projectVelocityToStrain()
    for n in ("e11", "e12", "e22"):
        assert rel_l2(g.getDGData(n, S), o.get_field(n, S)) < 1e-12, n
    o.run_pass("stress", 2)
    g.stressUpdateHighOrder(2)
    for n in ("s11", "s12", "s22"):
        assert rel_l2(g.getDGData(n, S), o.get_field(n, S)) < 1e-12, n
    o.run_pass("divergence")
    g.stressDivergence()
    for n in ("dStressX", "dStressY"):
        assert rel_l2(g.get_cg(n), o.get_cg(n)) < 1e-12, n
    o.run_pass("momentum")
    g.updateMomentum()
    o.run_pass("boundaries")
    g.applyBoundaries()
    for n in ("u", "v"):
        assert rel_l2(g.get_cg(n), o.get_cg(n)) < 1e-12, n
    for n in ("u", "v", "uiostress", "viostress"):
        assert rel_l2(g.getDGData(n, dg), o.get_field(n, dg)) < 1e-12, n


@pytest.mark.parametrize("patch", list(PATCHES))
def test_spherical_ten_steps_match_oracle(patch):
    """10 steps (advection with the metric, 100 subcycles each) under a steady wind."""
    _say(f"spherical 10 steps {patch}")
    nx, ny = 24, 20
    o, g = _pair(2, 6, nx, ny, patch)
    rng = np.random.default_rng(3)
    h = rng.uniform(0.5, 1.5, nx * ny)
    A = rng.uniform(0.85, 1.0, nx * ny)
    uw, vw = rng.normal(8, 2, nx * ny), rng.normal(-4, 2, nx * ny)
    for name, v in (("uwind", uw), ("vwind", vw), ("uocean", 0.02 * np.ones(nx * ny)),
                    ("vocean", np.zeros(nx * ny)), ("ssh", np.zeros(nx * ny)), ("hice", h), ("cice", A)):
        o.set_field(name, v, 1)
        g.setData(name, v, 1)
    for k in range(10):
        o.update(120.0)
        g.update(120.0)
        _say(f"  step {k}: du={rel_l2(g.get_cg('u'), o.get_cg('u')):.2e}")
    errs = {n: rel_l2(g.get_cg(n), o.get_cg(n)) for n in ("u", "v")}
    errs.update({n: rel_l2(g.getDGData(n, 6), o.get_field(n, 6)) for n in ("hice", "cice")})
    errs.update({n: rel_l2(g.getDGData(n, 8), o.get_field(n, 8)) for n in ("s11", "s12", "s22", "e11", "e12", "e22")})
    _say("  " + " ".join(f"{n}={e:.1e}" for n, e in errs.items()))
    assert np.abs(g.get_cg("u")).max() > 1e-3
    for n, e in errs.items():
        assert e < 1e-10, (n, e)


def test_spherical_meshes_run_fp64_only():
    """The FP32 variant is built for uniform meshes; both rheologies run on the sphere
    (BBM parity on spherical patches is in test_gpu_bbm.py)."""
    import nextsimdg_amd as nd

    nx, ny = 12, 10
    coords, mask = _coords(nx, ny, "midlat"), _mask(nx, ny)
    k = nd.MEVPDynamicsKernel(2, 6)
    k.initialise(coords, True, mask, nx, ny)
    with pytest.raises(nd.NsdgError):
        k.set_precision(nd.PREC_FP32)
    b = nd.BBMDynamicsKernel(2, 6)
    b.initialise(coords, True, mask, nx, ny)
    with pytest.raises(nd.NsdgError):
        b.set_precision(nd.PREC_FP32)
