"""Pins of the oracle's spherical meshes (ParametricMesh SPHERICAL: longitude, latitude in
radians, rotated by RotatePoleToGreenland; SphericalTools::massMatrix; the spherical
branches of ParametricMap.cpp and CGDynamicsKernel.cpp). The reference has no spherical
test, so each part is checked against the mathematics it discretises:

  * lumped mass: the metric cos(latitude) integrated over a longitude-latitude rectangle,
    dlon (sin lat1 - sin lat0);
  * strain: zero for rigid rotations of the sphere (about the polar and an equatorial
    axis, converging as h^2); the metric terms alone for a constant zonal or meridional
    velocity (e12 = U tan(lat) / 2R, e11 = -V tan(lat) / R);
  * divergence: an isotropic constant stress exerts no force on interior nodes;
  * CG2DG: constants exactly; the cos-weighted cell mean of sin(latitude);
  * SSH gradient of a constant height: -c tan(lat) / R at interior nodes. The reference's
    weak form, -(d_lat phi_j cos(lat), eta) / (phi_j, cos(lat)), is that of
    d_lat(eta cos(lat)) / cos(lat): the gradient plus -tan(lat) eta / R, which this pins;
  * transport: the cos-weighted DG0 mass, computed here by an independent quadrature, is
    conserved in a general flow, and a constant stays constant under a rigid zonal flow;
  * BBM projection: the reference projects Gauss-point values with the cos-latitude mass
    matrix against weights without cos(latitude) (ParametricMap.cpp:333-351), so a
    constant damage comes back scaled by about 1/cos(latitude); pinned against that
    product computed here.

Meshes are built in the rotated (internal) frame: the vertices handed to the kernel are
those of a longitude-latitude rectangle taken back through the inverse rotation
(RotatePoleFromGreenland, ParametricMesh.hpp:159-180).
"""
import numpy as np
import pytest

from helpers import DEFAULT_PARAMS
from oracle import OracleDyn

R = 6371000.0
LON0, LON1, LAT0, LAT1 = -0.30, 0.30, 0.55, 0.95


def _from_greenland(lon, lat):
    x, y, z = np.cos(lat) * np.cos(lon), np.cos(lat) * np.sin(lon), np.sin(lat)
    aw, bw = np.radians(-40.0), np.radians(-15.0)
    x1, y1, z1 = np.cos(bw) * x - np.sin(bw) * z, y, np.sin(bw) * x + np.cos(bw) * z
    x2, y2, z2 = np.cos(aw) * x1 - np.sin(aw) * y1, np.sin(aw) * x1 + np.cos(aw) * y1, z1
    return np.arctan2(y2, x2), np.arcsin(z2)


def _mesh(nx, ny):
    j, i = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    lon = LON0 + (LON1 - LON0) * i.ravel() / nx
    lat = LAT0 + (LAT1 - LAT0) * j.ravel() / ny
    return np.stack(_from_greenland(lon, lat), axis=1)


def _kernel(nx, ny, cg=2, dg=6, mask=None):
    o = OracleDyn(cg, dg, dict(DEFAULT_PARAMS))
    o.initialise(nx, ny, _mesh(nx, ny), np.ones(nx * ny) if mask is None else mask, spherical=True)
    return o


def _nodes(nx, ny, cg=2):
    j, i = np.meshgrid(np.arange(cg * ny + 1), np.arange(cg * nx + 1), indexing="ij")
    return (LON0 + (LON1 - LON0) * i.ravel() / (cg * nx), LAT0 + (LAT1 - LAT0) * j.ravel() / (cg * ny))


def _interior(nx, ny, cg=2, margin=2):
    j, i = np.meshgrid(np.arange(cg * ny + 1), np.arange(cg * nx + 1), indexing="ij")
    return ((i >= margin) & (i <= cg * nx - margin) & (j >= margin) & (j <= cg * ny - margin)).ravel()


def test_inverse_rotation_round_trip():
    lon, lat = np.meshgrid(np.linspace(-3, 3, 7), np.linspace(-1.4, 1.4, 5))
    blon, blat = _from_greenland(lon, lat)
    o = OracleDyn(2, 6, dict(DEFAULT_PARAMS))
    # the kernel's rotation of these vertices gives back the rectangle (checked through the
    # lumped mass below); here only the inverse is checked to be a rotation
    v = np.stack([np.cos(blat) * np.cos(blon), np.cos(blat) * np.sin(blon), np.sin(blat)])
    np.testing.assert_allclose((v ** 2).sum(0), 1.0, rtol=1e-15)
    del o


def test_spherical_lumped_mass_integrates_the_metric():
    nx, ny = 24, 16
    o = _kernel(nx, ny)
    area = (LON1 - LON0) * (np.sin(LAT1) - np.sin(LAT0))
    assert o.get_cg("lumpedcgmass").sum() == pytest.approx(area, rel=1e-11)


@pytest.mark.parametrize("axis", ["polar", "equatorial"])
def test_spherical_strain_of_rigid_rotation_vanishes(axis):
    om = 1e-6  # 1/s
    errs = []
    for nx, ny in ((12, 8), (24, 16)):
        o = _kernel(nx, ny)
        lon, lat = _nodes(nx, ny)
        if axis == "polar":
            u, v = om * R * np.cos(lat), 0 * lat
        else:
            u, v = -om * R * np.sin(lat) * np.cos(lon), om * R * np.sin(lon)
        o.set_cg("u", u)
        o.set_cg("v", v)
        o.run_pass("strain")
        e = np.concatenate([o.get_field(n, 8)[:, 0] for n in ("e11", "e12", "e22")])
        errs.append(np.abs(e).max() / om)
    assert errs[1] < 1e-3 and errs[0] / errs[1] > 3.0, errs  # O(h^2) to zero


def test_spherical_strain_metric_terms():
    nx, ny = 24, 16
    o = _kernel(nx, ny)
    lon, lat = _nodes(nx, ny)
    U, V = 0.1, 0.07
    o.set_cg("u", np.full(lon.size, U))
    o.set_cg("v", np.zeros(lon.size))
    o.run_pass("strain")
    # cos-weighted cell means of the DG strain against those of the exact tensor
    jc, ic = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    la0 = LAT0 + (LAT1 - LAT0) * jc.ravel() / ny
    la1 = la0 + (LAT1 - LAT0) / ny
    mean_tan = (np.cos(la0) - np.cos(la1)) / (np.sin(la1) - np.sin(la0))  # <tan lat>_cos
    m0 = _cos_mass0(nx, ny, 8)

    def mean(f):
        return (o.get_field(f, 8) * m0).sum(1) / m0[:, 0]

    np.testing.assert_allclose(mean("e12"), 0.5 * U * mean_tan / R, rtol=1e-8)
    assert np.abs(mean("e11")).max() < 1e-12 * U / R
    o.set_cg("u", np.zeros(lon.size))
    o.set_cg("v", np.full(lon.size, V))
    o.run_pass("strain")
    np.testing.assert_allclose(mean("e11"), -V * mean_tan / R, rtol=1e-8)
    assert np.abs(o.get_field("e22", 8)).max() < 1e-12 * V / R
    assert np.abs(o.get_field("e12", 8)).max() < 1e-12 * V / R


def test_spherical_isotropic_stress_exerts_no_interior_force():
    nx, ny = 16, 12
    o = _kernel(nx, ny)
    ne = nx * ny
    p = np.zeros((ne, 8))
    p[:, 0] = -2.0e4
    o.set_field("s11", p, 8)
    o.set_field("s22", p, 8)
    o.set_field("s12", np.zeros((ne, 8)), 8)
    o.run_pass("divergence")
    fx, fy = o.get_cg("dStressX"), o.get_cg("dStressY")
    inner = _interior(nx, ny)
    # the scale: the force of sigma22 = p alone, which the metric term of sigma11 cancels
    o.set_field("s11", np.zeros((ne, 8)), 8)
    o.run_pass("divergence")
    scale = np.abs(o.get_cg("dStressY")[inner]).max()
    assert scale > 0
    assert np.abs(fx[inner]).max() < 1e-12 * scale
    assert np.abs(fy[inner]).max() < 1e-6 * scale


def test_spherical_cg2dg_cell_means():
    nx, ny = 24, 16
    o = _kernel(nx, ny)
    lon, lat = _nodes(nx, ny)
    o.set_cg("u", np.full(lon.size, 3.5))
    got = o.get_field("u", 6)
    np.testing.assert_allclose(got[:, 0], 3.5, rtol=1e-13)
    assert np.abs(got[:, 1:]).max() < 1e-12
    o.set_cg("u", np.sin(lat))
    jc, ic = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    la0 = LAT0 + (LAT1 - LAT0) * jc.ravel() / ny
    la1 = la0 + (LAT1 - LAT0) / ny
    m0 = _cos_mass0(nx, ny, 6)
    mean = (o.get_field("u", 6) * m0).sum(1) / m0[:, 0]  # the projection keeps cos-weighted means
    np.testing.assert_allclose(mean, 0.5 * (np.sin(la0) + np.sin(la1)), rtol=1e-7)


def test_spherical_ssh_gradient_of_a_constant_height():
    nx, ny = 24, 16
    o = _kernel(nx, ny)
    c = 0.8
    o.set_field("ssh", np.full(nx * ny, c), 1)
    o.set_field("hice", np.ones(nx * ny), 1)
    o.set_field("cice", np.ones(nx * ny), 1)
    o.prepare_iteration()
    lon, lat = _nodes(nx, ny)
    inner = _interior(nx, ny, margin=3)
    gx, gy = o.get_cg("gradX"), o.get_cg("gradY")
    assert np.abs(gx[inner]).max() < 1e-12 * c / R
    np.testing.assert_allclose(gy[inner], -c * np.tan(lat[inner]) / R, rtol=2e-3)


def _cos_mass0(nx, ny, dg):
    """Row 0 of the spherical DG mass matrix of every cell, by an independent 5x5 Gauss
    rule on the longitude-latitude rectangle (the elements are affine in (lon, lat))."""
    g, w = np.polynomial.legendre.leggauss(5)
    t, w = 0.5 * (g + 1), 0.5 * w
    X, Y = np.meshgrid(t - 0.5, t - 0.5, indexing="ij")
    W = np.outer(w, w)
    basis = [np.ones_like(X), X, Y, X * X - 1 / 12, Y * Y - 1 / 12, X * Y, Y * (X * X - 1 / 12),
             X * (Y * Y - 1 / 12)][:dg]
    dlon, dlat = (LON1 - LON0) / nx, (LAT1 - LAT0) / ny
    out = np.zeros((nx * ny, dg))
    for e in range(nx * ny):
        la = LAT0 + (e // nx + t) * dlat
        cw = np.cos(la)[None, :] * W * dlon * dlat
        out[e] = [(b * cw).sum() for b in basis]
    return out


def test_spherical_transport_conserves_the_metric_mass():
    nx, ny, dg = 16, 12, 6
    o = OracleDyn(2, dg, dict(DEFAULT_PARAMS))
    mask = np.ones((ny, nx))
    mask[0, :] = mask[-1, :] = mask[:, 0] = mask[:, -1] = 0
    o.initialise(nx, ny, _mesh(nx, ny), mask.ravel(), spherical=True)
    lon, lat = _nodes(nx, ny)
    inner = _interior(nx, ny, margin=5)
    # a general flow, zero on the cells next to the land ring (no flux through the
    # outflow-only Dirichlet edges)
    o.set_cg("u", np.where(inner, 0.3 * np.sin(3 * lat) * np.cos(2 * lon), 0.0))
    o.set_cg("v", np.where(inner, 0.2 * np.cos(4 * lon), 0.0))
    rng = np.random.default_rng(4)
    h = np.zeros((nx * ny, dg))
    h[:, 0] = rng.uniform(0.5, 1.5, nx * ny)
    h[:, 1:3] = rng.normal(0, 0.05, (nx * ny, 2))
    o.set_field("hice", h, dg)
    m0 = _cos_mass0(nx, ny, dg)
    before = (m0 * o.get_field("hice", dg)).sum()
    for _ in range(5):
        o.advection(3000.0)
    after = (m0 * o.get_field("hice", dg)).sum()
    assert after == pytest.approx(before, rel=1e-13)
    # a constant under a rigid zonal flow stays constant away from the outflow-only
    # Dirichlet edges of the first and last ocean columns
    o.set_cg("u", 0.3 * np.cos(lat))
    o.set_cg("v", np.zeros(lon.size))
    o.set_field("cice", np.full(nx * ny, 0.9), 1)
    o.advection(3000.0)
    A = o.get_field("cice", dg)
    col, row = np.arange(nx * ny) % nx, np.arange(nx * ny) // nx
    far = (col >= 3) & (col <= nx - 4) & (row >= 1) & (row <= ny - 2)
    np.testing.assert_allclose(A[far, 0], 0.9, rtol=1e-12)
    assert np.abs(A[far, 1:]).max() < 1e-12


def test_spherical_transport_moves_with_the_rotation_rate():
    """Solid-body rotation about the polar axis at rate om carries a bump om t in longitude
    (this fixes the 1/R of the spherical inverse mass)."""
    nx, ny, dg = 32, 8, 6
    o = OracleDyn(2, dg, dict(DEFAULT_PARAMS))
    o.initialise(nx, ny, _mesh(nx, ny), np.ones(nx * ny), spherical=True)
    lon, lat = _nodes(nx, ny)
    om, dt, steps = 2.5e-6, 1000.0, 30  # Courant number 0.13 (DG2 with Heun needs well below 0.2)
    o.set_cg("u", om * R * np.cos(lat))
    o.set_cg("v", np.zeros(lon.size))
    col = np.arange(nx * ny) % nx
    lon_c = LON0 + (LON1 - LON0) * (col + 0.5) / nx
    h = np.exp(-((lon_c + 0.12) / 0.05) ** 2)
    o.set_field("hice", h, 1)
    m0 = _cos_mass0(nx, ny, dg)[:, 0]

    def centroid():  # first moment of the cell means and slopes in longitude
        hh = o.get_field("hice", dg)
        return ((lon_c * hh[:, 0] + hh[:, 1] * (LON1 - LON0) / nx / 12) * m0).sum() / (hh[:, 0] * m0).sum()

    c0 = centroid()
    for _ in range(steps):
        o.advection(dt)
    assert centroid() - c0 == pytest.approx(om * dt * steps, rel=0.02)


def test_spherical_bbm_projection_has_no_metric_weight():
    """One BBM stress pass with no strain and no stress leaves a constant damage d at every
    Gauss point (no failure: tau + tan(phi) sn = 0), and projects it with iMJwPSI_dam =
    M_cos^-1 PSI diag(w J): here J = dlon dlat on the longitude-latitude rectangle and
    M_cos the 3 x 3 Gauss mass matrix with cos(latitude), both computed independently.
    Cell means come out near d / cos(latitude), not d."""
    from oracle import MEB_DEFAULTS

    nx, ny, dg, d = 6, 4, 6, 0.5
    o = OracleDyn(2, dg, dict(DEFAULT_PARAMS), bbm=True, meb=dict(MEB_DEFAULTS))
    o.initialise(nx, ny, _mesh(nx, ny), np.ones(nx * ny), spherical=True)
    ne = nx * ny
    dmg = np.zeros((ne, dg))
    dmg[:, 0] = d
    o.set_field("damage", dmg, dg)
    o.set_field("hice", np.ones(ne), 1)
    o.set_field("cice", np.ones(ne), 1)
    o.prepare_iteration()
    o.begin_iteration(1.2)
    o.run_pass("stress", 1)
    got = o.get_field("damage", dg)
    g, w = np.polynomial.legendre.leggauss(3)
    t, w = 0.5 * (g + 1), 0.5 * w
    X, Y = np.meshgrid(t - 0.5, t - 0.5, indexing="ij")  # [x][y]
    W = np.outer(w, w)
    basis = np.array([np.ones_like(X), X, Y, X * X - 1 / 12, Y * Y - 1 / 12, X * Y])
    dlon, dlat = (LON1 - LON0) / nx, (LAT1 - LAT0) / ny
    J = dlon * dlat
    for e in range(ne):
        lat = LAT0 + (e // nx + t[None, :]) * dlat + 0 * X
        M = np.einsum("iab,jab,ab->ij", basis, basis, W * J * np.cos(lat))
        rhs = np.einsum("iab,ab->i", basis, W * J * d)
        want = np.linalg.solve(M, rhs)
        np.testing.assert_allclose(got[e], want, rtol=1e-12, atol=1e-14)
        latc = LAT0 + (e // nx + 0.5) * dlat
        assert abs(got[e, 0] * np.cos(latc) / d - 1.0) < 5e-3

