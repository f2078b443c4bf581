"""Pins of the FP64 CPU oracle against the reference's own golden values and fixtures,
and against properties the mathematics fixes (closed forms, invariants, brute force).

Every check here is chosen so that a plausible restatement error (dropped term, wrong
sign or index, transposed operand) fails it. No GPU is involved.
"""
import json
import os

import numpy as np
import pytest

from helpers import DEFAULT_PARAMS, GOLDEN, box_coords, cg_node_coords, read_nh_smesh, smesh_text
from oracle import Oracle, OracleDyn, OracleMesh

ORA = Oracle()

# --------------------------------------------------------------------------- golden values
# dynamics/test/Advection_test.cpp:274-283 (uniform) and :292-301 (distorted), used it<2
ADV_UNIFORM = {1: [5.1256149074257538e-02, 4.8288256703303903e-02],
               3: [2.9450967798560313e-02, 1.3427281824939470e-02],
               6: [9.9340386651904228e-03, 4.0274889287136816e-03]}
ADV_DISTORTED = {1: [5.0958748236875594e-02, 4.8461087243594887e-02],
                 3: [3.2033423984965226e-02, 1.5639052766007147e-02],
                 6: [1.1830071142946669e-02, 4.9207680503709564e-03]}
# dynamics/test/AdvectionPeriodicBC_test.cpp:281-290
ADV_RING = {1: [1.0338503986019776e+00, 1.1451366598186583e+00],
            3: [1.0949680727313791e+00, 8.3851748134278226e-01],
            6: [7.1704413076190610e-01, 4.6135258391583606e-01]}
TOL = 1e-7  # the reference's own tolerance (Advection_test.cpp:26)


@pytest.mark.parametrize("dg", [1, 3, 6])
@pytest.mark.parametrize("it", [0, 1])
def test_advection_golden_uniform(dg, it):
    err, _ = ORA.advection_test(dg, 0.0, it)
    assert err == pytest.approx(ADV_UNIFORM[dg][it], rel=TOL)


@pytest.mark.parametrize("dg", [1, 3, 6])
@pytest.mark.parametrize("it", [0, 1])
def test_advection_golden_distorted(dg, it):
    err, _ = ORA.advection_test(dg, 0.05, it)
    assert err == pytest.approx(ADV_DISTORTED[dg][it], rel=TOL)


@pytest.mark.parametrize("dg", [1, 3, 6])
@pytest.mark.parametrize("it", [0, 1])
def test_advection_golden_periodic_ring(dg, it):
    err, _ = ORA.advection_ring_test(dg, it)
    assert err == pytest.approx(ADV_RING[dg][it], rel=TOL)


# --------------------------------------------------------------------------- basis tables
def _table_args(key):
    name, rest = key.split("<")
    args = [int(a) for a in rest.rstrip(">").split(",")]
    return name, args


def test_basis_tables_match_reference_headers():
    """Every constant of codeGenerationDGinGauss.hpp / codeGenerationCGinGauss.hpp."""
    with open(os.path.join(GOLDEN, "basis_tables.json")) as f:
        tables = json.load(f)
    checked = 0
    for key, ref in tables.items():
        name, args = _table_args(key)
        if name == "PSIe":   # PSIe<E, G>
            mine = ORA.table("PSIe", args[0], args[1])
        elif name == "PSIe_w":  # PSIe_w<D, G, edge>
            mine = ORA.table("PSIe_w", args[0], args[1], args[2])
        elif name == "GAUSSWEIGHTS":
            mine = ORA.table("GAUSSWEIGHTS", args[0])
        else:
            mine = ORA.table(name, args[0], args[1])
        ref = np.array(ref)
        assert mine.shape == ref.shape, key
        np.testing.assert_allclose(mine, ref, rtol=0, atol=2e-15 * max(1.0, np.abs(ref).max()), err_msg=key)
        checked += 1
    assert checked >= 100


def test_dg_mass_is_diagonal_under_3x3_gauss():
    """Orthogonality of the DG basis: M_ref = diag(1, 1/12, 1/12, 1/180, 1/180, 1/144, 1/2160, 1/2160)."""
    psi = ORA.table("PSI", 8, 3).reshape(8, 9)
    w = ORA.table("GAUSSWEIGHTS", 3)
    M = psi @ np.diag(w) @ psi.T
    expect = np.diag([1, 1 / 12, 1 / 12, 1 / 180, 1 / 180, 1 / 144, 1 / 2160, 1 / 2160])
    np.testing.assert_allclose(M, expect, atol=1e-15)


def test_gauss_rules_integrate_monomials_exactly():
    """The tensor weights come from g-point Gauss-Legendre rules exact to degree 2g-1."""
    for g in range(1, 5):
        W = ORA.table("GAUSSWEIGHTS", g).reshape(g, g)
        w1 = W[0] / np.sqrt(W[0, 0])  # W[0, gx] = w[gx] w[0]
        x = np.polynomial.legendre.leggauss(g)[0] * 0.5 + 0.5
        for p in range(2 * g):
            assert np.dot(w1, np.sort(x) ** p) == pytest.approx(1.0 / (p + 1), abs=1e-14)


# --------------------------------------------------------------------------- mesh
def test_smesh_reader_matches_reference_test():
    """dynamics/test/ParametricMesh_test.cpp:39-79 on the 25 km NH mesh."""
    nx, ny, dx = 154, 121, 25000.0
    m = OracleMesh.from_smesh(read_nh_smesh())
    assert m.shape == (nx, ny)
    v = m.vertices()
    assert v[0, 0] == 0.0 and v[0, 1] == 0.0
    assert v[1, 0] == dx and v[1, 1] == 0.0
    assert v[nx + 1, 0] == 0.0 and v[nx + 1, 1] == dx
    assert v[-1, 0] == nx * dx and v[-1, 1] == ny * dx
    lm = m.landmask()
    assert not lm[0] and not lm[4] and lm[5] and lm[nx - 1] and not lm[(ny - 1) * nx]
    d = [m.dirichlet(e) for e in range(4)]
    assert d[0][0] == 5 and d[3][0] == 5 and d[1][0] == 7 and d[2][0] == 48
    assert len(d[0]) == 369 and len(d[1]) == 384
    assert len(d[0]) == len(d[2]) and len(d[1]) == len(d[3])
    assert len(m.periodic()) == 0


def test_mask_derived_dirichlet_equals_smesh_lists():
    """ParametricMesh_test.cpp:84-139: dirichletFromMask + dirichletFromEdge == file lists."""
    f = OracleMesh.from_smesh(read_nh_smesh())
    nx, ny = f.shape
    a = OracleMesh.from_coords(nx, ny, f.vertices(), f.landmask().astype(np.float64))
    np.testing.assert_array_equal(a.landmask(), f.landmask())
    for e in range(4):
        np.testing.assert_array_equal(a.dirichlet(e), f.dirichlet(e))


# --------------------------------------------------------------------------- interpolation
def _poly2(x, y):
    return 0.3 + 2e-6 * x - 3e-6 * y + 4e-12 * x * x - 1e-12 * y * y + 5e-12 * x * y


@pytest.mark.parametrize("cg", [1, 2])
def test_cg2dg_reproduces_p2_and_dg2cg_partition(cg):
    nx, ny, dx, dy = 7, 5, 1000.0, 1500.0
    m = OracleMesh.rectangle(nx * dx, ny * dy, nx, ny)
    X, Y = cg_node_coords(nx, ny, dx, dy, cg=cg)
    # CG1 reproduces bilinear, CG2 biquadratic; restrict to what each space holds exactly
    f = (lambda x, y: 0.3 + 2e-6 * x - 3e-6 * y + 1e-12 * x * y) if cg == 1 else _poly2
    dg = ORA.cg2dg(m, cg, 6, f(X, Y)).reshape(-1, 6)
    # compare cell means to the exact integral mean over each rectangle (numpy quadrature)
    g, w = np.polynomial.legendre.leggauss(4)
    g, w = 0.5 * g + 0.5, 0.5 * w
    for e in [0, 8, nx * ny - 1]:
        ix, iy = e % nx, e // nx
        xs, ys = (ix + g) * dx, (iy + g) * dy
        mean = sum(w[a] * w[b] * f(xs[a], ys[b]) for a in range(4) for b in range(4))
        assert dg[e, 0] == pytest.approx(mean, rel=1e-13)
    # DG2CG of a constant DG field is that constant at every node (boundary doubling)
    ones = np.zeros((nx * ny, 6))
    ones[:, 0] = 1.0
    np.testing.assert_allclose(ORA.dg2cg(m, cg, 6, ones), 1.0, rtol=1e-15)


def test_dg2cg_of_continuous_field_gives_nodal_values():
    nx, ny, dx, dy = 6, 4, 2000.0, 1000.0
    m = OracleMesh.rectangle(nx * dx, ny * dy, nx, ny)
    X, Y = cg_node_coords(nx, ny, dx, dy, cg=2)
    dg = ORA.cg2dg(m, 2, 6, _poly2(X, Y))  # exact P2 projection
    back = ORA.dg2cg(m, 2, 6, dg)
    np.testing.assert_allclose(back, _poly2(X, Y), rtol=1e-13)


# --------------------------------------------------------------------------- limiters
def test_limiters_bound_gauss_values_and_are_idempotent():
    rng = np.random.default_rng(3)
    phi = rng.normal(0.9, 0.2, (500, 6))
    psi = ORA.table("PSI", 6, 3).reshape(6, 9)
    lim = ORA.limit(6, phi, 1.0, True).reshape(-1, 6)
    assert (lim @ psi).max() <= 1.0 + 1e-13
    assert lim[:, 0].max() <= 1.0
    np.testing.assert_allclose(ORA.limit(6, lim, 1.0, True).reshape(-1, 6), lim, rtol=1e-13, atol=1e-15)
    # LimitMin clamps the mean; fields already >= min at every Gauss point are untouched.
    # (Its scale factor (c0 - min) / (minvalue - c0) is negative, so where it acts it flips
    # the higher moments, dgLimit.hpp:79-81; the periodic ring golden values pin that.)
    lo = ORA.limit(6, phi, 0.0, False).reshape(-1, 6)
    assert lo[:, 0].min() >= 0.0
    ok = (phi @ psi).min(axis=1) >= 0.0
    assert ok.sum() > 50
    np.testing.assert_array_equal(lo[ok], phi[ok])
    # DG1 (3 dofs): the corner bound 2 max(|c1+c2|, |c1-c2|) of dgLimit.hpp:22
    p3 = rng.normal(0.9, 0.3, (400, 3))
    l3 = ORA.limit(3, p3, 1.0, True).reshape(-1, 3)
    corners = l3[:, :1] + 0.5 * np.array([[s1 * a + s2 * b for s1 in (-1, 1) for s2 in (-1, 1)]
                                           for a, b in zip(l3[:, 1], l3[:, 2])])
    assert corners.max() <= 1.0 + 1e-13


# --------------------------------------------------------------------------- momentum map
def _dyn(cg, dg, nx, ny, dx, dy=None, skew=0.0, params=None, mask=None):
    d = OracleDyn(cg, dg, params or DEFAULT_PARAMS)
    mask = np.ones(nx * ny) if mask is None else mask
    d.initialise(nx, ny, box_coords(nx, ny, dx, dy, skew), mask)
    return d


@pytest.mark.parametrize("cg", [1, 2])
def test_lumped_mass_closed_form(cg):
    nx, ny, dx, dy = 6, 5, 1000.0, 800.0
    d = _dyn(cg, 6 if cg == 2 else 3, nx, ny, dx, dy)
    ml = d.get_cg("lumpedcgmass").reshape(cg * ny + 1, cg * nx + 1)
    assert ml.sum() == pytest.approx(nx * ny * dx * dy, rel=1e-14)
    J = dx * dy
    if cg == 2:
        assert ml[2, 2] == pytest.approx(J / 9, rel=1e-14)      # interior vertex: 4 x 1/36
        assert ml[2, 3] == pytest.approx(2 * J / 9, rel=1e-14)  # interior edge midpoint: 2 x 1/9
        assert ml[3, 3] == pytest.approx(4 * J / 9, rel=1e-14)  # element centre
        assert ml[0, 0] == pytest.approx(J / 36, rel=1e-14)
    else:
        assert ml[2, 2] == pytest.approx(J, rel=1e-14)
        assert ml[0, 0] == pytest.approx(J / 4, rel=1e-14)


@pytest.mark.parametrize("cg,dg", [(1, 3), (2, 6)])
@pytest.mark.parametrize("skew", [0.0, 0.3])
def test_strain_of_linear_velocity_is_exact(cg, dg, skew):
    """u = (a x + b y, c x + d y): e11 = a, e22 = d, e12 = (b+c)/2 in component 0 only."""
    nx, ny, dx, dy = 6, 4, 1000.0, 1300.0
    d = _dyn(cg, dg, nx, ny, dx, dy, skew)
    X, Y = cg_node_coords(nx, ny, dx, dy, cg, skew)
    a, b, c, e = 2e-6, -1e-6, 3e-6, -0.5e-6
    d.set_cg("u", a * X + b * Y)
    d.set_cg("v", c * X + e * Y)
    d.run_pass("strain")
    S = d.S
    e11, e12, e22 = (d.get_field(n, S) for n in ("e11", "e12", "e22"))
    np.testing.assert_allclose(e11[:, 0], a, rtol=1e-12)
    np.testing.assert_allclose(e22[:, 0], e, rtol=1e-12)
    np.testing.assert_allclose(e12[:, 0], 0.5 * (b + c), rtol=1e-12)
    for f in (e11, e12, e22):
        assert np.abs(f[:, 1:]).max() < 1e-12 * 1e-6
    # rigid rotation: zero strain
    d.set_cg("u", -1e-6 * Y)
    d.set_cg("v", 1e-6 * X)
    d.run_pass("strain")
    for n in ("e11", "e12", "e22"):
        assert np.abs(d.get_field(n, S)).max() < 1e-18


def _hibler(e11, e12, e22, h, A, p):
    """Hibler (1979) VP stress, ellipse ratio e = 2, written from the strain invariants
    (divergence and maximum shear) rather than the component form the code uses."""
    div = e11 + e22
    shear = np.sqrt((e11 - e22) ** 2 + 4 * e12 ** 2)
    delta = np.sqrt(p["DeltaMin"] ** 2 + div ** 2 + shear ** 2 / 4.0)
    P = p["Pstar"] * h * np.exp(-20.0 * (1.0 - A))
    zeta = P / (2 * delta)
    eta = zeta / 4.0
    s11 = 2 * eta * e11 + (zeta - eta) * div - P / 2
    s22 = 2 * eta * e22 + (zeta - eta) * div - P / 2
    s12 = 2 * eta * e12
    return s11, s12, s22, P


@pytest.mark.parametrize("cg,dg", [(1, 3), (2, 6)])
def test_stress_relaxation_closed_form(cg, dg):
    """Constant strain, h, A: sigma_k = r^k sigma_0 + (1 - r^k) sigma_VP, r = 1 - 1/alpha."""
    nx, ny, dx = 4, 3, 1000.0
    d = _dyn(cg, dg, nx, ny, dx, 1200.0, skew=0.2)
    ne, S = nx * ny, d.S
    rng = np.random.default_rng(7)
    e = rng.normal(0, 3e-7, (3, ne))
    h = rng.uniform(0.5, 2.0, ne)
    A = rng.uniform(0.8, 1.0, ne)
    d.set_field("hice", h, 1)
    d.set_field("cice", A, 1)
    for k, n in enumerate(("e11", "e12", "e22")):
        f = np.zeros((ne, S))
        f[:, 0] = e[k]
        d.set_field(n, f, S)
    s0 = rng.normal(0, 100.0, (3, ne))
    for k, n in enumerate(("s11", "s12", "s22")):
        f = np.zeros((ne, S))
        f[:, 0] = s0[k]
        d.set_field(n, f, S)
    K = 37
    d.run_pass("stress", K)
    vp = _hibler(e[0], e[1], e[2], h, A, DEFAULT_PARAMS)[:3]
    r = (1 - 1 / DEFAULT_PARAMS["alpha"]) ** K
    for k, n in enumerate(("s11", "s12", "s22")):
        got = d.get_field(n, S)
        np.testing.assert_allclose(got[:, 0], r * s0[k] + (1 - r) * vp[k], rtol=1e-11, atol=1e-9)
        assert np.abs(got[:, 1:]).max() < 1e-9


def test_stress_lies_on_yield_ellipse_in_plastic_limit():
    """Large strain rates: principal stresses on Hibler's ellipse (e = 2)."""
    nx, ny = 3, 3
    p = dict(DEFAULT_PARAMS, alpha=1.0)  # one relaxation step reaches sigma_VP
    d = _dyn(2, 6, nx, ny, 1000.0, params=p)
    ne, S = nx * ny, d.S
    rng = np.random.default_rng(11)
    e = rng.normal(0, 1e-5, (3, ne))
    d.set_field("hice", np.full(ne, 1.0), 1)
    d.set_field("cice", np.full(ne, 1.0), 1)
    for k, n in enumerate(("e11", "e12", "e22")):
        f = np.zeros((ne, S))
        f[:, 0] = e[k]
        d.set_field(n, f, S)
    d.run_pass("stress")
    s11, s12, s22 = (d.get_field(n, S)[:, 0] for n in ("s11", "s12", "s22"))
    P = DEFAULT_PARAMS["Pstar"]
    sI = 0.5 * (s11 + s22)
    sII = np.sqrt(0.25 * (s11 - s22) ** 2 + s12 ** 2)
    ell = ((sI + P / 2) / (P / 2)) ** 2 + (sII / (P / 4)) ** 2
    np.testing.assert_allclose(ell, 1.0, rtol=1e-6)


def _no_dirichlet_dyn(cg, dg, nx, ny, dx, dy, skew=0.0):
    coords = box_coords(nx, ny, dx, dy, skew)
    m = OracleMesh.from_smesh(smesh_text(coords, nx, ny))
    d = OracleDyn(cg, dg, DEFAULT_PARAMS)
    d.initialise_mesh(m)
    return d


@pytest.mark.parametrize("cg,dg", [(1, 3), (2, 6)])
def test_divergence_partition_of_unity_and_constant_stress(cg, dg):
    nx, ny = 5, 4
    d = _no_dirichlet_dyn(cg, dg, nx, ny, 1000.0, 1400.0, skew=0.25)
    ne, S = nx * ny, d.S
    rng = np.random.default_rng(5)
    for n in ("s11", "s12", "s22"):
        d.set_field(n, rng.normal(0, 1e3, (ne, S)), S)
    d.run_pass("divergence")
    for n in ("dStressX", "dStressY"):
        v = d.get_cg(n)
        assert abs(v.sum()) < 1e-9 * np.abs(v).sum()
    # constant stress: interior nodes feel no force
    for k, n in enumerate(("s11", "s12", "s22")):
        f = np.zeros((ne, S))
        f[:, 0] = [250.0, -80.0, 120.0][k]
        d.set_field(n, f, S)
    d.run_pass("divergence")
    row = cg * nx + 1
    for n in ("dStressX", "dStressY"):
        v = d.get_cg(n).reshape(cg * ny + 1, row)
        inner = v[1:-1, 1:-1]
        assert np.abs(inner).max() < 1e-9 * np.abs(v).max()


def _dg_basis(j, x, y):
    X, Y = x - 0.5, y - 0.5
    return [1.0, X, Y, X * X - 1 / 12, Y * Y - 1 / 12, X * Y, Y * (X * X - 1 / 12), X * (Y * Y - 1 / 12)][j]


def _q2(j, x, y):
    def b(k, t):
        return [2 * (t - 0.5) * (t - 1), 4 * t * (1 - t), 2 * t * (t - 0.5)][k]

    def db(k, t):
        return [4 * t - 3, 4 - 8 * t, 4 * t - 1][k]

    jx, jy = j % 3, j // 3
    return b(jx, x) * b(jy, y), db(jx, x) * b(jy, y), b(jx, x) * db(jy, y)


def test_divergence_matches_brute_force_quadrature():
    """-(sigma, grad phi_j) assembled independently with numpy on a distorted 2x2 mesh."""
    nx = ny = 2
    coords = box_coords(nx, ny, 1000.0, 1000.0)
    coords[4] += [130.0, -70.0]  # move the centre vertex: bilinear (non-affine) elements
    m = OracleMesh.from_smesh(smesh_text(coords, nx, ny))
    d = OracleDyn(2, 6, DEFAULT_PARAMS)
    d.initialise_mesh(m)
    rng = np.random.default_rng(9)
    sig = {n: rng.normal(0, 1e3, (4, 8)) for n in ("s11", "s12", "s22")}
    for n, v in sig.items():
        d.set_field(n, v, 8)
    d.run_pass("divergence")
    gx, gw = np.polynomial.legendre.leggauss(3)  # same 3x3 rule as the reference
    gx, gw = 0.5 * gx + 0.5, 0.5 * gw
    fx = np.zeros(25)
    fy = np.zeros(25)
    for e in range(4):
        ix, iy = e % 2, e // 2
        nid = iy * 3 + ix
        c = coords[[nid, nid + 1, nid + 3, nid + 4]]
        for a in range(3):
            for b in range(3):
                xi, et = gx[a], gx[b]
                w = gw[a] * gw[b]
                # bilinear map and its Jacobian
                dN = np.array([[-(1 - et), -(1 - xi)], [(1 - et), -xi], [-et, (1 - xi)], [et, xi]])
                Jm = c.T @ dN  # 2x2: [[dx/dxi, dx/deta], [dy/dxi, dy/deta]]
                detJ = np.linalg.det(Jm)
                inv = np.linalg.inv(Jm)
                s = {n: sum(sig[n][e, k] * _dg_basis(k, xi, et) for k in range(8)) for n in sig}
                for j in range(9):
                    _, dxi, deta = _q2(j, xi, et)
                    grad = inv.T @ np.array([dxi, deta])
                    node = (2 * iy + j // 3) * 5 + 2 * ix + j % 3
                    fx[node] -= w * detJ * (s["s11"] * grad[0] + s["s12"] * grad[1])
                    fy[node] -= w * detJ * (s["s12"] * grad[0] + s["s22"] * grad[1])
    np.testing.assert_allclose(d.get_cg("dStressX"), fx, rtol=1e-10, atol=1e-9 * np.abs(fx).max())
    np.testing.assert_allclose(d.get_cg("dStressY"), fy, rtol=1e-10, atol=1e-9 * np.abs(fy).max())


def test_momentum_pseudo_time_converges_to_implicit_euler_balance():
    """sigma = 0 (P* = 0): the mEVP iteration's fixed point solves the backward-Euler
    momentum balance node by node, with the reference's same-component Coriolis term."""
    p = dict(DEFAULT_PARAMS, Pstar=0.0, beta=30.0)
    nx, ny = 4, 4
    d = _dyn(2, 6, nx, ny, 1000.0, params=p, mask=np.ones(16))
    ne = nx * ny
    rng = np.random.default_rng(2)
    d.set_field("hice", rng.uniform(0.5, 2, ne), 1)
    d.set_field("cice", rng.uniform(0.7, 1, ne), 1)
    d.set_field("uwind", rng.normal(5, 3, ne), 1)
    d.set_field("vwind", rng.normal(-3, 3, ne), 1)
    d.set_field("uocean", rng.normal(0, 0.05, ne), 1)
    d.set_field("vocean", rng.normal(0, 0.05, ne), 1)
    d.set_field("ssh", np.zeros(ne), 1)
    d.prepare_iteration()
    dt = 120.0
    u0 = rng.normal(0, 0.02, d.ncg)
    v0 = rng.normal(0, 0.02, d.ncg)
    d.set_cg("u", u0)
    d.set_cg("v", v0)
    d.begin_iteration(dt)
    d.run_pass("subcycle", 3000)
    u, v = d.get_cg("u"), d.get_cg("v")
    H, A = d.get_cg("cgH"), d.get_cg("cgA")
    ua, va, uo, vo = (d.get_cg(n) for n in ("uwind", "vwind", "uocean", "vocean"))
    rho, fc = p["rho_ice"], p["fc"]
    du = np.hypot(u - uo, v - vo)
    ta = np.hypot(ua, va)
    ru = rho * H * (u - u0) / dt - A * (p["F_atm"] * ta * ua + p["F_ocean"] * du * (uo - u)) + rho * H * fc * u
    rv = rho * H * (v - v0) / dt - A * (p["F_atm"] * ta * va + p["F_ocean"] * du * (vo - v)) - rho * H * fc * v
    free = np.abs(u) > 0  # Dirichlet nodes are held at zero
    scale = np.abs(A * p["F_atm"] * ta * ua).max()
    assert np.abs(ru[free]).max() < 1e-10 * scale
    assert np.abs(rv[free]).max() < 1e-10 * scale
    assert free.sum() > 0


def test_ssh_gradient_of_linear_height_is_exact_inside():
    nx, ny, dx = 10, 8, 1000.0
    d = _dyn(2, 6, nx, ny, dx)
    xc = (np.arange(nx) + 0.5) * dx
    yc = (np.arange(ny) + 0.5) * dx
    Xc, Yc = np.meshgrid(xc, yc)
    a, b = 3e-7, -2e-7
    d.set_field("ssh", (a * Xc + b * Yc).ravel(), 1)
    d.set_field("hice", np.ones(nx * ny), 1)
    d.set_field("cice", np.ones(nx * ny), 1)
    d.prepare_iteration()
    gx = d.get_cg("gradX").reshape(2 * ny + 1, 2 * nx + 1)
    gy = d.get_cg("gradY").reshape(2 * ny + 1, 2 * nx + 1)
    np.testing.assert_allclose(gx[4:-4, 4:-4], a, rtol=1e-9)
    np.testing.assert_allclose(gy[4:-4, 4:-4], b, rtol=1e-9)
