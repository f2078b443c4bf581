"""GPU parity: the gfx950 library (through its C ABI) against the FP64 CPU oracle on
the same seeded inputs, at sizes the oracle finishes in seconds, plus the reference's
own advection golden values run on the GPU.

Tolerances: FP64 path within 1e-10 relative L2 per field (u, v, h, A, sigma, strain) after 10
model steps; single passes within 1e-12; transport golden values at the reference's
1e-7 (Advection_test.cpp:26).
"""
import numpy as np
import pytest

from helpers import DEFAULT_PARAMS, box_coords, rel_l2, smesh_text
from oracle import Oracle, OracleDyn, OracleMesh, OracleTransport

pytestmark = pytest.mark.gpu

ORA = Oracle()


def _gpu():
    import nextsimdg_amd as nd

    return nd


def _say(msg):
    print(msg, flush=True)


# --------------------------------------------------------------------------- transport
ADV_UNIFORM = {1: 5.1256149074257538e-02, 3: 2.9450967798560313e-02, 6: 9.9340386651904228e-03}
ADV_DISTORTED = {1: 5.0958748236875594e-02, 3: 3.2033423984965226e-02, 6: 1.1830071142946669e-02}
ADV_RING = {1: 1.0338503986019776e+00, 3: 1.0949680727313791e+00, 6: 7.1704413076190610e-01}


def _gpu_mesh_from_oracle(om):
    nd = _gpu()
    nx, ny = om.shape
    m = nd.ParametricMesh(nx, ny, om.vertices())
    m.landmaskFromModelArray(om.landmask().astype(np.float64))
    for e in range(4):
        m.set_dirichlet(e, om.dirichlet(e))
    per = om.periodic()
    for seg in sorted(set(per[:, 0].tolist())) if len(per) else []:
        m.add_periodic(per[per[:, 0] == seg][:, 1:])
    return m


@pytest.mark.parametrize("dg", [1, 3, 6])
@pytest.mark.parametrize("distort", [0.0, 0.05])
def test_advection_golden_on_gpu(dg, distort):
    """Advection_test.cpp:105-265 at 12 x 13 with the transport running on the GPU."""
    _say(f"advection golden dg={dg} distort={distort}")
    nd = _gpu()
    om = OracleMesh.rectangle(409600.0, 512000.0, 12, 13, distort)
    deg = {1: 0, 3: 1, 6: 2}[dg]
    NT = 100 * (deg + 1) ** 2
    dt = 409600.0 / NT
    phi0 = ORA.function2dg(om, dg, 0)
    vx, vy = ORA.function2dg(om, dg, 1), ORA.function2dg(om, dg, 2)
    t = nd.DGTransport(_gpu_mesh_from_oracle(om), dg)
    t.settimesteppingscheme("rk2" if dg < 3 else "rk3")
    t.set_velocity(vx, vy)
    phi = t.run(dt, phi0, NT)
    err = np.sqrt(ORA.l2error_dg(om, dg, 0, phi)) / 409600.0
    gold = (ADV_UNIFORM if distort == 0.0 else ADV_DISTORTED)[dg]
    assert err == pytest.approx(gold, rel=1e-7)
    # and the field itself against the oracle's
    ot = OracleTransport(om, dg)
    ot.set_scheme(2 if dg < 3 else 3)
    ot.set_velocity(vx, vy)
    ref = phi0.copy()
    for _ in range(NT):
        ot.reinit()
        ref = ot.step(dt, ref)
    assert rel_l2(phi, ref) < 1e-12


@pytest.mark.parametrize("dg", [1, 3, 6])
def test_periodic_ring_golden_on_gpu(dg):
    """AdvectionPeriodicBC_test.cpp:117-276 (periodic seam + limiters) on the GPU."""
    _say(f"ring golden dg={dg}")
    nd = _gpu()
    om = OracleMesh.ring(32, 4)
    deg = {1: 0, 3: 1, 6: 2}[dg]
    NT = 50 * (deg + 1) ** 2
    dt = 250000.0 / NT
    phi0 = ORA.limit(dg, ORA.function2dg(om, dg, 3), 1.0, True)
    vx, vy = ORA.function2dg(om, dg, 4), ORA.function2dg(om, dg, 5)
    t = nd.DGTransport(_gpu_mesh_from_oracle(om), dg)
    t.settimesteppingscheme("rk2" if dg < 3 else "rk3")
    t.set_velocity(vx, vy)
    phi = t.run(dt, phi0, NT, limit_max=True, limit_min=True)
    err = np.sqrt(ORA.l2error_dg(om, dg, 3, phi)) / 100000.0
    assert err == pytest.approx(ADV_RING[dg], rel=1e-7)


@pytest.mark.parametrize("dg", [1, 3, 6])
def test_limiters_match_oracle(dg):
    nd = _gpu()
    rng = np.random.default_rng(dg)
    phi = rng.normal(0.8, 0.4, (777, dg))
    for is_max, bound in ((True, 1.0), (False, 0.0)):
        got = (nd.LimitMax if is_max else nd.LimitMin)(dg, phi, bound)
        ref = ORA.limit(dg, phi, bound, is_max)
        np.testing.assert_allclose(got, ref, rtol=1e-14, atol=1e-15)


# --------------------------------------------------------------------------- dynamics
CASES = [(2, 6), (1, 3)]


def _pair(cg, dg, nx, ny, dx=4000.0, mask=None, params=None, coords=None):
    nd = _gpu()
    p = dict(DEFAULT_PARAMS, **(params or {}))
    coords = box_coords(nx, ny, dx) if coords is None else coords
    mask = np.ones(nx * ny) if mask is None else mask
    o = OracleDyn(cg, dg, p)
    o.initialise(nx, ny, coords, mask)
    g = nd.MEVPDynamicsKernel(cg, dg, params=nd.default_params(**p))
    g.initialise(coords, False, mask, nx, ny)
    return o, g


def _box_mask(nx, ny):
    m = np.ones((ny, nx))
    m[0, :] = m[-1, :] = m[:, 0] = m[:, -1] = 0
    return m.ravel()


def _set_random_state(o, g, rng, S, dg):
    ne = o.ne
    h = np.abs(rng.normal(1.0, 0.3, (ne, dg))) * np.r_[1.0, [0.05] * (dg - 1)]
    A = np.clip(rng.normal(0.9, 0.05, (ne, dg)), 0, 1) * np.r_[1.0, [0.02] * (dg - 1)]
    for name, v, nc in (("hice", h, dg), ("cice", A, dg)):
        o.set_field(name, v, nc)
        g.setData(name, v, nc)
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


@pytest.mark.parametrize("cg,dg", CASES)
@pytest.mark.parametrize("nx,ny", [(3, 2), (31, 7), (32, 8), (62, 14), (63, 15)])
def test_dg2cg_tiles_match_oracle(cg, dg, nx, ny):
    """DG2CG (Interpolations.cpp:129-204) runs in tiles of 31 x 7 elements: meshes smaller
    than a tile, tile multiples and one past them. The forcing's DG0 -> CG transfer and the
    clamped cgH, cgA of prepareIteration against the oracle."""
    _say(f"dg2cg tiles cg={cg} dg={dg} {nx}x{ny}")
    o, g = _pair(cg, dg, nx, ny, mask=_box_mask(nx, ny) if nx > 2 and ny > 2 else np.ones(nx * ny),
                 coords=box_coords(nx, ny, 4000.0))
    _set_random_state(o, g, np.random.default_rng(nx * 100 + ny), o.S, dg)
    for n in ("uwind", "vwind", "uocean", "vocean"):
        assert rel_l2(g.get_cg(n), o.get_cg(n)) < 1e-15, n
    o.prepare_iteration()
    g.prepareIteration()
    for n in ("cgH", "cgA"):
        assert rel_l2(g.get_cg(n), o.get_cg(n)) < 1e-14, n


@pytest.mark.parametrize("cg,dg", CASES)
@pytest.mark.parametrize("uniform", [True, False])
def test_single_passes_match_oracle(cg, dg, uniform):
    _say(f"single passes cg={cg} dg={dg} uniform={uniform}")
    nx, ny = 20, 14
    coords = box_coords(nx, ny, 4000.0, 4000.0 if uniform else 3000.0, skew=0.0 if uniform else 0.2)
    if not uniform:
        coords[:, 0] += 150.0 * np.sin(coords[:, 1] / 9000.0)  # bilinear, non-affine elements
    o, g = _pair(cg, dg, nx, ny, mask=_box_mask(nx, ny), coords=coords)
    S = o.S
    _set_random_state(o, g, np.random.default_rng(42 + cg), S, dg)
    # setData of the forcing goes through DG2CG
    for n in ("uwind", "vwind", "uocean", "vocean"):
        assert rel_l2(g.get_cg(n), o.get_cg(n)) < 1e-15
    np.testing.assert_array_equal(g.get_cg("lumpedcgmass") > 0, True)
    assert rel_l2(g.get_cg("lumpedcgmass"), o.get_cg("lumpedcgmass")) < 1e-14
    o.prepare_iteration()
    g.prepareIteration()
    for n in ("cgH", "cgA", "gradX", "gradY"):
        assert rel_l2(g.get_cg(n), o.get_cg(n)) < 1e-13, n
    o.begin_iteration(120.0)
    g.begin_iteration(120.0)
    o.run_pass("strain")
    g.projectVelocityToStrain()
    for n in ("e11", "e12", "e22"):
        assert rel_l2(g.getDGData(n, S), o.get_field(n, S)) < 1e-12, n
    o.run_pass("stress", 3)
    g.stressUpdateHighOrder(3)
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
    # DG outputs: cell means of u, v and the ice-ocean stress through CG2DG
    for n in ("u", "v", "uiostress", "viostress"):
        assert rel_l2(g.getDG0Data(n), o.get_field(n, 1)) < 1e-12, n
        assert rel_l2(g.getDGData(n, dg), o.get_field(n, dg)) < 1e-12, n


@pytest.mark.parametrize("mesh", ["box", "general", "sphere"])
@pytest.mark.parametrize("cg,dg", CASES)
def test_fused_subcycles_equal_split_passes(cg, dg, mesh):
    """The one-pass subcycle kernels (literal operators on the uniform box, streamed
    geometry on a non-affine mesh and on a spherical patch) reproduce the per-pass
    sequence of the reference run with the per-element operator tables."""
    _say(f"fused vs split cg={cg} dg={dg} mesh={mesh}")
    nd = _gpu()
    from nextsimdg_amd import inputs

    nx, ny = 37, 29  # not a multiple of the tile
    sph = mesh == "sphere"
    if mesh == "box":
        coords = box_coords(nx, ny, 2000.0)
    elif mesh == "general":
        coords = box_coords(nx, ny, 2000.0, 1700.0, skew=0.15)
        coords[:, 0] += 120.0 * np.sin(coords[:, 1] / 9000.0)  # bilinear, non-affine elements
    else:
        coords = inputs.sphere_patch_coords(nx, ny)
    mask = _box_mask(nx, ny)
    mask[5 * nx + 7] = 0.0  # an interior land cell
    p = nd.default_params()
    a = nd.MEVPDynamicsKernel(cg, dg, params=p)
    b = nd.MEVPDynamicsKernel(cg, dg, params=p)
    o = OracleDyn(cg, dg, DEFAULT_PARAMS)
    o.initialise(nx, ny, coords, mask, spherical=sph)
    for k in (a, b):
        k.initialise(coords, sph, mask, nx, ny)
    rng = np.random.default_rng(3)
    _set_random_state(o, a, rng, o.S, dg)
    _set_random_state(o, b, np.random.default_rng(3), o.S, dg)
    for k in (a, b):
        k.prepareIteration()
        k.begin_iteration(120.0)
    a.subcycles(7)
    for _ in range(7):
        b.projectVelocityToStrain()
        b.stressUpdateHighOrder(1)
        b.stressDivergence()
        b.updateMomentum()
        b.applyBoundaries()
    a.synchronize()
    # the box kernel associates the same sums as the split passes up to rounding (1e-13);
    # the geometry kernels apply the factored operators (products of the same per-Gauss-point
    # terms, in another order), which leaves a few more roundings (1e-12)
    tol = 1e-13 if mesh == "box" else 1e-12
    errs = {n: rel_l2(a.get_cg(n), b.get_cg(n)) for n in ("u", "v")}
    errs.update({n: rel_l2(a.getDGData(n, o.S), b.getDGData(n, o.S)) for n in ("s11", "s12", "s22", "e11", "e12", "e22")})
    _say("  " + " ".join(f"{n}={e:.1e}" for n, e in errs.items()))
    assert np.abs(a.get_cg("u")).max() > 0.0
    for n, e in errs.items():
        assert e < tol, (n, e)


RAGGED = [(2, 2), (2, 7), (7, 2), (3, 2), (31, 7), (32, 8), (62, 14), (63, 15)]


def _mesh_coords(mesh, nx, ny):
    from nextsimdg_amd import inputs

    if mesh == "box":
        return box_coords(nx, ny, 2000.0)
    if mesh == "general":
        c = box_coords(nx, ny, 2000.0, 1700.0, skew=0.15)
        c[:, 0] += 120.0 * np.sin(c[:, 1] / 9000.0)
        return c
    return inputs.sphere_patch_coords(nx, ny)


@pytest.mark.parametrize("mesh", ["box", "general", "sphere"])
@pytest.mark.parametrize("nx,ny", RAGGED)
def test_fused_subcycles_on_ragged_meshes(nx, ny, mesh):
    """Meshes smaller than one tile (down to the smallest the library takes, 2 x 2), exact
    tile multiples and one element past them: the one-pass kernels (tile halo clamping,
    the last tile's domain-edge nodes) against the split passes, with a land cell where
    the mesh has room for one."""
    cg, dg, S = 2, 6, 8
    _say(f"ragged {nx}x{ny} {mesh}")
    nd = _gpu()
    coords = _mesh_coords(mesh, nx, ny)
    mask = np.ones(nx * ny)
    if nx * ny >= 6:
        mask[(ny // 2) * nx + nx // 2] = 0.0
    p = nd.default_params()
    a, b = nd.MEVPDynamicsKernel(cg, dg, params=p), nd.MEVPDynamicsKernel(cg, dg, params=p)
    ne, ncg = nx * ny, (cg * nx + 1) * (cg * ny + 1)
    rng = np.random.default_rng(nx * 100 + ny)
    state = {"hice": np.abs(rng.normal(1.0, 0.3, (ne, dg))) * np.r_[1.0, [0.05] * (dg - 1)],
             "cice": np.clip(rng.normal(0.9, 0.05, (ne, dg)), 0, 1) * np.r_[1.0, [0.02] * (dg - 1)]}
    for n in ("s11", "s12", "s22", "e11", "e12", "e22"):
        state[n] = rng.normal(0, 1e3 if n[0] == "s" else 1e-7, (ne, S))
    forcing = {n: rng.normal(0, sc, ne) for n, sc in (("uwind", 10), ("vwind", 10), ("uocean", 0.05),
                                                       ("vocean", 0.05), ("ssh", 0.1))}
    vel = {n: rng.normal(0, 0.1, ncg) for n in ("u", "v")}
    for k in (a, b):
        k.initialise(coords, mesh == "sphere", mask, nx, ny)
        for n, v in state.items():
            k.setData(n, v, v.shape[1])
        for n, v in forcing.items():
            k.setData(n, v, 1)
        for n, v in vel.items():
            k.set_cg(n, v)
        k.prepareIteration()
        k.begin_iteration(120.0)
    a.subcycles(5)
    for _ in range(5):
        b.projectVelocityToStrain()
        b.stressUpdateHighOrder(1)
        b.stressDivergence()
        b.updateMomentum()
        b.applyBoundaries()
    a.synchronize()
    tol = 1e-13 if mesh == "box" else 1e-12
    for n in ("u", "v"):
        ga, gb = a.get_cg(n), b.get_cg(n)
        assert np.isfinite(ga).all()
        assert np.abs(ga - gb).max() <= tol * max(np.abs(gb).max(), 1e-30), n
    for n in ("s11", "s12", "s22", "e11", "e12", "e22"):
        assert rel_l2(a.getDGData(n, S), b.getDGData(n, S)) < tol, n


def _module_step_inputs(nx, ny, steps, dt, seed):
    from nextsimdg_amd import inputs

    setup = inputs.benchmark_setup(nx, ny, seed)
    xoff, yoff = inputs.cyclone_offset(seed)
    winds = [ORA.benchmark_wind(nx, ny, inputs.L_BOX, k * dt, xoff, yoff, 30.0) for k in range(steps)]
    return setup, winds


def _run_module_level(o, g, setup, winds, dt):
    """MEVPDynamics::update (MEVPDynamics.cpp:59-87) on both sides with the same inputs."""
    h0, A0 = setup["h"], setup["A"]
    gh, gA, oh, oA = h0, A0, h0, A0
    for name in ("uocean", "vocean"):
        o.set_field(name, setup[name], 1)
        g.setData(name, setup[name], 1)
    zero = np.zeros_like(h0)
    for k, (uw, vw) in enumerate(winds):
        o.set_field("hice", oh, 1)
        o.set_field("cice", oA, 1)
        g.setData("hice", gh, 1)
        g.setData("cice", gA, 1)
        for name, v in (("uwind", uw), ("vwind", vw), ("ssh", zero)):
            o.set_field(name, v, 1)
            g.setData(name, v, 1)
        o.update(dt)
        g.update(dt)
        oh, oA = o.get_field("hice", 1), o.get_field("cice", 1)
        gh, gA = g.getDG0Data("hice"), g.getDG0Data("cice")
        _say(f"  step {k}: du={rel_l2(g.get_cg('u'), o.get_cg('u')):.3e}")


@pytest.mark.parametrize("cg,dg", CASES)
def test_ten_model_steps_match_oracle(cg, dg):
    """10 module-level steps of the benchmark box: <= 1e-10 relative L2 per field."""
    _say(f"10 steps cg={cg} dg={dg}")
    from nextsimdg_amd import inputs

    nx = ny = 24 if cg == 2 else 32
    dt = 120.0
    o, g = _pair(cg, dg, nx, ny, coords=inputs.box_coords(nx, ny), mask=inputs.box_mask(nx, ny))
    setup, winds = _module_step_inputs(nx, ny, 10, dt, seed=5)
    _run_module_level(o, g, setup, winds, dt)
    S = o.S
    errs = {n: rel_l2(g.get_cg(n), o.get_cg(n)) for n in ("u", "v")}
    errs.update({n: rel_l2(g.getDGData(n, dg), o.get_field(n, dg)) for n in ("hice", "cice")})
    # the strain output is that of the velocity entering the last subcycle
    errs.update({n: rel_l2(g.getDGData(n, S), o.get_field(n, S)) for n in ("s11", "s12", "s22", "e11", "e12", "e22")})
    _say(f"  rel L2 errors: {errs}")
    assert np.abs(g.get_cg("u")).max() > 1e-3  # the cyclone moved the ice
    for n, e in errs.items():
        assert e < 1e-10, (n, e)


def test_general_mesh_steps_match_oracle():
    """Per-element operators streamed (non-uniform mesh) through two full steps."""
    _say("general mesh steps")
    nx, ny = 18, 16
    coords = box_coords(nx, ny, 5000.0, 4000.0)
    coords[:, 0] += 300.0 * np.sin(coords[:, 1] / 11000.0) * np.sin(coords[:, 0] / 13000.0)
    mask = _box_mask(nx, ny)
    o, g = _pair(2, 6, nx, ny, coords=coords, mask=mask)
    rng = np.random.default_rng(8)
    h = rng.uniform(0.5, 1.5, nx * ny)
    A = rng.uniform(0.8, 1.0, nx * ny)
    uw, vw = rng.normal(8, 2, nx * ny), rng.normal(-5, 2, nx * ny)
    for name, v in (("hice", h), ("cice", A), ("uwind", uw), ("vwind", vw), ("uocean", 0 * h), ("vocean", 0 * h),
                    ("ssh", 0 * h)):
        o.set_field(name, v, 1)
        g.setData(name, v, 1)
    for _ in range(2):
        o.update(120.0)
        g.update(120.0)
    for n in ("u", "v"):
        assert rel_l2(g.get_cg(n), o.get_cg(n)) < 1e-10, n
    for n in ("hice", "cice"):
        assert rel_l2(g.getDGData(n, 6), o.get_field(n, 6)) < 1e-10, n


def test_periodic_mesh_dynamics_match_oracle():
    """A mesh with a periodic seam uses the split node passes; compare two steps."""
    _say("periodic mesh dynamics")
    nd = _gpu()
    nx, ny = 12, 10
    coords = box_coords(nx, ny, 4000.0)
    per = [[(j + 1) * nx - 1, j * nx, 1] for j in range(ny)]
    text = smesh_text(coords, nx, ny, dirichlet=[(i, 0) for i in range(nx)] + [(nx * (ny - 1) + i, 2) for i in range(nx)],
                      periodic=[per])
    om = OracleMesh.from_smesh(text)
    o = OracleDyn(2, 6, DEFAULT_PARAMS)
    o.initialise_mesh(om)
    g = nd.MEVPDynamicsKernel(2, 6)
    g.initialise_mesh(_gpu_mesh_from_oracle(om))
    rng = np.random.default_rng(4)
    for name, v in (("hice", rng.uniform(0.5, 1.5, nx * ny)), ("cice", rng.uniform(0.8, 1, nx * ny)),
                    ("uwind", rng.normal(8, 2, nx * ny)), ("vwind", rng.normal(-3, 2, nx * ny)),
                    ("uocean", np.zeros(nx * ny)), ("vocean", np.zeros(nx * ny)), ("ssh", np.zeros(nx * ny))):
        o.set_field(name, v, 1)
        g.setData(name, v, 1)
    for _ in range(2):
        o.update(120.0)
        g.update(120.0)
    for n in ("u", "v"):
        assert rel_l2(g.get_cg(n), o.get_cg(n)) < 1e-10, n
    assert rel_l2(g.getDGData("hice", 6), o.get_field("hice", 6)) < 1e-10


def test_benchmark_wind_kernel_matches_oracle_forcing():
    from nextsimdg_amd import inputs

    nx = ny = 40
    o, g = _pair(2, 6, nx, ny, coords=inputs.box_coords(nx, ny), mask=inputs.box_mask(nx, ny))
    for t in (0.0, 7200.0, 86400.0 * 1.5):
        uw, vw = ORA.benchmark_wind(nx, ny, inputs.L_BOX, t, 1000.0, -2000.0, 30.0)
        o.set_field("uwind", uw, 1)
        o.set_field("vwind", vw, 1)
        g.benchmark_wind(inputs.L_BOX, t, 1000.0, -2000.0, 30.0)
        for n in ("uwind", "vwind"):
            assert rel_l2(g.get_cg(n), o.get_cg(n)) < 1e-14, (n, t)


def test_mass_sums_match_tools_meanvalue():
    from nextsimdg_amd import inputs

    nx = ny = 30
    o, g = _pair(2, 6, nx, ny, coords=inputs.box_coords(nx, ny), mask=inputs.box_mask(nx, ny))
    rng = np.random.default_rng(1)
    h = rng.normal(1.0, 0.1, (nx * ny, 6))
    A = rng.normal(0.9, 0.05, (nx * ny, 6))
    g.setData("hice", h, 6)
    g.setData("cice", A, 6)
    om = OracleMesh.from_coords(nx, ny, inputs.box_coords(nx, ny), inputs.box_mask(nx, ny))
    vol, area = g.mass()
    assert vol == pytest.approx(ORA.mean_value(om, 6, h), rel=1e-13)
    assert area == pytest.approx(ORA.mean_value(om, 6, A), rel=1e-13)


def test_state_round_trip_and_dg0_reset():
    from nextsimdg_amd import inputs

    nx = ny = 16
    nd = _gpu()
    g = nd.MEVPDynamicsKernel(2, 6)
    g.initialise(inputs.box_coords(nx, ny), False, inputs.box_mask(nx, ny), nx, ny)
    rng = np.random.default_rng(2)
    g.setData("hice", rng.normal(1, 0.1, (nx * ny, 6)), 6)
    g.setData("s11", rng.normal(0, 1, (nx * ny, 8)), 8)
    g.set_cg("u", rng.normal(0, 1, g.ncg))
    blob = g.save_state()
    h1, s1, u1 = g.getDGData("hice"), g.getDGData("s11"), g.get_cg("u")
    g.setData("hice", np.zeros(nx * ny), 1)
    g.set_cg("u", np.zeros(g.ncg))
    g.load_state(blob)
    np.testing.assert_array_equal(g.getDGData("hice"), h1)
    np.testing.assert_array_equal(g.getDGData("s11"), s1)
    np.testing.assert_array_equal(g.get_cg("u"), u1)
    g.reset_dg0("hice")
    d = g.getDGData("hice")
    np.testing.assert_array_equal(d[:, 0], h1[:, 0])
    assert np.all(d[:, 1:] == 0)
    with pytest.raises(Exception):
        g.getDG0Data("no_such_field")


@pytest.mark.parametrize("cg,dg", [(2, 6), (1, 3)])
@pytest.mark.parametrize("nx,ny", [(2, 7), (7, 2), (2, 2), (3, 9), (13, 13), (14, 14), (26, 26), (27, 27),
                                   (40, 29), (31, 7)])
def test_paired_subcycles_equal_single_passes(cg, dg, nx, ny):
    """Two subcycles per pass (k_mevp_pair, forced on), against one per pass (k_mevp_fused,
    forced on) on the same state, on two-element-wide meshes (the smallest the library
    takes), meshes smaller than a pair tile (13 x 13), exact tile multiples and one element
    past them, with land cells. Nine subcycles: four pairs and the last one alone.

    The pair kernel runs the arithmetic of the single one in the same order, but the
    compiler may fuse a multiply-add differently in the two kernels: single values then
    differ by an ulp (1.4e-16 relative in u measured). Bound: 1e-14 relative (max norm)
    per field. A second pair run must give the same bits (no race between the sub-steps)."""
    _say(f"pairs vs single passes cg={cg} dg={dg} {nx}x{ny}")
    nd = _gpu()
    S = 8 if cg == 2 else 3
    coords = box_coords(nx, ny, 2000.0)
    mask = np.ones(nx * ny)
    if nx * ny >= 6:
        mask[(ny // 2) * nx + nx // 2] = 0.0
        mask[0] = 0.0
    ne, ncg = nx * ny, (cg * nx + 1) * (cg * ny + 1)
    rng = np.random.default_rng(nx * 1000 + ny * 10 + cg)
    state = {"hice": np.abs(rng.normal(1.0, 0.3, (ne, dg))) * np.r_[1.0, [0.05] * (dg - 1)],
             "cice": np.clip(rng.normal(0.9, 0.05, (ne, dg)), 0, 1) * np.r_[1.0, [0.02] * (dg - 1)]}
    for n in ("s11", "s12", "s22", "e11", "e12", "e22"):
        state[n] = rng.normal(0, 1e3 if n[0] == "s" else 1e-7, (ne, S))
    forcing = {n: rng.normal(0, sc, ne) for n, sc in (("uwind", 10), ("vwind", 10), ("uocean", 0.05),
                                                       ("vocean", 0.05), ("ssh", 0.1))}
    vel = {n: rng.normal(0, 0.1, ncg) for n in ("u", "v")}
    ks = []
    for mode in (1, 1, 2):  # NSDG_PAIR_ALWAYS twice, NSDG_PAIR_NEVER
        k = nd.MEVPDynamicsKernel(cg, dg)
        k.initialise(coords, False, mask, nx, ny)
        k.set_pairing(mode)
        for n, v in state.items():
            k.setData(n, v, v.shape[1])
        for n, v in forcing.items():
            k.setData(n, v, 1)
        for n, v in vel.items():
            k.set_cg(n, v)
        k.prepareIteration()
        k.begin_iteration(120.0)
        k.subcycles(9)
        k.synchronize()
        ks.append(k)
    a, a2, b = ks
    if min(nx, ny) > 2:  # a developed state (two-wide meshes are mostly Dirichlet nodes)
        assert np.abs(a.get_cg("u")).max() > 1e-3

    def fields(k):
        out = {n: k.get_cg(n) for n in ("u", "v")}
        out.update({n: k.getDGData(n, S) for n in ("s11", "s12", "s22", "e11", "e12", "e22")})
        return out

    fa, fa2, fb = fields(a), fields(a2), fields(b)
    for n in fa:
        assert np.array_equal(fa[n], fa2[n]), n
        err = np.abs(fa[n] - fb[n]).max() / max(np.abs(fb[n]).max(), 1e-300)
        assert err <= 1e-14, (n, err)


def test_pairing_mode_is_checked():
    nd = _gpu()
    k = nd.MEVPDynamicsKernel(2, 6)
    with pytest.raises(nd.NsdgError):
        k.set_pairing(3)
