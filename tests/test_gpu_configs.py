"""The benchmark configurations' own checks at their own sizes.

Config 2 (512 x 512 box, dx = 1 km, CG2 / DG2 / 8-dof stress, FP64, seeded cyclone track).
At this resolution the fixed mEVP parameters (alpha = beta = 1500, 100 subcycles,
VPCGDynamicsKernel.hpp:125-126) leave the pseudo-time iteration amplifying perturbations
by roughly e^0.2 per subcycle, so two FP64 codes that round one operation differently
drift apart within a step. The oracle shows it against itself: with h perturbed by one
ulp it departs by ~1e-8 after one step and ~1e-4 after ten. Hence two checks:
  * from one shared developed state, 1 and 5 subcycles of the GPU equal the oracle's
    within 1e-12 and 1e-11 relative L2 per field (u, v, sigma): the arithmetic itself;
  * after 10 module-level steps (MEVPDynamics::update, MEVPDynamics.cpp:59-87) from the
    same inputs, every field (u, v, h, A, sigma) lies within 1e-10 or within 10x the
    spread of the oracle under that one-ulp perturbation, whichever is larger.

Config 5 (2048 x 2048, three seeded storms of 10-25 m/s -- copies of the reference cyclone,
BenchmarkAtmosphere.cpp:29-76 -- heterogeneous ice with open-water leads, FP32 bulk): the
storm forcing equals the oracle's cyclones summed; five steps stay finite with exact
Dirichlet zeros, repeat bit for bit, close the mass budget to 1e-13, and keep FP32 within
3x the FP64 path's own spread.

Config 3 (1024 x 1024, heterogeneous ice, FP32 variant): the same amplification holds, so
  * from one shared FP64 state, 1 and 10 FP32 subcycles stay within 1e-5 relative L2 of
    FP64 per field: the FP32 arithmetic's own error;
  * after 1 and 10 model steps the FP32 error stays within 3x the FP64 path's own spread
    when h is perturbed by a relative 1e-7 (one FP32 rounding) -- at this size that spread
    is ~4e-4 in u after one step, far above the 1e-5 the FP32 arithmetic alone produces;
  * at 24 x 24, where the iteration is well conditioned, FP32 on the GPU stays within 1e-5
    of the FP64 oracle after 10 steps.
"""
import numpy as np
import pytest

from helpers import DEFAULT_PARAMS, rel_l2
from oracle import Oracle, OracleDyn

pytestmark = pytest.mark.gpu

ORA = Oracle()
S_NAMES = ("s11", "s12", "s22")


def _say(msg):
    print(msg, flush=True)


def _gpu(nx, s, prec=0):
    import nextsimdg_amd as nd

    g = nd.MEVPDynamicsKernel(2, 6)
    g.initialise(s["coords"], False, s["mask"], nx, nx)
    for name in ("uocean", "vocean"):
        g.setData(name, s[name], 1)
    if prec:
        g.set_precision(prec)
    return g


def _oracle(nx, s):
    o = OracleDyn(2, 6, dict(DEFAULT_PARAMS))
    o.initialise(nx, nx, s["coords"], s["mask"])
    for name in ("uocean", "vocean"):
        o.set_field(name, s[name], 1)
    return o


def _module_steps(k, setf, getf, h, A, winds, dt=120.0):
    """MEVPDynamics::update per step: DG0 h and A, forcing, update, DG0 read-back."""
    zero = np.zeros_like(h)
    for uw, vw in winds:
        setf("hice", h, 1)
        setf("cice", A, 1)
        for name, v in (("uwind", uw), ("vwind", vw), ("ssh", zero)):
            setf(name, v, 1)
        k.update(dt)
        h, A = getf("hice"), getf("cice")
    return h, A


def _fields_gpu(g):
    f = {n: g.get_cg(n) for n in ("u", "v")}
    f.update({n: g.getDGData(n, 6) for n in ("hice", "cice")})
    f.update({n: g.getDGData(n, 8) for n in S_NAMES})
    return f


def _fields_oracle(o):
    f = {n: o.get_cg(n) for n in ("u", "v")}
    f.update({n: o.get_field(n, 6) for n in ("hice", "cice")})
    f.update({n: o.get_field(n, 8) for n in S_NAMES})
    return f


def _winds(nx, steps, seed, dt=120.0):
    from nextsimdg_amd import inputs

    xoff, yoff = inputs.cyclone_offset(seed)
    return [ORA.benchmark_wind(nx, nx, inputs.L_BOX, k * dt, xoff, yoff, 30.0) for k in range(steps)]


def test_config2_subcycles_at_512_match_oracle_from_a_shared_state():
    from nextsimdg_amd import inputs

    nx, seed = 512, 7
    _say(f"config 2: {nx}^2 subcycles from a shared state")
    s = inputs.benchmark_setup(nx, nx, seed)
    g = _gpu(nx, s)
    winds = _winds(nx, 4, seed)
    _module_steps(g, g.setData, g.getDG0Data, s["h"], s["A"], winds[:3])
    o = _oracle(nx, s)
    # the GPU's developed state, handed to the oracle whole
    for n in ("hice", "cice"):
        o.set_field(n, g.getDGData(n, 6), 6)
    for n in S_NAMES + ("e11", "e12", "e22"):
        o.set_field(n, g.getDGData(n, 8), 8)
    for n in ("u", "v"):
        o.set_cg(n, g.get_cg(n))
    uw, vw = winds[3]
    for k, setf in ((g, g.setData), (o, o.set_field)):
        setf("uwind", uw, 1)
        setf("vwind", vw, 1)
        setf("ssh", np.zeros(nx * nx), 1)
        k.prepareIteration() if k is g else k.prepare_iteration()
        k.begin_iteration(120.0)
    for total, tol in ((1, 1e-12), (5, 1e-11)):
        g.subcycles(1 if total == 1 else 4)
        o.run_pass("subcycle", 1 if total == 1 else 4)
        errs = {n: rel_l2(g.get_cg(n), o.get_cg(n)) for n in ("u", "v")}
        errs.update({n: rel_l2(g.getDGData(n, 8), o.get_field(n, 8)) for n in S_NAMES})
        _say(f"  {total} subcycles: " + " ".join(f"{n}={e:.1e}" for n, e in errs.items()))
        assert np.abs(g.get_cg("u")).max() > 1e-3
        for n, e in errs.items():
            assert e < tol, (total, n, e)


def test_config2_ten_steps_at_512_within_the_oracles_own_spread():
    from nextsimdg_amd import inputs

    nx, steps, seed = 512, 10, 7
    _say(f"config 2: {nx}^2, {steps} module-level steps against the oracle")
    s = inputs.benchmark_setup(nx, nx, seed)
    winds = _winds(nx, steps, seed)
    g = _gpu(nx, s)
    _module_steps(g, g.setData, g.getDG0Data, s["h"], s["A"], winds)
    res = {"gpu": _fields_gpu(g)}
    del g
    ulp = np.where(np.random.default_rng(seed).random(nx * nx) < 0.5, -1.0, 1.0) * np.finfo(float).eps
    for tag, h in (("oracle", s["h"]), ("oracle_ulp", s["h"] * (1.0 + ulp))):
        o = _oracle(nx, s)
        _module_steps(o, o.set_field, lambda n: o.get_field(n, 1), h, s["A"], winds)
        res[tag] = _fields_oracle(o)
        del o
    err = {n: rel_l2(res["gpu"][n], res["oracle"][n]) for n in res["gpu"]}
    spread = {n: rel_l2(res["oracle_ulp"][n], res["oracle"][n]) for n in res["gpu"]}
    _say("  gpu vs oracle:    " + " ".join(f"{n}={e:.2e}" for n, e in err.items()))
    _say("  oracle 1-ulp h:   " + " ".join(f"{n}={e:.2e}" for n, e in spread.items()))
    assert np.abs(res["oracle"]["u"]).max() > 1e-3  # the cyclone moved the ice
    for n in err:
        assert err[n] <= max(1e-10, 10.0 * spread[n]), (n, err[n], spread[n])


def _fp32_state(nx, kind, seed, prec, perturb=0.0, steps=2):
    import bench

    k, off, L = bench.make_kernel(nx, nx, seed, prec, kind, perturb)
    bench.run_steps(k, off, L, 0, steps, 120.0)
    return k, off, L


def test_config3_fp32_subcycles_at_1024_from_a_shared_state():
    nx = 1024
    _say(f"config 3: FP32 subcycles at {nx}^2 from a shared FP64 state")
    ks = {}
    for prec in (0, 1):
        k, off, L = _fp32_state(nx, "heterogeneous", 0, 0)
        k.reset_dg0("hice")
        k.reset_dg0("cice")
        k.benchmark_wind(L, 2 * 120.0, off[0], off[1], 30.0)
        k.advectionAndLimits(120.0)
        k.prepareIteration()
        k.begin_iteration(120.0)
        if prec:
            k.set_precision(1)
        ks[prec] = k
    done = 0
    for total in (1, 10):
        for k in ks.values():
            k.subcycles(total - done)
        done = total
        errs = {n: rel_l2(ks[1].get_cg(n), ks[0].get_cg(n)) for n in ("u", "v")}
        errs.update({n: rel_l2(ks[1].getDGData(n, 8), ks[0].getDGData(n, 8)) for n in S_NAMES})
        _say(f"  {total} subcycles: " + " ".join(f"{n}={e:.1e}" for n, e in errs.items()))
        for n, e in errs.items():
            assert e < 1e-5, (total, n, e)
        assert max(errs.values()) > 1e-12  # FP32 arithmetic, not the FP64 kernels again


def test_config3_fp32_steps_at_1024_within_the_conditioning_floor():
    import bench

    nx = 1024
    _say(f"config 3: FP32 steps at {nx}^2 against FP64 and its 1e-7 floor")
    ks = {tag: bench.make_kernel(nx, nx, 0, prec, "heterogeneous", pert)
          for tag, prec, pert in (("fp64", 0, 0.0), ("fp32", 1, 0.0), ("floor", 0, 1e-7))}
    done = 0
    for total in (1, 10):
        f = {}
        for tag, (k, off, L) in ks.items():
            bench.run_steps(k, off, L, done, total - done, 120.0)
            f[tag] = _fields_gpu(k)
        done = total
        e32 = {n: rel_l2(f["fp32"][n], f["fp64"][n]) for n in f["fp64"]}
        efl = {n: rel_l2(f["floor"][n], f["fp64"][n]) for n in f["fp64"]}
        _say(f"  {total} steps fp32:  " + " ".join(f"{n}={e:.1e}" for n, e in e32.items()))
        _say(f"  {total} steps floor: " + " ".join(f"{n}={e:.1e}" for n, e in efl.items()))
        for n in e32:
            assert e32[n] <= max(1e-5, 3.0 * efl[n]), (total, n, e32[n], efl[n])


def test_config3_fp32_against_the_oracle_at_24():
    from nextsimdg_amd import inputs

    nx, steps, seed = 24, 10, 3
    _say(f"FP32 GPU against the FP64 oracle at {nx}^2, {steps} steps")
    s = inputs.benchmark_setup(nx, nx, seed, kind="heterogeneous")
    winds = _winds(nx, steps, seed)
    g = _gpu(nx, s, prec=1)
    _module_steps(g, g.setData, g.getDG0Data, s["h"], s["A"], winds)
    o = _oracle(nx, s)
    _module_steps(o, o.set_field, lambda n: o.get_field(n, 1), s["h"], s["A"], winds)
    fg, fo = _fields_gpu(g), _fields_oracle(o)
    errs = {n: rel_l2(fg[n], fo[n]) for n in fo}
    _say("  " + " ".join(f"{n}={e:.1e}" for n, e in errs.items()))
    assert np.abs(fo["u"]).max() > 1e-3
    for n, e in errs.items():
        assert e < 1e-5, (n, e)
    assert max(errs.values()) > 1e-12


def test_storm_wind_is_the_sum_of_reference_cyclones():
    """nsdg_dyn_storm_wind against the oracle's cyclone (BenchmarkAtmosphere.cpp:29-76)
    summed over the storms, on the CG nodes after the DG0 -> CG transfer."""
    import nextsimdg_amd as nd
    from nextsimdg_amd import inputs

    _say("storm wind against summed oracle cyclones")
    nx = ny = 40
    s = inputs.benchmark_setup(nx, ny, 1)
    st = inputs.storms(9)
    g = _gpu(nx, s)
    o = _oracle(nx, s)
    for t in (0.0, 3 * 3600.0, 2.5 * 86400.0):
        uw = sum(ORA.benchmark_wind(nx, ny, inputs.L_BOX, t, x, y, v)[0] for x, y, v in st)
        vw = sum(ORA.benchmark_wind(nx, ny, inputs.L_BOX, t, x, y, v)[1] for x, y, v in st)
        o.set_field("uwind", uw, 1)
        o.set_field("vwind", vw, 1)
        g.storm_wind(inputs.L_BOX, t, st)
        for n in ("uwind", "vwind"):
            assert rel_l2(g.get_cg(n), o.get_cg(n)) < 1e-14, (n, t)
    # peak speeds in the configured 10-25 m/s range (vmax / e at 100 km from a centre)
    assert np.all((st[:, 2] / np.e >= 10.0) & (st[:, 2] / np.e <= 25.0))


def test_config5_storms_and_leads_at_2048():
    """Config 5's workload at its size: 2048 x 2048, three seeded storms, heterogeneous ice
    with open-water leads, FP32 bulk. Five steps: finite, Dirichlet nodes exactly zero,
    bit-for-bit repeatable, the advection mass budget closed to 1e-13 in every step, and
    the FP32 state within 3x the FP64 path's own spread under a 1e-7 perturbation of h
    (or 1e-5)."""
    import bench
    from nextsimdg_amd import inputs

    nx, steps = 2048, 5
    _say(f"config 5: {nx}^2, 3 storms, leads, FP32, {steps} steps")
    st = inputs.storms(5)
    runs = {}
    for tag, prec, pert in (("fp32", 1, 0.0), ("fp32b", 1, 0.0), ("fp64", 0, 0.0), ("floor", 0, 1e-7)):
        k, _, L = bench.make_kernel(nx, nx, 5, prec, "leads", pert)
        if tag == "fp32":
            k.set_budget(True)
        for i in range(steps):
            k.reset_dg0("hice")
            k.reset_dg0("cice")
            k.storm_wind(L, 120.0 * i, st)
            if tag == "fp32":
                h0, a0 = k.mass()
            k.update(120.0)
            if tag == "fp32":
                h1, a1 = k.mass()
                b = k.budget(reset=True)
                eh = abs((h1 - h0) - (b["limiter_h"] - b["outflow_h"])) / h0
                ea = abs((a1 - a0) - (b["limiter_A"] - b["outflow_A"])) / a0
                _say(f"  step {i}: budget closure h {eh:.1e} A {ea:.1e}")
                assert eh <= 1e-13 and ea <= 1e-13, (i, eh, ea)
        runs[tag] = _fields_gpu(k)
        if tag == "fp32":
            u = runs[tag]["u"].reshape(2 * nx + 1, 2 * nx + 1)
            assert np.all(np.isfinite(u)) and np.abs(u).max() > 1e-3
            # the CG nodes on the edges between the land ring and the ocean are Dirichlet nodes
            lo, hi = 2, 2 * nx - 2
            assert np.all(u[lo, lo:hi + 1] == 0) and np.all(u[hi, lo:hi + 1] == 0)
            assert np.all(u[lo:hi + 1, lo] == 0) and np.all(u[lo:hi + 1, hi] == 0)
        del k
    for n in runs["fp32"]:
        assert np.array_equal(runs["fp32"][n], runs["fp32b"][n]), n
    e32 = {n: rel_l2(runs["fp32"][n], runs["fp64"][n]) for n in runs["fp64"]}
    efl = {n: rel_l2(runs["floor"][n], runs["fp64"][n]) for n in runs["fp64"]}
    _say("  fp32:  " + " ".join(f"{n}={e:.1e}" for n, e in e32.items()))
    _say("  floor: " + " ".join(f"{n}={e:.1e}" for n, e in efl.items()))
    for n in e32:
        assert e32[n] <= max(1e-5, 3.0 * efl[n]), (n, e32[n], efl[n])
