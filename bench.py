"""Benchmark of the mEVP dynamics step on one MI355X.

A step is one model timestep of the benchmark box as MEVPDynamics::update runs it
(MEVPDynamics.cpp:59-87): DG0 reset of h and A, the moving-cyclone wind evaluated on
the device, then VPCGDynamicsKernel::update (advection of A and h with limiters,
prepareIteration, 100 mEVP subcycles). Inputs are resident in HBM before timing.

Metric: Mcell-mEVP-iters/s = elements x subcycles / wall time / 1e6, FP64 path, on
the 512 x 512 CG2/DG2 box (dx = 1 km, dt = 120 s). Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def byte_model(nx, ny, cg=2, we=8):
    """Algorithmic HBM bytes of one subcycle: every value a pass needs, moved once.
    Two-pass form: element pass = stress read + written (3 S coefficients each), ice
    strength at the Q Gauss points, u and v at the nodes; node pass = stress read, per node
    u, v and the 6 frozen momentum coefficients read, u, v written. One-pass (fused) form:
    stress read + written, strength, per node u, v and the frozen coefficients read, u, v
    written. Returns (element, node, fused)."""
    S, Q = (8, 9) if cg == 2 else (3, 4)
    ne = nx * ny
    nn = (cg * nx + 1) * (cg * ny + 1)
    elem = we * (6 * S + Q) * ne + 8 * 2 * nn
    node = we * 3 * S * ne + 8 * 10 * nn
    fused = we * (6 * S + Q) * ne + 8 * 10 * nn
    return elem, node, fused


def survey_byte_model(nx, ny, cg=2, dg=6, we=8):
    """Algorithmic bytes of one mEVP subcycle of the reference's two-pass formulation (the
    roofline figure): element pass u, v at the nodes, h and A, stress read and written;
    node pass the stress gathered and 13 node fields read, u and v written:
    B = we (9 S + 2 D) N_e + 8 * 17 N_n (we = 8 for FP64 element data, 4 for FP32)."""
    S = 8 if cg == 2 else 3
    return we * (9 * S + 2 * dg) * nx * ny + 8 * 17 * (cg * nx + 1) * (cg * ny + 1)


def make_kernel(nx, ny, seed, prec, kind="plain", perturb=0.0, brittle=False, warp=0.0):
    """The benchmark box; warp > 0 shifts the vertices by warp * sin(y / 9 km) in x, which
    makes every element a non-affine quadrilateral (the general-mesh kernels)."""
    import numpy as np
    import nextsimdg_amd as nd
    from nextsimdg_amd import inputs

    s = inputs.benchmark_setup(nx, ny, seed, kind=kind)
    k = nd.BBMDynamicsKernel(2, 6) if brittle else nd.MEVPDynamicsKernel(2, 6)
    coords = s["coords"]
    if warp:
        coords = coords.copy()
        coords[:, 0] += warp * np.sin(coords[:, 1] / 9000.0)
    k.initialise(coords, False, s["mask"], nx, ny)
    h = s["h"] * (1.0 + perturb * np.random.default_rng(seed + 777).standard_normal(s["h"].size))
    k.setData("hice", h, 1)
    k.setData("cice", s["A"], 1)
    k.setData("uocean", s["uocean"], 1)
    k.setData("vocean", s["vocean"], 1)
    k.setData("ssh", 0 * s["h"], 1)
    if prec:
        k.set_precision(prec)
    return k, inputs.cyclone_offset(seed), inputs.L_BOX


def run_steps(k, off, L, t0, n, dt):
    for i in range(n):
        k.reset_dg0("hice")
        k.reset_dg0("cice")
        k.benchmark_wind(L, (t0 + i) * dt, off[0], off[1], 30.0)
        k.update(dt)


def measure(nx, ny, steps, warmup, dt=120.0, seed=0, prec=0, kind="plain"):
    k, off, L = make_kernel(nx, ny, seed, prec, kind)
    run_steps(k, off, L, 0, warmup, dt)
    k.synchronize()
    t = time.perf_counter()
    run_steps(k, off, L, warmup, steps, dt)
    k.synchronize()
    el = time.perf_counter() - t
    # per-kernel durations with HIP events on the library's stream
    k.set_profiling(True)
    run_steps(k, off, L, warmup + steps, 1, dt)
    k.synchronize()
    stats = {n: k.kernel_stats(n) for n in ("mevp_fused", "mevp_pair", "mevp_element", "mevp_node", "transport_op",
                                            "dg2cg", "cg2dg", "frozen_terms")}
    k.set_profiling(False)
    return el, stats


def geometry_byte_model(nx, ny, spherical=True):
    """Algorithmic HBM bytes of one one-pass subcycle on a general mesh (CG2): per element
    the stress read and written (3 S coefficients each), the strength at the Q Gauss points
    and the geometry table (the 6 bilinear map coefficients, on the sphere cos and sin of
    the latitude at the Q Gauss points, the upper triangle of the inverse mass matrix); per
    node u, v read and written and the 6 frozen values read."""
    S, Q = 8, 9
    ne, nn = nx * ny, (2 * nx + 1) * (2 * ny + 1)
    geo = 6 + (2 * Q if spherical else 0) + S * (S + 1) // 2
    return 8 * (6 * S + Q + geo) * ne + 80 * nn


def spherical_leg(nx, steps, warmup, dt=120.0, seed=0, traffic_path=None):
    """The mEVP step on a spherical longitude-latitude patch (per-element geometry with the
    metric terms streamed from HBM): throughput and the one-pass subcycle kernel against
    its byte model."""
    import nextsimdg_amd as nd
    from nextsimdg_amd import inputs

    s = inputs.benchmark_setup(nx, nx, seed)
    k = nd.MEVPDynamicsKernel(2, 6)
    k.initialise(inputs.sphere_patch_coords(nx, nx), True, s["mask"], nx, nx)
    for name, key in (("hice", "h"), ("cice", "A"), ("uocean", "uocean"), ("vocean", "vocean")):
        k.setData(name, s[key], 1)
    k.setData("ssh", 0 * s["h"], 1)
    off = inputs.cyclone_offset(seed)
    run_steps(k, off, inputs.L_BOX, 0, warmup, dt)
    k.synchronize()
    t = time.perf_counter()
    run_steps(k, off, inputs.L_BOX, warmup, steps, dt)
    k.synchronize()
    el = time.perf_counter() - t
    k.set_profiling(True)
    run_steps(k, off, inputs.L_BOX, warmup + steps, 1, dt)
    k.synchronize()
    stats = {n: k.kernel_stats(n) for n in ("mevp_fused", "strain", "transport_op", "cg2dg", "dg2cg",
                                            "frozen_terms")}
    k.set_profiling(False)
    alg = geometry_byte_model(nx, nx)
    out = {"workload": f"mEVP dynamics step on a spherical longitude-latitude patch {nx}x{nx} (1.2 x 0.4 rad "
                       f"around 63N), CG2/DG2, FP64, per-element geometry with metric terms, 100 subcycles",
           "value": round(nx * nx * 100 * steps / el / 1e6, 3), "unit": "Mcell-mEVP-iters/s",
           "steps": steps, "ms_per_step": round(el / steps * 1e3, 4), "kernels": {}}
    c, tot = stats["mevp_fused"]
    avg = tot / c
    out["kernels"]["mevp_fused_geom"] = {"avg_launch_ms": round(avg, 5), "algorithmic_bytes_per_launch": alg,
                                         "achieved": round(alg / (avg * 1e-3) / 1e9, 1), "unit": "GB/s",
                                         "hbm_frac": round(alg / (avg * 1e-3) / HBM_PEAK, 4),
                                         "traffic": _traffic(traffic_path, f"mevp_fused_geom_{nx}")}
    out["kernels_ms"] = {n: (round(v[1] / v[0], 5) if v[0] else None) for n, v in stats.items()}
    return out


def brittle_byte_model(nx, ny, cg=2, dg=6):
    """Algorithmic HBM bytes of one one-pass brittle substep: stress read + written (3 S
    coefficients each), damage read + written, h and A read (D coefficients each) per
    element; per node u, v read + written, the 7 terms fixed over the substeps (ocean
    velocity, dt / (rho H) and its ratio to the lumped mass, the wind and SSH forcing, A F_ocean)
    read, the averaged velocity read + written."""
    S = 8 if cg == 2 else 3
    ne, nn = nx * ny, (cg * nx + 1) * (cg * ny + 1)
    return 8 * (6 * S + 4 * dg) * ne + 8 * 15 * nn


def brittle_general_byte_model(nx, ny, cg=2, dg=6):
    """brittle_byte_model on a general (planar) mesh: plus the geometry table (6 bilinear map
    coefficients and the upper triangle of the inverse stress mass matrix) and the brittle
    table (h_e and the upper triangle of the inverse DG mass matrix) per element."""
    S = 8 if cg == 2 else 3
    return brittle_byte_model(nx, ny, cg, dg) + 8 * (6 + S * (S + 1) // 2 + 1 + dg * (dg + 1) // 2) * nx * ny


def brittle_leg(nx, steps, warmup, dt=120.0, seed=0, traffic_path=None, warp=0.0):
    """The BBM (brittle) dynamics step on the same box and forcing: advection, stress and
    damage transport, 100 substeps of dt / 100. Mcell-BBM-substeps/s and the one-pass
    substep kernel against its byte model. warp > 0: the same box with non-affine elements
    (the one-pass kernel from per-element geometry)."""
    k, off, L = make_kernel(nx, nx, seed, 0, brittle=True, warp=warp)
    run_steps(k, off, L, 0, warmup, dt)
    k.synchronize()
    t = time.perf_counter()
    run_steps(k, off, L, warmup, steps, dt)
    k.synchronize()
    el = time.perf_counter() - t
    k.set_profiling(True)
    run_steps(k, off, L, warmup + steps, 1, dt)
    k.synchronize()
    names = ("bbm_fused", "strain", "transport_op", "limit", "dg2cg", "cg2dg", "ssh_gradient")
    stats = {n: k.kernel_stats(n) for n in names}
    k.set_profiling(False)
    c, tot = stats["bbm_fused"]
    avg_ms = tot / c
    alg = brittle_general_byte_model(nx, nx) if warp else brittle_byte_model(nx, nx)
    mesh = f"benchmark box {nx}x{nx} with non-affine elements (x += {warp:g} m sin(y / 9 km))" if warp else \
        f"benchmark box {nx}x{nx}"
    return {
        "workload": f"BBM brittle dynamics step, {mesh}, CG2 velocity, DG2 advection, "
                    f"8-dof DG stress, FP64, dt=120 s, 100 substeps of 1.2 s, moving cyclone",
        "value": round(nx * nx * 100 * steps / el / 1e6, 3),
        "unit": "Mcell-BBM-substeps/s",
        "steps": steps,
        "ms_per_step": round(el / steps * 1e3, 4),
        "kernel": "bbm_fused_geom" if warp else "bbm_fused",
        "avg_launch_ms": round(avg_ms, 5),
        "algorithmic_bytes_per_launch": alg,
        "achieved": round(alg / (avg_ms * 1e-3) / 1e9, 1),
        "unit_achieved": "GB/s",
        "hbm_frac": round(alg / (avg_ms * 1e-3) / HBM_PEAK, 4),
        "traffic": _traffic(traffic_path, f"bbm_fused_geom_{nx}" if warp else f"bbm_fused_{nx}"),
        "kernels_ms": {n: (round(v[1] / v[0], 5) if v[0] else None) for n, v in stats.items()},
    }


def _traffic(path, key):
    """Measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)."""
    if not path or not os.path.exists(path):
        return None
    try:
        return json.load(open(path)).get(key)
    except (ValueError, OSError):
        return None


def fp32_variant(nx, steps, warmup, err_steps, seed=0):
    """The FP32 variant on heterogeneous ice (h = 0.3 + LogNormal(0, 0.3), A ~ Beta(8, 1)):
    throughput next to the FP64 path on the same inputs, and the per-field relative L2
    error of FP32 against FP64 after 1, 10 and err_steps model steps, each beside the
    conditioning floor: the same FP64 run with h perturbed by a relative 1e-7 (one FP32
    rounding)."""
    import numpy as np

    kind = "heterogeneous"
    t64, _ = measure(nx, nx, steps, warmup, seed=seed, prec=0, kind=kind)
    t32, st32 = measure(nx, nx, steps, warmup, seed=seed, prec=1, kind=kind)
    horizons = sorted({h for h in (1, 10, err_steps) if 1 <= h <= err_steps})
    fields = {}
    for prec, pert in ((0, 0.0), (1, 0.0), ("floor", 1e-7)):
        k, off, L = make_kernel(nx, nx, seed, 0 if prec == "floor" else prec, kind, pert)
        done = 0
        for h in horizons:
            run_steps(k, off, L, done, h - done, 120.0)
            done = h
            k.synchronize()
            f = {n: k.get_cg(n) for n in ("u", "v")}
            f.update({n: k.getDGData(n, 6) for n in ("hice", "cice")})
            f.update({n: k.getDGData(n, 8) for n in ("s11", "s12", "s22")})
            fields[(prec, h)] = f
        del k
    def rel(a, h=err_steps):
        ref = fields[(0, h)]
        return {n: float(f"{np.linalg.norm(fields[(a, h)][n] - ref[n]) / np.linalg.norm(ref[n]):.3e}") for n in ref}
    iters = nx * nx * 100 * steps
    return {
        "workload": f"mEVP dynamics step, {nx}x{nx} CG2/DG2, heterogeneous ice, FP32 stress/strength storage "
                    "and element arithmetic (compensated relaxed-stress sum, stochastic rounding of the "
                    "stored stress), FP64 velocity update",
        "value": round(iters / t32 / 1e6, 3),
        "fp64_value": round(iters / t64 / 1e6, 3),
        "speedup_vs_fp64": round(t64 / t32, 3),
        "kernels_ms": {n: (round(t / c, 5) if c else None) for n, (c, t) in st32.items()},
        "rel_l2_vs_fp64": rel(1),
        "fp64_rel_l2_under_1e-7_input_perturbation": rel("floor"),
        "error_after_steps": err_steps,
        "rel_l2_by_steps": {str(h): {"fp32_vs_fp64": rel(1, h), "fp64_floor": rel("floor", h)} for h in horizons},
    }


WEIGHTS = os.path.join(ROOT, "nextsimdg_amd", "data", "warm_start_256.npy")


def learned_vs_plain(nx, kinds=("heterogeneous", "leads", "plain"), seed=5000, spinup=2, check_every=100):
    """Speed at equal momentum residual: plain FP64 mEVP (100 subcycles from the cold start)
    against the FP32 path started by the trained network with the residual gate, iterating
    until r <= r(plain). Both from the same saved state at the start of a timestep. The
    residual loop starts from the residual the gate computed and checks first after the
    plain path's own 100 subcycles, then in chunks fitted to the rate of decrease."""
    from nextsimdg_amd import learned as L

    if not os.path.exists(WEIGHTS):
        return {"error": f"no trained weights at {os.path.relpath(WEIGHTS, ROOT)}"}
    net = L.WarmStartNet(0)
    net.set_params(__import__("numpy").load(WEIGHTS))
    out = {"workload": f"{nx}x{nx} CG2/DG2 box, one timestep after {spinup} plain steps, seeded cyclone "
                       f"(track {seed}+); network trained at 256x256 (include/nsdg_cnn.h)",
           "check_every": check_every, "cases": []}
    for i, kind in enumerate(kinds):
        k, off = L.new_track(nx, seed + i, kind)
        for st in range(spinup):
            L.plain_step(k, off, st)
        L.begin_step(k, off, spinup)
        blob = k.save_state()
        r_cold = k.momentum_residual()[0]
        k.subcycles(100)  # warm
        k.load_state(blob)
        k.synchronize()
        t = time.perf_counter()
        k.subcycles(100)
        k.synchronize()
        t_plain = time.perf_counter() - t
        r_plain = k.momentum_residual()[0]
        k.load_state(blob)
        k.set_precision(1)
        t = time.perf_counter()
        rp, rc, used = L.warm_start(k, net)
        done, r = L.subcycles_to_residual(k, r_plain, 1000, check_every, rp if used else rc)
        t_learned = time.perf_counter() - t
        out["cases"].append({"ice": kind, "r_cold": float(f"{r_cold:.5e}"), "r_plain_100": float(f"{r_plain:.5e}"),
                             "r_predicted": float(f"{rp:.5e}"), "prediction_kept": used,
                             "fp32_subcycles_to_r_plain": done, "r_reached": float(f"{r:.5e}"),
                             "t_plain_ms": round(t_plain * 1e3, 3), "t_learned_ms": round(t_learned * 1e3, 3),
                             "speedup": round(t_plain / t_learned, 3)})
        del k
    out["speedup_min"] = min(c["speedup"] for c in out["cases"])
    return out


def cpu_info():
    """CPU model, logical CPUs of the host, and the CPUs this process may run on."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    avail = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "cpus_available": avail}


def _oracle_threads():
    """All the cores this run may use: OMP_NUM_THREADS when set, else the CPUs available."""
    env = os.environ.get("OMP_NUM_THREADS")
    return int(env) if env and env.isdigit() else cpu_info()["cpus_available"]


def _oracle_subcycles(nx, ny, seed, target_s):
    """mEVP subcycles of the config-2 box on the FP64 oracle for about target_s seconds."""
    from nextsimdg_amd import inputs
    from oracle import Oracle, OracleDyn

    s = inputs.benchmark_setup(nx, ny, seed)
    o = OracleDyn(2, 6)
    o.initialise(nx, ny, s["coords"], s["mask"])
    uw, vw = Oracle().benchmark_wind(nx, ny, inputs.L_BOX, 0.0, *inputs.cyclone_offset(seed), 30.0)
    for name, v in (("hice", s["h"]), ("cice", s["A"]), ("uocean", s["uocean"]), ("vocean", s["vocean"]),
                    ("uwind", uw), ("vwind", vw), ("ssh", 0 * uw)):
        o.set_field(name, v, 1)
    o.advection(120.0)
    o.prepare_iteration()
    o.begin_iteration(120.0)
    o.run_pass("subcycle", 1)  # warm
    n, t = 0, time.perf_counter()
    while time.perf_counter() - t < target_s or n < 2:
        o.run_pass("subcycle", 1)
        n += 1
    return n, time.perf_counter() - t


def cpu_baseline(nx, ny, seed=0, target_s=10.0):
    """The FP64 oracle (OpenMP C++ restatement of the reference) on the host cores: mEVP
    subcycles of the same box for about target_s seconds, on all cores and on one."""
    import oracle

    threads = _oracle_threads()
    oracle.set_threads(threads)
    n, el = _oracle_subcycles(nx, ny, seed, target_s)
    oracle.set_threads(1)
    n1, el1 = _oracle_subcycles(nx, ny, seed, target_s)
    oracle.set_threads(threads)
    return dict({
        "value": round(nx * ny * n / el / 1e6, 4),
        "unit": "Mcell-mEVP-iters/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{n} mEVP subcycles of the {nx}x{ny} CG2/DG2 box on the FP64 oracle (advection excluded), "
                  f"{el:.1f} s on {threads} threads",
        "one_thread": {"value": round(nx * ny * n1 / el1 / 1e6, 4), "unit": "Mcell-mEVP-iters/s", "cores": 1,
                       "sample": f"{n1} subcycles, {el1:.1f} s"},
    }, **cpu_info())


def _config1_steps(k, setf, getf, nx, winds, h, A, t_limit=None):
    """Module-level steps of config 1 (DG0 h and A, forcing, update); stops early after
    t_limit seconds. Returns (steps run, seconds)."""
    import numpy as np

    zero = np.zeros(nx * nx)
    t = time.perf_counter()
    n = 0
    for uw, vw in winds:
        setf("hice", h, 1)
        setf("cice", A, 1)
        setf("uwind", uw, 1)
        setf("vwind", vw, 1)
        setf("ssh", zero, 1)
        k.update(120.0)
        h, A = getf("hice"), getf("cice")
        n += 1
        if t_limit and time.perf_counter() - t > t_limit:
            break
    if hasattr(k, "synchronize"):
        k.synchronize()
    return n, time.perf_counter() - t


def config1(nx=128, steps=720, cpu_window_s=10.0):
    """Config 1, the plain method on the reference's CPU path: 128 x 128 box (dx = 4 km),
    CG1 velocity, DG1 advection, 3-dof stress, dt = 120 s, 100 subcycles, 24 h (720 steps)
    of the moving cyclone. The FP64 oracle runs a window of the 720 steps (about
    cpu_window_s seconds) on all cores and on one; the GPU runs all 720 steps. Module-level
    steps both sides (host round trips of h, A and the forcing included)."""
    import numpy as np
    import oracle
    import nextsimdg_amd as nd
    from nextsimdg_amd import inputs

    s = inputs.benchmark_setup(nx, nx, 0)
    off = inputs.cyclone_offset(0)
    ora = oracle.Oracle()
    winds = [ora.benchmark_wind(nx, nx, inputs.L_BOX, i * 120.0, off[0], off[1], 30.0) for i in range(steps)]
    out = {"workload": f"config 1: {nx}x{nx} box, dx=4 km, CG1/DG1, 3-dof stress, FP64, dt=120 s, 100 subcycles, "
                       f"moving cyclone, {steps} steps (24 h), module-level steps with host I/O",
           "unit": "Mcell-mEVP-iters/s"}
    g = nd.MEVPDynamicsKernel(1, 3)
    g.initialise(s["coords"], False, s["mask"], nx, nx)
    for name in ("uocean", "vocean"):
        g.setData(name, s[name], 1)
    _config1_steps(g, g.setData, g.getDG0Data, nx, winds[:5], s["h"], s["A"])  # warm
    n, el = _config1_steps(g, g.setData, g.getDG0Data, nx, winds, s["h"], s["A"])
    out["gpu"] = {"steps": n, "seconds": round(el, 3), "value": round(nx * nx * 100 * n / el / 1e6, 3)}
    threads = _oracle_threads()
    for tag, th, lim in (("cpu_all_cores", threads, cpu_window_s), ("cpu_one_thread", 1, cpu_window_s)):
        oracle.set_threads(th)
        o = oracle.OracleDyn(1, 3)
        o.initialise(nx, nx, s["coords"], s["mask"])
        for name in ("uocean", "vocean"):
            o.set_field(name, s[name], 1)
        n, el = _config1_steps(o, o.set_field, lambda name: o.get_field(name, 1), nx, winds, s["h"], s["A"], lim)
        out[tag] = {"threads": th, "steps": n, "seconds": round(el, 3),
                    "value": round(nx * nx * 100 * n / el / 1e6, 4),
                    "projected_24h_seconds": round(el / n * steps, 1)}
    oracle.set_threads(threads)
    out["gpu_vs_cpu_all_cores"] = round(out["gpu"]["value"] / out["cpu_all_cores"]["value"], 1)
    out.update(cpu_info())
    del np
    return out


def budget_leg(nx, steps, seed=0):
    """The advection mass budget of the benchmark box over `steps` module-level steps:
    change of the ice volume and area (Tools::MeanValue) against the Dirichlet outflow and
    the limiters' change, relative to the starting mass; closure = |change - (limiter -
    outflow)| / mass, the worst step."""
    k, off, L = make_kernel(nx, nx, seed, 0)
    k.set_budget(True)
    m0 = k.mass()
    worst = [0.0, 0.0]
    tot = dict(outflow_h=0.0, outflow_A=0.0, limiter_h=0.0, limiter_A=0.0)
    for i in range(steps):
        k.reset_dg0("hice")
        k.reset_dg0("cice")
        k.benchmark_wind(L, i * 120.0, off[0], off[1], 30.0)
        a = k.mass()
        k.update(120.0)
        b = k.mass()
        bud = k.budget(reset=True)
        for f, key in ((0, "h"), (1, "A")):
            err = abs((b[f] - a[f]) - (bud["limiter_" + key] - bud["outflow_" + key])) / a[f]
            worst[f] = max(worst[f], err)
        for key in tot:
            tot[key] += bud[key]
    m1 = k.mass()
    out = {"workload": f"benchmark box {nx}x{nx}, CG2/DG2 FP64, {steps} module-level steps of the moving cyclone"}
    for f, key, name in ((0, "h", "volume"), (1, "A", "area")):
        out[name] = {"rel_change": float(f"{(m1[f] - m0[f]) / m0[f]:.6e}"),
                     "rel_outflow": float(f"{tot['outflow_' + key] / m0[f]:.6e}"),
                     "rel_limiter": float(f"{tot['limiter_' + key] / m0[f]:.6e}"),
                     "closure_worst_step": float(f"{worst[f]:.3e}")}
    return out


def config5_leg(nx, seed=5, spinup=2, check_every=100):
    """Config 5: 2048 x 2048 box, three seeded storms (10-25 m/s), heterogeneous ice with
    open-water leads. One timestep after `spinup` plain steps, from one saved state: plain
    FP64 mEVP (100 subcycles) against FP32 with the gated learned start iterating until the
    momentum residual reaches plain's: subcycles to the residual and wall time."""
    import numpy as np
    from nextsimdg_amd import inputs
    from nextsimdg_amd import learned as Lrn

    st = inputs.storms(seed)
    k, _, L = make_kernel(nx, nx, seed, 0, "leads")
    for i in range(spinup):
        k.reset_dg0("hice")
        k.reset_dg0("cice")
        k.storm_wind(L, i * 120.0, st)
        k.update(120.0)
    k.reset_dg0("hice")
    k.reset_dg0("cice")
    k.storm_wind(L, spinup * 120.0, st)
    k.advectionAndLimits(120.0)
    k.prepareIteration()
    k.begin_iteration(120.0)
    blob = k.save_state()
    r_cold = k.momentum_residual()[0]
    k.subcycles(100)  # warm
    k.load_state(blob)
    k.synchronize()
    t = time.perf_counter()
    k.subcycles(100)
    k.synchronize()
    t_plain = time.perf_counter() - t
    r_plain = k.momentum_residual()[0]
    out = {"workload": f"config 5: {nx}x{nx} CG2/DG2 box, 3 seeded storms (peaks "
                       f"{', '.join(f'{v / np.e:.1f}' for v in st[:, 2])} m/s), heterogeneous ice with open-water "
                       f"leads, one timestep after {spinup} plain steps",
           "r_cold": float(f"{r_cold:.5e}"), "r_plain_100": float(f"{r_plain:.5e}"),
           "t_plain_ms": round(t_plain * 1e3, 3)}
    net = None
    if os.path.exists(WEIGHTS):
        net = Lrn.WarmStartNet(0)
        net.set_params(np.load(WEIGHTS))
    for tag, learned in (("fp32", False), ("fp32_learned", True)):
        if learned and net is None:
            continue
        k.load_state(blob)
        k.set_precision(1)
        k.synchronize()
        t = time.perf_counter()
        rec = {}
        r_start = r_cold  # the saved start state's residual, computed above
        if learned:
            rp, rc, used = Lrn.warm_start(k, net)
            rec.update(r_predicted=float(f"{rp:.5e}"), prediction_kept=used)
            r_start = rp if used else rc
        done, r = Lrn.subcycles_to_residual(k, r_plain, 1000, check_every, r_start)
        el = time.perf_counter() - t
        rec.update(subcycles_to_r_plain=done, r_reached=float(f"{r:.5e}"), t_ms=round(el * 1e3, 3),
                   speedup_vs_plain_fp64=round(t_plain / el, 3))
        out[tag] = rec
        k.set_precision(0)
    return out


MFMA_BF16_PEAK = 2.5e15  # dense BF16 FLOP/s of MI355X


def network_leg(nx, reps=10):
    """The warm-start network's forward pass at nx x nx (one image, 32 BF16 channels): the
    3 x 3 convolution kernels' times (HIP events), their dense BF16 MFMA rate against the
    2.5 PFLOP/s peak (2 x 32 x 9 x 32 FLOP per pixel and hidden layer; the output layer
    computes one 16-channel MFMA tile, 2 x 32 x 9 x 16) and their HBM rate on the
    activations they read and write (64 B in per pixel; 64 B out, 32 B for the FP32 output
    layer)."""
    import numpy as np
    from nextsimdg_amd import learned as L

    net = L.WarmStartNet(0)
    if os.path.exists(WEIGHTS):
        net.set_params(np.load(WEIGHTS))
    rng = np.random.default_rng(0)
    xf = rng.normal(0.0, 1.0, (1, nx, nx, L.CHANNELS)).astype(np.float32)
    x = L.DeviceArray.from_numpy((xf.view(np.uint32) >> 16).astype(np.uint16))
    y = L.DeviceArray((1, nx, nx, L.OUTPUTS), np.float32)
    net.forward(x, y)
    net.set_profiling(True)
    for _ in range(reps):
        net.forward(x, y)
    npix = nx * nx
    out = {"workload": f"warm-start network forward, {nx}x{nx} element grid, 4 layers of 3x3 convolutions, "
                       "32 channels, BF16 on MFMA (v_mfma_f32_16x16x32_bf16), FP32 accumulation",
           "kernels": {}}
    for name, layers, cout, out_bytes in (("cnn_conv", 3, 32, 64), ("cnn_conv_out", 1, 16, 32)):
        c, t = net.kernel_stats(name)
        if not c:
            continue
        ms = t / c
        flop = 2.0 * L.CHANNELS * 9 * cout * npix
        out["kernels"][name] = {"layers_per_forward": layers, "avg_launch_ms": round(ms, 5),
                                "tflops": round(flop / (ms * 1e-3) / 1e12, 1),
                                "mfma_frac": round(flop / (ms * 1e-3) / MFMA_BF16_PEAK, 4),
                                "hbm_gbs": round((64 + out_bytes) * npix / (ms * 1e-3) / 1e9, 1),
                                "hbm_frac": round((64 + out_bytes) * npix / (ms * 1e-3) / HBM_PEAK, 4)}
    c1, t1 = net.kernel_stats("cnn_conv")
    c2, t2 = net.kernel_stats("cnn_conv_out")
    out["forward_ms"] = round((t1 + t2) / reps, 4) if c1 and c2 else None
    return out


def calibrate_copy(mib=1024, reps=3):
    """A plain device-to-device copy of known size, for scaling the FETCH_SIZE / WRITE_SIZE
    counters of rocprofv3 --pmc runs to bytes."""
    import numpy as np
    from nextsimdg_amd.learned import DeviceArray

    n = mib * 1024 * 1024 // 8
    a = DeviceArray.from_numpy(np.ones((1, n)))
    b = DeviceArray((1, n), np.float64)
    for _ in range(reps):
        a.gather([0], b)
    print(json.dumps({"copy_bytes_each_way": n * 8, "reps": reps}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nx", type=int, default=512)
    ap.add_argument("--roofline-nx", type=int, default=1024,
                    help="grid of the roofline measurement (quoted at >= 1024^2, where the working set "
                         "exceeds the 256 MB last-level cache)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fp32-nx", type=int, default=1024, help="FP32 variant size (0 skips it)")
    ap.add_argument("--fp32-error-steps", type=int, default=100)
    ap.add_argument("--learned-nx", type=int, default=1024, help="learned vs plain size (0 skips it)")
    ap.add_argument("--bbm-nx", type=int, default=512, help="brittle (BBM) step size (0 skips it)")
    ap.add_argument("--bbm-general-nx", type=int, default=512,
                    help="brittle step on a non-affine mesh of this size (0 skips it)")
    ap.add_argument("--spherical-nx", type=int, default=512, help="spherical patch size (0 skips it)")
    ap.add_argument("--budget-steps", type=int, default=10, help="mass budget steps at --nx (0 skips it)")
    ap.add_argument("--config5-nx", type=int, default=2048, help="config 5 (storms, leads) size (0 skips it)")
    ap.add_argument("--network-nx", type=int, default=1024, help="network forward size (0 skips it)")
    ap.add_argument("--no-config1", action="store_true", help="skip config 1 (128^2 CG1/DG1, CPU and GPU)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="measured HBM bytes per launch (from rocprofv3 --pmc) to report as traffic")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a 1 GiB device copy (for counter calibration) and exit")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the path is built and measured on one "
              "GPU's share of the grid; sharding across GPUs is a design, not a measured path", file=sys.stderr)
        sys.exit(2)
    nx = ny = args.nx
    log(f"bench: {nx}x{ny} CG2/DG2 FP64, warmup {args.warmup}, steps {args.steps}")
    el, stats = measure(nx, ny, args.steps, args.warmup)
    iters = nx * ny * 100 * args.steps
    value = iters / el / 1e6
    eb, nb, fb = byte_model(nx, ny)
    ms = {n: (c, t / c if c else None) for n, (c, t) in stats.items()}
    # roofline of the dominant kernel at the roofline grid: one profiled step's HIP events
    rnx = args.roofline_nx or nx
    if rnx != nx:
        log(f"bench: roofline step at {rnx}x{rnx} ...")
        _, rstats = measure(rnx, rnx, 1, 1)
    else:
        rstats = stats
    fused_of = dict(zip(("mevp_element", "mevp_node", "mevp_fused"), byte_model(rnx, rnx)))
    fused_of["mevp_pair"] = 2 * fused_of["mevp_fused"]
    # subcycles per launch of each kernel: the pair kernel runs two (reading and writing the
    # state once per two subcycles is what it saves)
    per_launch = {"mevp_element": 1, "mevp_node": 1, "mevp_fused": 1, "mevp_pair": 2}
    dom = max(per_launch, key=lambda n: rstats[n][1])
    avg_ms = rstats[dom][1] / rstats[dom][0]
    alg = per_launch[dom] * survey_byte_model(rnx, rnx)
    achieved = alg / (avg_ms * 1e-3)
    traffic = _traffic(args.traffic, f"{dom}_{rnx}")
    # % of HBM peak on the two-pass byte model of the reference's formulation (element pass:
    # u, v, h, A, sigma r+w; node pass: sigma and 13 node fields r, u, v w), 8 (9 S + 2 D) per
    # element plus 8 * 17 per node, FP64: whole-step throughput measured against that model.
    ne, nn = nx * ny, (2 * nx + 1) * (2 * ny + 1)
    ref_bytes = 8 * (9 * 8 + 2 * 6) * ne + 8 * 17 * nn
    out = {
        "metric": "Mcell-mEVP-iters/s, % HBM peak; learned vs plain speedup at equal residual",
        "value": round(value, 3),
        "unit": "Mcell-mEVP-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"mEVP dynamics step, benchmark box {nx}x{ny}, CG2 velocity, DG2 advection, "
                               f"8-dof DG stress, FP64, dt=120 s, 100 subcycles, moving cyclone",
                   "nx": nx, "ny": ny, "subcycles": 100, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": dom, "grid": f"{rnx}x{rnx}", "achieved": round(achieved / 1e9, 1),
                     "peak": 8000.0, "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": alg,
                     "byte_model": "8 (9 S + 2 D) N_e + 8 * 17 N_n per subcycle x subcycles per launch "
                                   f"({per_launch[dom]})",
                     "avg_launch_ms": round(avg_ms, 5), "launches": rstats[dom][0],
                     # achieved / frac are effective rates on the reference formulation's bytes:
                     # the two-subcycle kernel reads and writes the state once per two
                     # subcycles, so they can pass the peak; the physical HBM rate is counter_rate
                     "note": "achieved and frac: the two-pass byte model of the reference formulation over the "
                             "launch time (an effective rate; above 1 when the kernel needs fewer bytes than the "
                             "model, as the two-subcycle kernel does); counter_rate and counter_frac: the HBM "
                             "bytes the counters measure per launch over the same time",
                     # the one-pass kernels' own compulsory bytes (stress r+w, strength, 10 node
                     # values): a tighter model than the two-pass one above
                     "achieved_fused_model": round(fused_of[dom] / (avg_ms * 1e-3) / 1e9, 1),
                     "frac_fused_model": round(fused_of[dom] / (avg_ms * 1e-3) / HBM_PEAK, 4),
                     # HBM bytes the counters measured per launch, over the launch time
                     "counter_rate": round(traffic / (avg_ms * 1e-3) / 1e9, 1) if traffic else None,
                     "counter_frac": round(traffic / (avg_ms * 1e-3) / HBM_PEAK, 4) if traffic else None},
        "kernels_ms": {n: (round(v[1], 5) if v[1] else None) for n, v in ms.items()},
        "hbm_peak_frac": round(value * 1e6 / ne * ref_bytes / HBM_PEAK, 4),
        "hbm_bytes_per_subcycle_reference_model": ref_bytes,
        "subcycle_ms": round(sum(stats[n][1] for n in ("mevp_fused", "mevp_pair", "mevp_element", "mevp_node"))
                             / (max(stats["mevp_fused"][0], stats["mevp_element"][0]) + 2 * stats["mevp_pair"][0]), 5),
    }
    if args.fp32_nx:
        log(f"bench: FP32 variant {args.fp32_nx}^2 ...")
        out["fp32"] = fp32_variant(args.fp32_nx, max(2, args.steps // 2), 1, args.fp32_error_steps)
    if args.learned_nx:
        log(f"bench: learned warm start vs plain at {args.learned_nx}^2 ...")
        out["learned"] = learned_vs_plain(args.learned_nx)
        out["learned_vs_plain_speedup_at_equal_residual"] = out["learned"].get("speedup_min")
    if args.bbm_nx:
        log(f"bench: BBM brittle step at {args.bbm_nx}^2 ...")
        out["bbm"] = brittle_leg(args.bbm_nx, max(2, args.steps // 2), 1, traffic_path=args.traffic)
    if args.bbm_general_nx:
        log(f"bench: BBM brittle step on a non-affine {args.bbm_general_nx}^2 mesh ...")
        out["bbm_general_mesh"] = brittle_leg(args.bbm_general_nx, max(2, args.steps // 2), 1,
                                              traffic_path=args.traffic, warp=50.0)
    if args.spherical_nx:
        log(f"bench: spherical patch at {args.spherical_nx}^2 ...")
        out["spherical"] = spherical_leg(args.spherical_nx, max(2, args.steps // 2), 1, traffic_path=args.traffic)
    if args.budget_steps:
        log(f"bench: mass budget at {nx}^2 ...")
        out["budget"] = budget_leg(nx, args.budget_steps)
    if args.config5_nx:
        log(f"bench: config 5 (storms, leads) at {args.config5_nx}^2 ...")
        out["config5"] = config5_leg(args.config5_nx)
    if args.network_nx:
        log(f"bench: network forward at {args.network_nx}^2 ...")
        out["network"] = network_leg(args.network_nx)
    if not args.no_cpu_baseline:
        log("bench: CPU baseline (oracle) ...")
        try:
            out["cpu_baseline"] = cpu_baseline(nx, ny)
        except Exception as e:  # noqa: BLE001 - the baseline is reported, never required
            out["cpu_baseline"] = {"value": None, "error": str(e)}
        if not args.no_config1:
            log("bench: config 1 (128^2 CG1/DG1, oracle and GPU) ...")
            try:
                out["config1"] = config1()
            except Exception as e:  # noqa: BLE001
                out["config1"] = {"error": str(e)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
