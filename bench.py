"""Throughput of the DVM time step on one MI355X.

Headline workload (the 2-D lid-driven cavity microflow): 64x64 cells, 25^3
velocity nodes on +-5, hard spheres, Kn = 0.1, four diffuse walls at T_w = 1,
lid at 0.3 along x, uniform equilibrium start, 20 000 collision samples per
cell per step. One step = CDF build + VR collision sampling + wall ghosts +
upwind sweep + moments, with the distributions resident in HBM before timing
starts. It runs with FP32 collision arithmetic (the value) and in plain FP64.

Also measured (unless --no-relaxation / --no-batched / --no-serving):
  relaxation_c1  the space-homogeneous relaxation (1 cell, 21^3, 100 000 samples
            per step, 200 steps, FP64) with the L1 distance from equilibrium
            every step, beside the oracle on one core
  batched   4096 space-homogeneous instances, grid edge 15..41 odd, mixtures of
            1-3 Maxwellians, 50 samples per node, stepped by one replayed HIP
            graph per step with autotuned collision tiles
  serving   Poisson arrivals of solve requests (70 % homogeneous, 30 % 1-D
            shocks) served by continuous batching; p50 / p99 latency

Prints one JSON line: cell-node updates/s (cells x velocity nodes x steps /
wall time of the timed steps), per-phase device times from HIP events on the
library's stream, the roofline of the dominant kernel (the collision kernel)
and of the sweep, and the CPU oracle's rate on a bounded sample of the same
per-cell work.
"""
import argparse
import ctypes as C
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per cell-node of one launch:
#   sweep     read f once, write the convected f once (FP64)
#   collide   FP64: read f, the cached equilibrium (the deviation signs) and
#             both CDFs once, write f + change once; FP32 (fast path): read f,
#             both CDFs and the deviation sign bit once, write the corrected
#             f + change once
SWEEP_BYTES_PER_NODE = 2 * 8
COLLIDE_BYTES_PER_NODE = {"fp64": 5 * 8, "fp32": 4 * 8 + 1 / 8}
TRAFFIC_TABLE = os.path.join(ROOT, "profiles", "hbm_traffic.txt")
VALU_TABLE = os.path.join(ROOT, "profiles", "valu_per_launch.txt")
SIMDS = 1024        # MI355X: 256 CUs x 4 SIMDs
CLOCK_GHZ = 2.4     # MI355X peak engine clock
METRIC = "cell-node updates/s, FP32-collision vs plain FP64 at 1e-5; p50/p99 solve latency"


def hip_sync():
    hip = C.CDLL("libamdhip64.so")
    rc = hip.hipDeviceSynchronize()
    if rc != 0:
        raise RuntimeError(f"hipDeviceSynchronize failed: {rc}")


def copy_calibration(nbytes=512 << 20):
    """One device-to-device copy of a known size (read nbytes, write nbytes),
    so that the HBM byte counters of a profile can be scaled against it."""
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3)  # hipMemcpyDeviceToDevice
    hip.hipDeviceSynchronize()
    hip.hipFree(a)
    hip.hipFree(b)


def copy_bandwidth(nbytes=1 << 30, reps=10):
    """HBM bandwidth of a plain device-to-device copy on this GPU (hipMemcpyAsync
    of 1 GiB, timed with HIP events): 2 x bytes / time, in GB/s. The practical
    ceiling for a read-once / write-once kernel beside the 8 TB/s nominal."""
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
    hip.hipEventSynchronize.argtypes = [C.c_void_p]
    hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
    hip.hipEventDestroy.argtypes = [C.c_void_p]
    a, b, e0, e1 = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    try:
        assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
        assert hip.hipEventCreate(C.byref(e0)) == 0 and hip.hipEventCreate(C.byref(e1)) == 0
        assert hip.hipMemcpyAsync(b, a, nbytes, 3, None) == 0  # warm-up, hipMemcpyDeviceToDevice
        assert hip.hipEventRecord(e0, None) == 0
        for _ in range(reps):
            assert hip.hipMemcpyAsync(b, a, nbytes, 3, None) == 0
        assert hip.hipEventRecord(e1, None) == 0
        assert hip.hipEventSynchronize(e1) == 0
        ms = C.c_float()
        assert hip.hipEventElapsedTime(C.byref(ms), e0, e1) == 0
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    return 2.0 * nbytes * reps / (ms.value * 1e-3) / 1e9


def run_gpu(pkg, cfg, f, steps, warmup, final=None):
    """Times `steps` fused steps after `warmup` ones; returns wall seconds, mean
    per-phase device ms (HIP events on the library's stream) and the moments.
    With a dict `final`, also stores the final f and output moments in it
    (after the timed region)."""
    s = pkg.Solver(cfg)
    s.initialize(f)
    for n in range(1, warmup + 1):
        s.step(n)
    s.timing_reset()
    hip_sync()
    t0 = time.perf_counter()
    for n in range(warmup + 1, warmup + steps + 1):
        s.step(n)
    hip_sync()
    t1 = time.perf_counter()
    phase, calls = s.timing_read()
    props = s.calculate_physical_properties()
    if final is not None:
        final["f"] = s.download_f()
        final["ext"] = s.output_moments()
    s.close()
    return (t1 - t0), {k: v / max(calls, 1) for k, v in phase.items()}, props


def output_moment_error(e32, e64, f32, f64):
    """Max over cells of the FP32 error of the heat flux, the viscous stress
    (each relative to the largest magnitude of that field over the cavity, as
    both vanish at equilibrium) and the entropy (relative per cell), and the
    relative L1 difference of the distributions, sum |f32 - f64| / sum |f64|."""
    q32, q64 = e32[:, 0:3], e64[:, 0:3]
    t32, t64 = e32[:, 3:9], e64[:, 3:9]
    return {"heat_flux": float(np.max(np.abs(q32 - q64)) / max(np.max(np.abs(q64)), 1e-300)),
            "viscous_stress": float(np.max(np.abs(t32 - t64)) / max(np.max(np.abs(t64)), 1e-300)),
            "entropy": float(np.max(np.abs(e32[:, 9] - e64[:, 9]) / np.maximum(np.abs(e64[:, 9]), 1e-300))),
            "f_l1": float(np.sum(np.abs(f32 - f64)) / np.sum(np.abs(f64)))}


_NATIVE = {}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def native_oracle_solver(pkg):
    """A factory of oracle contexts from oracle/dvm_oracle.c compiled on this
    host with gcc -O3 -march=native (-ffp-contract=off keeps its arithmetic
    the reference's; the library travels prebuilt, so the native build is
    made here, into a temporary folder). Falls back to the prebuilt -O2
    oracle if gcc fails. Returns (factory, compile flags)."""
    if "make" not in _NATIVE:
        import ctypes as C
        import subprocess
        import tempfile
        src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oracle", "dvm_oracle.c")
        out = os.path.join(tempfile.mkdtemp(prefix="dvm_oracle_"), "liboracle_native.so")
        flags = ["-O3", "-march=native", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math"]
        r = subprocess.run(["gcc"] + flags + ["-shared", "-o", out, src, "-lm"], capture_output=True, text=True)
        dvm = importlib.import_module(pkg.__name__ + ".dvm")
        if r.returncode == 0:
            lib = pkg.capi.declare(C.CDLL(out), "orc")
            _NATIVE["make"] = (lambda cfg: dvm.Solver(cfg, lib=lib, prefix="orc")), " ".join(flags)
        else:
            from tests._oracle import oracle_solver
            _NATIVE["make"] = oracle_solver, "-O2 (prebuilt; the native build failed)"
    return _NATIVE["make"]


def cpu_baseline(pkg, cfg_full, steps=12):
    """The oracle (serial C, one core) on a 16x16-cell cavity with the same
    velocity grid and samples per cell; rate in cell-node updates/s."""
    make, flags = native_oracle_solver(pkg)
    cfg, f = pkg.lid_driven_cavity(nx=16, ny=16, n_v=cfg_full.nvx, vmax=float(cfg_full.vx[-1]),
                                   samples=cfg_full.samples_per_cell, kn=cfg_full.kn)
    s = make(cfg)
    s.initialize(f)
    t0 = time.perf_counter()
    for n in range(1, steps + 1):
        s.step(n)
    dt = time.perf_counter() - t0
    s.close()
    return {"value": cfg.nx * cfg.ny * cfg.num_points * steps / dt, "unit": "cell-node updates/s", "cores": 1,
            "kind": "port", "cpu": cpu_model(), "compile": f"gcc {flags}",
            "sample": f"CPU oracle (serial C, gcc {flags}) on one core of {cpu_model()}, 16x16-cell cavity, "
                      f"{cfg.nvx}^3 nodes, {cfg.samples_per_cell} samples/cell, {steps} full steps, {dt:.1f} s"}


def cpu_baseline_threads(pkg, cfg_full, steps=6):
    """The same bounded sample on every host core the process may use (at
    most 16): one oracle context per thread, each stepping its own 16x16-cell
    cavity (the oracle keeps per-context scratch, so contexts are the unit of
    parallelism); ctypes releases the GIL during each step. Rate in
    cell-node updates/s over all threads."""
    import threading
    make, flags = native_oracle_solver(pkg)
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    threads = max(1, min(16, cores))
    cfg, f = pkg.lid_driven_cavity(nx=16, ny=16, n_v=cfg_full.nvx, vmax=float(cfg_full.vx[-1]),
                                   samples=cfg_full.samples_per_cell, kn=cfg_full.kn)
    solvers = [make(cfg) for _ in range(threads)]
    for s in solvers:
        s.initialize(f)
    errors = []

    def run(s):
        try:
            for n in range(1, steps + 1):
                s.step(n)
        except Exception as e:  # reported below; a failed thread voids the figure
            errors.append(e)

    ts = [threading.Thread(target=run, args=(s,)) for s in solvers]
    t0 = time.perf_counter()
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    dt = time.perf_counter() - t0
    for s in solvers:
        s.close()
    if errors:
        raise errors[0]
    return {"value": threads * cfg.nx * cfg.ny * cfg.num_points * steps / dt, "unit": "cell-node updates/s",
            "cores": threads, "kind": "port", "cpu": cpu_model(), "compile": f"gcc {flags}",
            "sample": f"{threads} threads, each the CPU oracle (serial C, gcc {flags}) on its own 16x16-cell cavity, "
                      f"{cfg.nvx}^3 nodes, {cfg.samples_per_cell} samples/cell, {steps} full steps, {dt:.1f} s "
                      f"({cores} cores in the affinity mask)"}


def measured_traffic(kernel, workload):
    """HBM bytes per launch of a kernel from the table of rocprofv3 PMC
    measurements (FETCH_SIZE and WRITE_SIZE, scaled by the copy calibration),
    or None when it has not been measured."""
    if not os.path.exists(TRAFFIC_TABLE):
        return None
    for line in open(TRAFFIC_TABLE):
        f = line.split()
        if len(f) >= 3 and not line.startswith("#") and f[0] == kernel and f[1] == workload:
            return float(f[2])
    return None


def measured_valu(kernel, workload):
    """Wave-level VALU instructions per launch from the rocprofv3 counter
    table, or None when it has not been measured."""
    if not os.path.exists(VALU_TABLE):
        return None
    for line in open(VALU_TABLE):
        f = line.split()
        if len(f) >= 3 and not line.startswith("#") and f[0] == kernel and f[1] == workload:
            return float(f[2])
    return None


def roofline(kernel, workload, alg_bytes, ms):
    achieved = alg_bytes / (ms * 1e-3) / 1e9 if ms > 0 else None
    out = {"bound": "hbm", "kernel": kernel, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": (achieved / HBM_PEAK_GBS) if achieved else None,
           "traffic": measured_traffic(kernel, workload), "algorithmic_bytes_per_launch": alg_bytes,
           "launch_ms": ms}
    # the issue rate that actually bounds the collision kernels: measured VALU
    # instructions per launch against one wave64 VALU instruction per 4 cycles
    # on each of the 1024 SIMDs at the 2.4 GHz peak engine clock
    valu = measured_valu(kernel, workload)
    if valu and ms > 0:
        out["valu_issue"] = {"instructions_per_launch": valu, "simds": SIMDS, "clock_ghz": CLOCK_GHZ,
                             "frac": valu * 4.0 / (SIMDS * CLOCK_GHZ * 1e9 * ms * 1e-3)}
    # which of the two rates the kernel runs closer to: "bound" stays the HBM
    # roofline the record is priced against; a kernel whose VALU issue rate is
    # the higher fraction is limited by its instruction stream, not by memory
    vf = out.get("valu_issue", {}).get("frac")
    out["limiter"] = "valu_issue" if (vf is not None and out["frac"] is not None and vf > out["frac"]) else "hbm"
    return out


def run_batched(pkg, count, steps, precision, tune=True):
    """The batched workload: admit `count` instances, autotune the collision
    tiles on them (the state is restored), time `steps` graph replays."""
    inst = pkg.batched_homogeneous(count, precision=precision)
    nodes = sum(c.num_points for c, _ in inst)
    samples = sum(c.samples_per_cell for c, _ in inst)
    b = pkg.Batch(inst[0][0], max_slots=count, node_capacity=nodes, cell_capacity=count)
    for c, f in inst:
        b.admit(c, f)
    b.step(1)
    b.synchronize()
    tuned = b.tune(1)[0] if tune else None
    b.step(1)
    b.synchronize()
    t0 = time.perf_counter()
    b.step(steps)
    b.synchronize()
    wall = time.perf_counter() - t0
    l1 = b.l1()
    tiles = b.tiles()
    b.close()
    return {"instances": count, "nodes": nodes, "samples_per_step": samples, "steps": steps,
            "ms_per_step": wall / steps * 1e3, "cell_node_updates_per_s": nodes * steps / wall,
            "collision_samples_per_s": samples * steps / wall, "tiles": tiles, "tuned": tuned,
            "l1_median_after": float(np.median(l1[l1 >= 0]))}


def run_batched_cavity(pkg, cfg, f, steps, warmup, precision):
    """The cavity as one slot of a batch: the same cells, collisions and sweep,
    but every step is one replay of the batch's captured HIP graph (shared
    persistent collision kernel fed from per-XCD work queues, per-slot 2-D
    walls and sweep)."""
    c = cfg.with_(precision=precision)
    b = pkg.Batch(c, max_slots=2, node_capacity=c.nx * c.ny * c.num_points, cell_capacity=c.nx * c.ny)
    try:
        b.admit(c, f)
        b.step(warmup)
        b.synchronize()
        t0 = time.perf_counter()
        b.step(steps)
        b.synchronize()
        wall = time.perf_counter() - t0
        tiles = b.tiles()
    finally:
        b.close()
    return {"ms_per_step": wall / steps * 1e3, "value": c.nx * c.ny * c.num_points * steps / wall,
            "precision": "fp32" if precision == pkg.capi.DVM_PREC_FP32 else "fp64", "tiles": tiles}


def run_shock(pkg, steps, warmup):
    """The 1-D normal shock: 200 cells x 31^3 nodes on +-6, VHS omega 0.81,
    Mach 2 with the Rankine-Hugoniot downstream state and a tanh profile,
    in/out flow at both ends, 50 000 samples per cell per step, the limited
    second-order sweep, plain FP64."""
    cfg, f = pkg.normal_shock_1d(nx=200, n_v=31, vmax=6.0, mach=2.0, samples=50_000)
    cfg = cfg.with_(conv_order=pkg.capi.DVM_CONV_MINMOD2, precision=pkg.capi.DVM_PREC_FP64)
    wall, phase, _ = run_gpu(pkg, cfg, f, steps, warmup)
    nodes = cfg.nx * cfg.num_points
    return {"workload": f"1-D normal shock Mach 2, {cfg.nx} cells x {cfg.nvx}^3 nodes on +-6, VHS omega {cfg.omega}, "
                        f"{cfg.samples_per_cell} samples/cell/step, minmod second-order sweep, in/out flow, FP64",
            "steps": steps, "ms_per_step": wall / steps * 1e3, "cell_node_updates_per_s": nodes * steps / wall,
            "collision_samples_per_s": cfg.samples_per_cell * cfg.nx * steps / wall,
            "phase_ms": {k: round(v, 4) for k, v in phase.items()}}


def run_relaxation(pkg, steps=200, cpu=True, cpu_steps=20):
    """The space-homogeneous relaxation (the reference's CPU-runnable case): 21^3
    nodes on +-5, two Maxwellians of equal density at T and 2T moving at +-1
    along x, hard spheres, 100 000 samples per step, dt = 0.05, FP64. One pass
    timed without host reads, then the same steps (same seed) with the moments
    and the L1 distance from equilibrium read every step, as the workload asks;
    the oracle on one core times a bounded number of the same steps."""
    cfg, f = pkg.homogeneous_relaxation()
    wall, phase, _ = run_gpu(pkg, cfg, f, steps, 0)
    props, l1 = [], []
    with pkg.Solver(cfg) as s:
        props.append(s.initialize(f)[0])
        l1.append(s.l1_to_equilibrium()[0])
        for n in range(1, steps + 1):
            s.step(n)
            props.append(s.calculate_physical_properties()[0])
            l1.append(s.l1_to_equilibrium()[0])
    props = np.array(props)
    out = {"workload": f"space-homogeneous relaxation, 1 cell x {cfg.nvx}^3 nodes on +-5, bi-Maxwellian T and 2T "
                       f"at +-1 along x, hard sphere, {cfg.samples_per_cell} samples/step, dt {cfg.dt}, FP64",
           "steps": steps, "ms_per_step": wall / steps * 1e3,
           "collision_samples_per_s": cfg.samples_per_cell * steps / wall,
           "phase_ms": {k: round(v, 4) for k, v in phase.items()},
           "l1_to_equilibrium": {str(n): float(l1[n]) for n in (0, 1, 10, 50, 100, steps) if n <= steps},
           "max_rel_drift": {"density": float(np.max(np.abs(props[:, 0] / props[0, 0] - 1.0))),
                             "velocity": float(np.max(np.abs(props[:, 1:4] - props[0, 1:4]))),
                             "temperature": float(np.max(np.abs(props[:, 4] / props[0, 4] - 1.0)))}}
    if cpu:
        make, flags = native_oracle_solver(pkg)
        o = make(cfg)
        o.initialize(f)
        t0 = time.perf_counter()
        for n in range(1, cpu_steps + 1):
            o.step(n)
        dt = time.perf_counter() - t0
        o.close()
        out["cpu_baseline"] = {"ms_per_step": dt / cpu_steps * 1e3, "cores": 1, "kind": "port", "cpu": cpu_model(),
                               "sample": f"CPU oracle (serial C, gcc {flags}), the same workload, {cpu_steps} steps, "
                                         f"{dt:.2f} s"}
    return out


def run_shock_steady(pkg, steps=2000):
    """The 1-D normal shock run to steady state, untimed: moments every 10
    steps over the last 200. A stationary shock carries the same mass flux
    n u through every cell; reported are its spread over the cells (in units
    of the inflow's) and the largest relative density change per cell over
    the last 100 steps."""
    cfg, f = pkg.normal_shock_1d(nx=200, n_v=31, vmax=6.0, mach=2.0, samples=50_000)
    cfg = cfg.with_(conv_order=pkg.capi.DVM_CONV_MINMOD2, precision=pkg.capi.DVM_PREC_FP64)
    hist = {}
    with pkg.Solver(cfg) as s:
        s.initialize(f)
        for n in range(1, steps + 1):
            s.step(n)
            if n > steps - 200 and n % 10 == 0:
                hist[n] = s.calculate_physical_properties()
    avg = np.mean([hist[n] for n in hist], axis=0)
    flux = avg[:, 0] * avg[:, 1]
    inflow = cfg.wall[0][0] * cfg.wall[0][1]
    return {"steps": steps, "mass_flux_spread": float((flux.max() - flux.min()) / inflow),
            "density_ratio_out_in": float(avg[-1, 0] / avg[0, 0]),
            "max_density_change_last_100_steps": float(np.max(np.abs(hist[steps][:, 0] / hist[steps - 100][:, 0] - 1.0)))}


def run_trajectory(pkg, cfg, f, steps):
    """FP32 against FP64 along one trajectory of the cavity (same seed), beside
    two FP64 runs with different seeds: after a few steps a pick that lands
    within rounding of a CDF boundary differs, and the runs are then
    independent Monte Carlo trajectories, so their difference is the method's
    noise."""
    def final(c):
        with pkg.Solver(c) as s:
            s.initialize(f)
            for n in range(1, steps + 1):
                s.step(n)
            return s.calculate_physical_properties(), s.output_moments(), s.download_f()
    p64, e64, f64 = final(cfg.with_(precision=pkg.capi.DVM_PREC_FP64))
    p32, e32, f32 = final(cfg.with_(precision=pkg.capi.DVM_PREC_FP32))
    same = moment_error(p32, p64)
    same.update(output_moment_error(e32, e64, f32, f64))
    del p32, e32, f32
    p64b, e64b, f64b = final(cfg.with_(precision=pkg.capi.DVM_PREC_FP64, seed=cfg.seed + 1000003))
    spread = moment_error(p64b, p64)
    spread.update(output_moment_error(e64b, e64, f64b, f64))
    return {"steps": steps, "fp32_vs_fp64_same_seed": same, "fp64_seed_spread": spread}


def run_serving(pkg, count, rate, precision):
    reqs = pkg.synthetic_requests(count=count, rate=rate, seed=5, precision=precision)
    pools = dict(node_capacity=160_000_000, cell_capacity=65536, max_slots=512)
    with pkg.Server(reqs[0].cfg, check_every=10, keep_moments=True, keep_f=True, **pools) as srv:
        st = srv.run(reqs)
    # error of the solves against the same solve in plain FP64 (same seed; every
    # other request, to bound the replay time), beside the spread between two
    # FP64 runs with different seeds (the Monte Carlo noise of the method) of
    # four of the shocks
    from importlib import import_module
    serve = import_module(pkg.__name__ + ".serve")
    sample = [o for o in st.outcomes if o.rid % 2 == 0]
    errs = serve.replay_fp64(reqs, sample, check_every=10, **pools)
    noise = serve.replay_fp64(reqs, [o for o in sample if o.kind == "shock"][:4], check_every=10,
                              seed_offset=1_000_000, **pools)
    for o in st.outcomes:
        o.f = None

    def summary(table, kind):
        e = [table[o.rid] for o in st.outcomes if o.kind == kind and o.rid in table]
        if not e:
            return None
        return {q: {"median": float(np.median([x[q] for x in e])), "max": float(np.max([x[q] for x in e]))}
                for q in e[0]}

    err_summary = {k: summary(errs, k) for k in ("homogeneous", "shock")}
    err_summary["shock_fp64_seed_spread"] = summary(noise, "shock")
    lat = {k: [o.latency for o in st.outcomes if o.kind == k] for k in ("homogeneous", "shock")}
    reasons = {}
    for o in st.outcomes:
        reasons[f"{o.kind}:{o.reason}"] = reasons.get(f"{o.kind}:{o.reason}", 0) + 1
    return {"requests": count, "rate_per_s": rate, "wall_s": st.wall_s,
            "latency_p50_s": st.latency_p50, "latency_p99_s": st.latency_p99,
            "latency_p50_s_by_kind": {k: float(np.percentile(v, 50)) for k, v in lat.items() if v},
            "latency_p99_s_by_kind": {k: float(np.percentile(v, 99)) for k, v in lat.items() if v},
            "collision_samples_per_s": st.samples / st.wall_s, "cell_node_updates_per_s": st.node_updates / st.wall_s,
            "max_resident": st.max_resident, "stop_reasons": reasons,
            "fp32_error_vs_fp64": err_summary,
            "workload": f"{count} requests, Poisson {rate}/s; 70% space-homogeneous mixtures (edge 15..41 odd, "
                        f"50 samples/node, stop at L1 <= 1e-4 or L1 plateau or 300 steps), 30% 1-D shocks "
                        f"(Mach U[1.2,3], 50-400 cells, edge 21..31 odd, 20000 samples/cell, run to steady "
                        f"state: mass flux within 2% of the inflow's and the 100-step mean density profile "
                        f"changing by at most 1e-3 of its peak, in two consecutive windows, or 6000 steps); "
                        f"FP32 collision arithmetic"}


def moment_error(p32, p64):
    """Max over cells of the relative error of density, velocity (in units of
    the thermal speed sqrt(T)) and temperature, FP32 path against FP64."""
    th = np.sqrt(np.abs(p64[:, 4]))
    errs = {"density": np.abs(p32[:, 0] - p64[:, 0]) / np.abs(p64[:, 0]),
            "velocity": np.max(np.abs(p32[:, 1:4] - p64[:, 1:4]), axis=1) / th,
            "temperature": np.abs(p32[:, 4] - p64[:, 4]) / np.abs(p64[:, 4])}
    return {k: float(np.max(v)) for k, v in errs.items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--nx", type=int, default=64)
    ap.add_argument("--nv", type=int, default=25)
    ap.add_argument("--samples", type=int, default=20000)
    ap.add_argument("--precision", choices=["both", "fp32", "fp64"], default="both")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calib", action="store_true", help="also run a 512 MiB device copy (profiling)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-batched-cavity", action="store_true", help="skip the cavity as a batch slot")
    ap.add_argument("--no-remap32", action="store_true", help="skip the FP32-remap variant of the cavity")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-shock", action="store_true")
    ap.add_argument("--no-fmc", action="store_true")
    ap.add_argument("--no-relaxation", action="store_true")
    ap.add_argument("--no-trajectory", action="store_true")
    ap.add_argument("--no-shard-tile", action="store_true")
    ap.add_argument("--trajectory-steps", type=int, default=500)
    ap.add_argument("--no-tune", action="store_true", help="batched: keep the default tiles (profiling runs)")
    ap.add_argument("--batched-instances", type=int, default=4096)
    ap.add_argument("--batched-steps", type=int, default=3)
    ap.add_argument("--serving-requests", type=int, default=40)
    ap.add_argument("--serving-rate", type=float, default=2.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    pkg = importlib.import_module("2d_dvm_microflows_amd")
    if args.copy_calib:
        copy_calibration()
    cfg, f = pkg.lid_driven_cavity(nx=args.nx, ny=args.nx, n_v=args.nv, samples=args.samples)
    ncell, N = cfg.nx * cfg.ny, cfg.num_points
    updates = ncell * N * args.steps
    runs, finals = {}, {}
    for prec in (("fp64", "fp32") if args.precision == "both" else (args.precision,)):
        c = cfg.with_(precision=pkg.capi.DVM_PREC_FP32 if prec == "fp32" else pkg.capi.DVM_PREC_FP64)
        finals[prec] = {}
        runs[prec] = run_gpu(pkg, c, f, args.steps, args.warmup, finals[prec])
        print(f"[bench] {prec}: {runs[prec][0] / args.steps * 1e3:.3f} ms/step", file=sys.stderr, flush=True)
    head = "fp32" if "fp32" in runs else "fp64"
    wall, phase, props = runs[head]
    value = updates / wall
    samples_per_step = cfg.samples_per_cell * ncell
    workload = f"cavity{cfg.nx}_{head}"
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-node updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head == "fp32" else "f64",
        "data": "synthetic",
        "config": {"workload": f"2-D lid-driven cavity {cfg.nx}x{cfg.ny} cells, {cfg.nvx}^3 velocity nodes on +-5, "
                               f"{cfg.samples_per_cell} VR samples/cell/step, hard sphere, Kn {cfg.kn}, lid 0.3, "
                               f"diffuse walls, {head.upper()} collision arithmetic (FP64 remap, moments"
                               f"{', conservation correction' if head == 'fp32' else ''})",
                   "cells": ncell, "velocity_nodes": N, "samples_per_cell": cfg.samples_per_cell,
                   "precision": head},
        "collision_samples_per_s": samples_per_step * args.steps / wall,
        "phase_ms": {k: round(v, 4) for k, v in phase.items()},
        # the collision kernel is the dominant one; it is bound by VALU work and
        # LDS atomics per sample rather than by HBM, so its HBM fraction is low
        "roofline": roofline("k_collide", workload, COLLIDE_BYTES_PER_NODE[head] * ncell * N, phase["collide"]),
        "roofline_sweep": roofline("k_convect2d", workload, SWEEP_BYTES_PER_NODE * ncell * N, phase["sweep"]),
    }
    copy_gbs = copy_bandwidth()
    out["device_copy_gbs"] = copy_gbs
    out["roofline_sweep"]["frac_of_device_copy"] = out["roofline_sweep"]["achieved"] / copy_gbs
    if "fp64" in runs and "fp32" in runs:
        w64 = runs["fp64"][0]
        out["fp64_path"] = {"value": updates / w64, "ms_per_step": w64 / args.steps * 1e3,
                            "phase_ms": {k: round(v, 4) for k, v in runs["fp64"][1].items()}}
        out["speedup_fp32_vs_fp64"] = w64 / wall
        out["fp32_moment_rel_error_vs_fp64"] = moment_error(runs["fp32"][2], runs["fp64"][2])
        out["fp32_moment_rel_error_vs_fp64"].update(
            output_moment_error(finals["fp32"]["ext"], finals["fp64"]["ext"], finals["fp32"]["f"], finals["fp64"]["f"]))
        out["fp32_moment_error_after_steps"] = args.warmup + args.steps
        if not args.no_remap32:
            # DVM_PREC_FP32_REMAP32: the remap's nearest nodes and weights in
            # FP32 as well (equally spaced axes). The workload names an FP64
            # remap, so this is reported beside the value, not as it.
            c = cfg.with_(precision=pkg.capi.DVM_PREC_FP32_REMAP32)
            fin = {}
            w, ph, pr = run_gpu(pkg, c, f, args.steps, args.warmup, fin)
            err = moment_error(pr, runs["fp64"][2])
            err.update(output_moment_error(fin["ext"], finals["fp64"]["ext"], fin["f"], finals["fp64"]["f"]))
            out["fp32_remap32"] = {"value": updates / w, "ms_per_step": w / args.steps * 1e3,
                                   "speedup_vs_fp64": w64 / w,
                                   "phase_ms": {k: round(v, 4) for k, v in ph.items()},
                                   "moment_rel_error_vs_fp64": err,
                                   "precision": "FP32 collision and remap, FP64 accumulation, moments and "
                                                "conservation correction"}
            print(f"[bench] fp32 remap32: {w / args.steps * 1e3:.3f} ms/step", file=sys.stderr, flush=True)
    finals.clear()
    if not args.no_batched and not args.no_batched_cavity:
        # the batched form of the cavity (one graph replay per step), both
        # precisions, against the plain FP64 context path above
        bc = {}
        for prec in ("fp32", "fp64"):
            code = pkg.capi.DVM_PREC_FP32 if prec == "fp32" else pkg.capi.DVM_PREC_FP64
            bc[prec] = run_batched_cavity(pkg, cfg, f, args.steps, args.warmup, code)
            print(f"[bench] batched cavity {prec}: {bc[prec]['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
        if "fp64" in runs:
            bc["speedup_batched_fp32_vs_plain_fp64"] = (runs["fp64"][0] / args.steps * 1e3) / bc["fp32"]["ms_per_step"]
        out["batched_cavity"] = bc
    if not args.no_fmc:
        # the same cavity with the full Monte Carlo kernel (iColl = 2): one
        # integral of samples drawn from the |f| CDF
        c = cfg.with_(coll_solver=pkg.capi.DVM_COLL_FMC,
                      precision=pkg.capi.DVM_PREC_FP32 if head == "fp32" else pkg.capi.DVM_PREC_FP64)
        wf, pf, _ = run_gpu(pkg, c, f, args.steps, args.warmup)
        out["fmc_path"] = {"value": updates / wf, "ms_per_step": wf / args.steps * 1e3, "precision": head,
                           "collision_samples_per_s": samples_per_step * args.steps / wf,
                           "phase_ms": {k: round(v, 4) for k, v in pf.items()}}
        print(f"[bench] fmc {head}: {wf / args.steps * 1e3:.3f} ms/step", file=sys.stderr, flush=True)
    if not args.no_shard_tile:
        # one GPU's share of the cavity on 8 GPUs (a 2 x 4 grid of 32 x 16-cell
        # tiles), at the same per-cell work; the halo each tile would exchange
        # per step is the upwind half of its interior faces' nodes
        ct, ft = pkg.lid_driven_cavity(nx=32, ny=16, n_v=args.nv, samples=args.samples)
        ct = ct.with_(precision=pkg.capi.DVM_PREC_FP32 if head == "fp32" else pkg.capi.DVM_PREC_FP64)
        wt, pt, _ = run_gpu(pkg, ct, ft, args.steps, args.warmup)
        face_cells = 2 * 16 + 2 * 32  # an interior tile's four faces
        out["shard_tile"] = {"workload": f"cavity tile 32x16 cells (one GPU's share of 64x64 on 8), {ct.nvx}^3 nodes, "
                                         f"{ct.samples_per_cell} samples/cell, {head}",
                             "ms_per_step": wt / args.steps * 1e3,
                             "cell_node_updates_per_s": 32 * 16 * N * args.steps / wt,
                             "phase_ms": {k: round(v, 4) for k, v in pt.items()},
                             "halo_bytes_per_step_interior_tile": face_cells * N * 8 // 2}
        print(f"[bench] shard tile: {out['shard_tile']['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    if not args.no_trajectory and "fp64" in runs and "fp32" in runs:
        out["fp32_trajectory"] = run_trajectory(pkg, cfg, f, args.trajectory_steps)
        print(f"[bench] trajectory: {out['fp32_trajectory']}", file=sys.stderr, flush=True)
    if not args.no_shock:
        out["shock_1d"] = run_shock(pkg, args.steps, args.warmup)
        out["shock_1d"]["steady_state"] = run_shock_steady(pkg)
        print(f"[bench] shock 1-D: {out['shock_1d']['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    if not args.no_relaxation:
        out["relaxation_c1"] = run_relaxation(pkg, cpu=not args.no_cpu)
        print(f"[bench] relaxation: {out['relaxation_c1']['ms_per_step']:.3f} ms/step", file=sys.stderr, flush=True)
    if not args.no_batched:
        out["batched"] = {}
        for prec in ("fp32", "fp64"):
            code = pkg.capi.DVM_PREC_FP32 if prec == "fp32" else pkg.capi.DVM_PREC_FP64
            out["batched"][prec] = run_batched(pkg, args.batched_instances, args.batched_steps, code,
                                               tune=not args.no_tune)
            print(f"[bench] batched {prec}: {out['batched'][prec]['ms_per_step']:.1f} ms/step",
                  file=sys.stderr, flush=True)
        out["batched"]["workload"] = (f"{args.batched_instances} space-homogeneous instances, edge U{{15..41 odd}}, "
                                      "1-3 Maxwellian mixtures, 50 samples/node, one replayed HIP graph per step, "
                                      "autotuned collision tiles")
    if not args.no_serving:
        out["serving"] = run_serving(pkg, args.serving_requests, args.serving_rate, pkg.capi.DVM_PREC_FP32)
        print(f"[bench] serving: p50 {out['serving']['latency_p50_s']:.3f} s, p99 {out['serving']['latency_p99_s']:.3f} s",
              file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(pkg, cfg)
        out["cpu_baseline"]["all_cores"] = cpu_baseline_threads(pkg, cfg)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
