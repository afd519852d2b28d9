"""Throughput of the DVM time step on one MI355X.

Workload (the 2-D lid-driven cavity microflow): 64x64 cells, 25^3 velocity
nodes on +-5, hard spheres, Kn = 0.1, four diffuse walls at T_w = 1, lid at
0.3 along x, uniform equilibrium start, 20 000 collision samples per cell per
step. One step = CDF build + VR collision sampling + wall ghosts + upwind sweep
+ moments, with the distributions resident in HBM before timing starts.

Prints one JSON line: cell-node updates/s (cells x velocity nodes x steps /
wall time of the timed steps), the per-phase device times from HIP events on
the library's stream, the sweep's HBM roofline and the CPU oracle's rate on a
bounded sample of the same per-cell work.
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


def run_gpu(pkg, cfg, f, steps, warmup):
    """Times `steps` fused steps after `warmup` ones; returns wall seconds, mean
    per-phase device ms (HIP events on the library's stream) and the moments."""
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
    s.close()
    return (t1 - t0), {k: v / max(calls, 1) for k, v in phase.items()}, props


def cpu_baseline(pkg, cfg_full, steps=4):
    """The oracle (serial C, one core) on a 16x16-cell cavity with the same
    velocity grid and samples per cell; rate in cell-node updates/s."""
    from tests._oracle import oracle_solver
    cfg, f = pkg.lid_driven_cavity(nx=16, ny=16, n_v=cfg_full.nvx, vmax=float(cfg_full.vx[-1]),
                                   samples=cfg_full.samples_per_cell, kn=cfg_full.kn)
    s = oracle_solver(cfg)
    s.initialize(f)
    t0 = time.perf_counter()
    for n in range(1, steps + 1):
        s.step(n)
    dt = time.perf_counter() - t0
    s.close()
    return {"value": cfg.nx * cfg.ny * cfg.num_points * steps / dt, "unit": "cell-node updates/s", "cores": 1,
            "kind": "port",
            "sample": f"CPU oracle (serial C, -O2), 16x16-cell cavity, {cfg.nvx}^3 nodes, "
                      f"{cfg.samples_per_cell} samples/cell, {steps} full steps, {dt:.1f} s"}


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
    runs = {}
    for prec in (("fp64", "fp32") if args.precision == "both" else (args.precision,)):
        c = cfg.with_(precision=pkg.capi.DVM_PREC_FP32 if prec == "fp32" else pkg.capi.DVM_PREC_FP64)
        runs[prec] = run_gpu(pkg, c, f, args.steps, args.warmup)
        print(f"[bench] {prec}: {runs[prec][0] / args.steps * 1e3:.3f} ms/step", file=sys.stderr, flush=True)
    head = "fp32" if "fp32" in runs else "fp64"
    wall, phase, props = runs[head]
    value = updates / wall
    sweep_ms = phase["sweep"]
    alg_bytes = 2 * 8 * ncell * N
    achieved = alg_bytes / (sweep_ms * 1e-3) / 1e9 if sweep_ms > 0 else None
    samples_per_step = cfg.samples_per_cell * ncell
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
        "roofline": {"bound": "hbm", "kernel": "k_convect2d (upwind sweep)",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": None,
                     "algorithmic_bytes_per_launch": alg_bytes},
    }
    if "fp64" in runs and "fp32" in runs:
        w64 = runs["fp64"][0]
        out["fp64_path"] = {"value": updates / w64, "ms_per_step": w64 / args.steps * 1e3,
                            "phase_ms": {k: round(v, 4) for k, v in runs["fp64"][1].items()}}
        out["speedup_fp32_vs_fp64"] = w64 / wall
        out["fp32_moment_rel_error_vs_fp64"] = moment_error(runs["fp32"][2], runs["fp64"][2])
        out["fp32_moment_error_after_steps"] = args.warmup + args.steps
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(pkg, cfg)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
