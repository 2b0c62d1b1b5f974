#!/usr/bin/env python3
"""Benchmark of the fire/smoke step on one MI355X (voxel-steps/s).

A step is one pass of the simulation step (velocity advection, advect-react,
force/boundary/divergence, red-black SOR, projection, light volume) over one
synthetic scene, through the C ABI (include/firesim.h). The workload is the
128^3 FP32 scene: one emitter sphere (r=12 voxels at (64,12,64), fuel 1/s,
1500 K), seeded wind of 0-2 voxels/s in a random xz direction, turbulence
forcing of amplitude 0.5, SOR 80 iterations at omega 1.7, dt = 1/60. Fields are
resident in HBM; the timed region is `--steps` steps bracketed by device
synchronisation, after `--warmup` untimed steps.

Printed: one JSON line with the contract keys plus
  roofline      the dominant pass (largest GPU time) against HBM peak, its
                time per launch from HIP events recorded on the scene stream
                around every pass during the timed steps,
  cpu_baseline  the FP64 oracle (OpenMP) on this host, bounded sample,
  stages        per-pass times and algorithmic-bandwidth fractions,
  step_latency  p50/p99 of the per-step GPU times in the timed region,
  render        one HDR raymarch + tonemap frame after the timed steps.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "native-firesim_amd"))

import firesim  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s

# Algorithmic (compulsory) HBM bytes per cell of each pass, FP32 fields:
# velocity advection reads u,v,w and writes them (24 B); advect-react reads 10
# scalars + 3 velocities and writes 10 scalars (92 B); force/boundary/divergence
# reads heat, soot, k, progress, read-modify-writes u,v,w and writes d (44 B);
# one SOR iteration reads p and d and writes p (12 B); projection reads p and
# read-modify-writes u,v,w (28 B). The light volume (~6 B) counts as 0.
PASS_BYTES = {"advect_velocity": 24.0, "advect_react": 92.0, "force": 44.0, "pressure": 12.0, "project": 28.0, "light": 0.0}
PASS_KERNELS = {
    "advect_velocity": ["k_advect_velocity"],
    "advect_react": ["k_advect_react", "k_finalize_release"],
    "force": ["k_force_contrib", "k_force_apply"],
    "pressure": ["k_sor_tb", "k_sor_half"],
    "project": ["k_project"],
    "light": ["k_light_build", "k_light_propagate"],
}

CONFIGS = {
    "c1": dict(grid=(64, 64, 64), sor=40, turbulence=0.0, wind=False, emitters="single", render=(256, 256, 104),
               desc="64^3 FP32, emitter r=6 at (32,6,32), SOR 40 x omega 1.7"),
    "c2": dict(grid=(128, 128, 128), sor=80, turbulence=0.5, wind=True, emitters="single", render=(512, 512, 128),
               desc="128^3 FP32, emitter r=12 at (64,12,64), seeded wind 0-2 voxel/s, turbulence 0.5, SOR 80 x omega 1.7"),
    "c3": dict(grid=(256, 384, 256), sor=120, turbulence=0.5, wind=True, emitters="eight", render=(1280, 720, 104),
               desc="256x384x256 FP32, 8 seeded floor emitters, SOR 120 x omega 1.7"),
}


def workload(name: str, seed: int):
    cfg = CONFIGS[name]
    nx, ny, nz = cfg["grid"]
    rng = np.random.default_rng(seed)
    wind_x = wind_z = 0.0
    if cfg["wind"]:
        speed, theta = rng.uniform(0.0, 2.0), rng.uniform(0.0, 2.0 * math.pi)
        # voxels/s -> domain fractions/s per axis (the step's velocity unit)
        wind_x, wind_z = speed * math.cos(theta) / nx, speed * math.sin(theta) / nz
    if cfg["emitters"] == "single":
        r = 6.0 * nx / 64.0
        emitters = [(nx / 2, r, nz / 2, r, 1.0, 1500.0)]
    else:
        emitters = []
        for _ in range(8):
            r = rng.uniform(4.0, 10.0)
            emitters.append((rng.uniform(r, nx - r), r, rng.uniform(r, nz - r), r, rng.uniform(0.5, 1.5), rng.uniform(1200.0, 1800.0)))
    params = firesim.default_params(scene_id=firesim.SCENE_EMITTERS_ONLY, turbulence=cfg["turbulence"], wind_x=wind_x, wind_z=wind_z,
                                    sor_omega=1.7, sor_iterations=cfg["sor"], raymarch_steps=cfg["render"][2], noise_seed=seed,
                                    camera_distance=3.0, cinematic_mode=1)
    return cfg, params, emitters


def cpu_baseline(name: str, seed: int, budget_s: float):
    """FP64 oracle on this host's cores: 2 warm-up steps, then timed steps until
    `budget_s` seconds or 20 steps."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    oracle.set_threads(threads)
    cfg, params, emitters = workload(name, seed)
    o = oracle.OracleScene(*cfg["grid"], 32, 32)
    o.set_params(params)
    o.set_emitters(emitters)
    o.reset()
    o.step(2)
    n, t0 = 0, time.perf_counter()
    while n < 20:
        o.step(1)
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    cells = math.prod(cfg["grid"])
    o.close()
    return {"value": cells * n / dt, "unit": "voxel-steps/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle, same {cfg['grid'][0]}x{cfg['grid'][1]}x{cfg['grid'][2]} workload from reset: 2 untimed + {n} timed steps, {dt:.1f} s"}


def read_traffic(kernel_prefixes, grid):
    """HBM bytes per dispatch measured with rocprofv3 PMC counters (FETCH_SIZE,
    WRITE_SIZE, copy-calibrated), from the committed profile summary."""
    path = os.path.join(REPO, "profiles", "pmc_traffic.txt")
    if not os.path.exists(path):
        return None
    gtag = "x".join(str(g) for g in grid)
    for line in open(path):
        if line.startswith("#") or "=" not in line:
            continue
        kv = dict(tok.split("=", 1) for tok in line.split() if "=" in tok)
        if kv.get("grid") == gtag and any(kv.get("kernel", "").startswith(k) for k in kernel_prefixes):
            return float(kv["fetch_bytes"]) + float(kv["write_bytes"])
    return None


def calibrate_copy():
    """A plain 256 MiB device copy (torch elementwise copy kernel), for scaling
    the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = 64 * 1024 * 1024
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": n * 4, "launches": 3}))


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--sor-kernel", type=int, default=None, help="enum fs_sor_kernel (default: library default)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the library's default SOR kernel")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of timed oracle work for cpu_baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-render", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true", help="run only the 256 MiB copy used to scale PMC counters")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(scenes shard one per GPU with no exchange, see the design notes).", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return

    cfg, params, emitters = workload(args.config, args.seed)
    nx, ny, nz = cfg["grid"]
    fw, fh, _ = cfg["render"]
    cells = nx * ny * nz
    scene = firesim.Scene(nx, ny, nz, fw, fh)
    if args.sor_kernel is not None:
        scene.set_sor_kernel(args.sor_kernel)
    scene.set_params(params)
    scene.set_emitters(emitters)
    scene.reset()
    if args.warmup:
        scene.step(args.warmup)
    scene.synchronize()
    tuned = None
    if args.sor_kernel is None and not args.no_autotune:
        tuned = scene.autotune(reps=2)  # picks the fastest bit-identical SOR kernel for this grid

    # Timed region: the production path (captured step graph replayed per
    # step), HIP events between steps for the latency distribution.
    t0 = time.perf_counter()
    per_step = scene.step_latency(args.steps)  # synchronises at the end
    wall = time.perf_counter() - t0
    scene.synchronize()
    value = cells * args.steps / wall
    ms_per_step = wall * 1e3 / args.steps

    # Measured region: the same number of steps with HIP events recorded on the
    # scene stream between passes (direct launches), for the per-pass times and
    # the dominant kernel's time per launch.
    times, _ = scene.step_timed(args.steps)
    stages = {}
    for i, name in enumerate(firesim.STAGE_NAMES):
        ms = times.stage_ms[i] / args.steps
        per_cell = PASS_BYTES[name] * (params.sor_iterations if name == "pressure" else 1)
        gbs = per_cell * cells / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        stages[name] = {"ms": round(ms, 4), "alg_GBps": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    dominant = max(stages, key=lambda k: stages[k]["ms"])
    launches = times.pressure_launches if dominant == "pressure" else len(PASS_KERNELS[dominant])
    launch_ms = stages[dominant]["ms"] / launches
    bytes_per_launch = PASS_BYTES[dominant] * cells * (params.sor_iterations / launches if dominant == "pressure" else 1.0 / launches)
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9
    traffic = read_traffic(PASS_KERNELS[dominant], cfg["grid"])
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": PASS_KERNELS[dominant][0], "pass": dominant, "launch_ms": round(launch_ms, 5),
                "alg_bytes_per_launch": bytes_per_launch}

    render = None
    if not args.no_render:
        scene.render_bgra8()  # warm
        r0 = time.perf_counter()
        frame = scene.render_bgra8()
        render = {"ms": round((time.perf_counter() - r0) * 1e3, 3), "width": fw, "height": fh, "samples": cfg["render"][2],
                  "mean_luma": float(np.mean(((frame >> 16) & 255) * 0.2126 + ((frame >> 8) & 255) * 0.7152 + (frame & 255) * 0.0722))}
    metrics = scene.metrics()
    scene_sor_kernel = scene.sor_kernel()
    scene.close()

    cpu = None if args.no_cpu_baseline else cpu_baseline(args.config, args.seed, args.cpu_budget)
    out = {
        "metric": "voxel-steps/s at 128³ + served p50/p99 step latency vs plain FP32 path",
        "value": value,
        "unit": "voxel-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": cfg["desc"], "grid": list(cfg["grid"]), "sor_iterations": params.sor_iterations,
                   "sor_omega": round(params.sor_omega, 4), "dt": round(params.dt, 6), "seed": args.seed,
                   "sor_kernel": firesim.SOR_KERNELS[scene_sor_kernel],
                   "sor_autotune_ms": {k: round(v, 4) for k, v in tuned["sor_ms"].items()} if tuned else None,
                   "step_path": "captured HIP graph replayed per step",
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stages": stages,
        "step_latency_ms": {"p50": float(np.percentile(per_step, 50)), "p99": float(np.percentile(per_step, 99)),
                            "mean": float(np.mean(per_step))},
        "render": render,
        "sanity": {"invalid_cells": metrics["invalid_cells"], "heat_release_total": metrics["heat_release_total"],
                   "divergence_after_l2": metrics["divergence_after_l2"]},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
