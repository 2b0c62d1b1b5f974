#!/usr/bin/env python3
"""Benchmark of the fire/smoke step on one MI355X (voxel-steps/s).

A step is one pass of the simulation step (velocity advection, advect-react,
force/boundary/divergence, red-black SOR, projection, light volume) over one
synthetic scene, through the C ABI (include/firesim.h). The workload is the
128^3 FP32 scene: one emitter sphere (r=12 voxels at (64,12,64), fuel 1/s,
1500 K), seeded wind of 0-2 voxels/s in a random xz direction, divergence-free
curl-noise forcing of amplitude 0.5, SOR 80 iterations at omega 1.7,
dt = 1/60, a 200-step run. Fields are resident in HBM; the timed region is
`--steps` steps bracketed by device synchronisation: the last `--steps` of the
200-step run (or those after `--warmup` untimed steps, if that is later).

Printed: one JSON line with the contract keys plus
  roofline      the dominant pass (largest GPU time) against HBM peak, its
                time per launch from HIP events recorded on the scene stream
                around every pass during the timed steps,
  cpu_baseline  the FP64 oracle (OpenMP) on this host, bounded sample,
  stages        per-pass times and algorithmic-bandwidth fractions,
  step_latency  p50/p99 of the per-step GPU times in the timed region,
  render        one HDR raymarch + tonemap frame after the timed steps,
  tall          the 256x384x256 scene (8 emitters, SOR 120): its last 10
                steps, and its pressure kernel's rate against HBM peak,
  served        a Poisson stream of requests (cubic grids 32..192, lognormal
                step counts, BF16 heat/fuel/soot, a 256x256 frame every 10
                steps) through the continuous-batching server: p50/p99 step
                latency and aggregate voxel-steps/s, against the same requests
                run one at a time on the plain FP32 path,
  batched       48 uneven scenes (48^3..128^3, 64x128x64, SOR 30-100) in one
                persistent launch per pass, FP32 and BF16, against the same
                scenes one at a time on the plain FP32 path.
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
import workloads  # noqa: E402

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
    "pressure": ["k_sor_tb", "k_sor_col", "k_sor_half"],
    "project": ["k_project"],
    "light": ["k_light_build", "k_light_propagate"],
}

CONFIGS = workloads.SINGLE


def workload(name: str, seed: int):
    return workloads.single_scene(name, seed)


def cpu_share() -> int:
    """CPUs this process may use: its affinity mask, capped by a cgroup CPU quota."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            n = min(n, max(1, int(int(quota) // int(period))))
    except (OSError, ValueError):
        pass
    return max(1, n)


def native_oracle() -> tuple[str | None, str]:
    """Builds the oracle for this host's CPU (-march=native) into a temporary
    directory; returns (library path or None, how it was built)."""
    import subprocess
    import tempfile
    out = os.path.join(tempfile.mkdtemp(prefix="oracle_native_"), "liboracle.so")
    cmd = ["g++", "-O3", "-march=native", "-fopenmp", "-fPIC", "-std=c++17", "-shared", "-o", out,
           os.path.join(REPO, "oracle", "oracle.cpp")]
    try:
        r = subprocess.run(cmd, capture_output=True, timeout=300)
    except (OSError, subprocess.TimeoutExpired):
        return None, "prebuilt x86-64-v3 (native build failed to run)"
    if r.returncode != 0:
        return None, "prebuilt x86-64-v3 (native build failed)"
    return out, "g++ -O3 -march=native -fopenmp"


def cpu_baseline(name: str, seed: int, budget_s: float, state=None, start_step: int = 0):
    """FP64 oracle on this host's CPU share, built for this CPU: from `state`
    (the GPU's fields at step `start_step`, so both time the same developed
    fire) or from reset, one untimed step, then timed steps until `budget_s`
    seconds or 20 steps."""
    lib, build = native_oracle()
    if lib:
        os.environ["ORACLE_LIB"] = lib
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle
    threads = cpu_share()
    oracle.set_threads(threads)
    cfg, params, emitters = workload(name, seed)
    o = oracle.OracleScene(*cfg["grid"], 32, 32)
    o.set_params(params)
    o.set_emitters(emitters)
    o.reset()
    if state is not None:
        fields, t, rt = state
        for k, a in fields.items():
            o.write(k, a.astype(np.float64))
        o.set_time(t, rt)
    o.step(1)
    n, t0 = 0, time.perf_counter()
    while n < 20:
        o.step(1)
        n += 1
        if time.perf_counter() - t0 > budget_s:
            break
    dt = time.perf_counter() - t0
    cells = math.prod(cfg["grid"])
    o.close()
    origin = f"the GPU state at step {start_step}" if state is not None else "reset"
    return {"value": cells * n / dt, "unit": "voxel-steps/s", "cores": threads, "kind": "port", "cpu_model": cpu_model(),
            "host_logical_cpus": os.cpu_count(), "host_physical_cores": physical_cores(), "build": build,
            "sample": f"FP64 oracle, same {cfg['grid'][0]}x{cfg['grid'][1]}x{cfg['grid'][2]} workload from {origin}: "
                      f"1 untimed + {n} timed steps, {dt:.1f} s, {threads} OpenMP threads (this job's CPU share)"}


def profile_kernel_name(sor_kernel: str) -> str:
    """rocprofv3's short name of the SOR kernel a scene runs: tb_t24_z16_k4 ->
    k_sor_tb<24,4,16>, col_t24_z12_k4 -> k_sor_col<24,4,12> (tile edge,
    half-sweeps per launch, tile depth)."""
    if sor_kernel == "half_sweep":
        return "k_sor_half"
    form, *rest = sor_kernel.split("_")
    parts = dict((tok[0], int(tok[1:])) for tok in rest)
    return f"k_sor_{form}<{parts['t']},{parts['k']},{parts.get('z', parts['t'])}>"


def read_traffic(kernel_names, grid):
    """HBM bytes per dispatch measured with rocprofv3 PMC counters (FETCH_SIZE,
    WRITE_SIZE, copy-calibrated), from the committed profile summary; None when
    the summary has no line for this exact kernel and grid."""
    path = os.path.join(REPO, "profiles", "pmc_traffic.txt")
    if not os.path.exists(path):
        return None
    gtag = "x".join(str(g) for g in grid)
    for line in open(path):
        if line.startswith("#") or "=" not in line:
            continue
        kv = dict(tok.split("=", 1) for tok in line.split() if "=" in tok)
        if kv.get("grid") == gtag and kv.get("kernel") in kernel_names:
            return float(kv["fetch_bytes"]) + float(kv["write_bytes"])
    return None


def _cpuinfo_first() -> dict:
    """Key/value pairs of the first processor in /proc/cpuinfo."""
    out = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if not line.strip():
                    break
                if ":" in line:
                    k, v = line.split(":", 1)
                    out[k.strip()] = v.strip()
    except OSError:
        pass
    return out


def cpu_model() -> str:
    """Model name with vendor, family, model and stepping numbers and the widest SIMD."""
    info = _cpuinfo_first()
    flags = info.get("flags", "").split()
    simd = "avx512f" if "avx512f" in flags else ("avx2" if "avx2" in flags else "sse")
    return (f"{info.get('model name', 'unknown')} ({info.get('vendor_id', '?')} family {info.get('cpu family', '?')} "
            f"model {info.get('model', '?')} stepping {info.get('stepping', '?')}, {simd})")


def physical_cores() -> int | None:
    """Distinct (physical id, core id) pairs in /proc/cpuinfo."""
    cores, phys = set(), None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("physical id"):
                    phys = line.split(":", 1)[1].strip()
                elif line.startswith("core id"):
                    cores.add((phys, line.split(":", 1)[1].strip()))
    except OSError:
        return None
    return len(cores) or None


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


def make_scene(grid, params, emitters, storage=firesim.STORAGE_FP32, frame=(64, 64)):
    s = firesim.Scene(*grid, frame[0], frame[1], storage)
    s.set_params(params)
    s.set_emitters(emitters)
    s.reset()
    return s


def progress(msg: str) -> None:
    """One line on stderr per finished section (stdout carries only the JSON line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def pcts(ms) -> dict:
    ms = np.asarray(ms, dtype=np.float64)
    return {"p50": round(float(np.percentile(ms, 50)), 4), "p99": round(float(np.percentile(ms, 99)), 4),
            "mean": round(float(np.mean(ms)), 4)} if ms.size else None


def plain_serial_step_ms(grid, params, emitters, steps):
    """The plain one-scene-at-a-time FP32 path: one scene, reference launch
    structure (one launch per SOR half-sweep, direct launches, no graph)."""
    s = make_scene(grid, params, emitters)
    s.set_tuning(sor_kernel=0, use_graph=False)
    s.step(1)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(steps)
    s.synchronize()
    dt = (time.perf_counter() - t0) / steps
    s.close()
    return dt * 1e3


def bench_batched(seed: int, steps: int, count: int, use_graph: bool = True):
    """Uneven batch: `count` scenes of the batch shapes in one persistent
    launch per pass, FP32 and BF16 storage, against the same scenes stepped
    one at a time on the plain FP32 path. (Per-field error of batched FP32
    and BF16 scenes against the FP64 oracle is a GPU test.)"""
    specs = workloads.batch_scenes(seed, count)
    cells = sum(math.prod(sp["grid"]) for sp in specs)
    result = {"scenes": count, "shapes": [list(s) for s in workloads.BATCH_SHAPES], "cells_per_step": cells,
              "sor_iterations": [int(sp["params"].sor_iterations) for sp in specs]}
    for mode, storage in (("fp32", firesim.STORAGE_FP32), ("bf16", firesim.STORAGE_BF16)):
        scenes = [make_scene(sp["grid"], sp["params"], sp["emitters"], storage) for sp in specs]
        with firesim.Batch() as batch:
            for sc in scenes:
                batch.add(sc)
            if not use_graph:
                batch.set_tuning(use_graph=False)
            batch.step(2)
            tuned = batch.autotune(reps=1)
            if not use_graph:
                batch.set_tuning(sor_kernel=tuned["chosen"], use_graph=False)
            batch.step(3)  # the tuned kernel's step graph is captured here, not in the timed steps
            batch.synchronize()
            t0 = time.perf_counter()
            ms = batch.step_latency(steps)
            wall = time.perf_counter() - t0
        for sc in scenes:
            sc.close()
        result[mode] = {"voxel_steps_per_s": cells * steps / wall, "ms_per_step": wall * 1e3 / steps, "step_ms": pcts(ms),
                        "sor_kernel": tuned["chosen"], "step_path": "captured HIP graph" if use_graph else "direct launches"}
    # the same scenes one at a time on the plain FP32 path (3 timed steps each)
    plain_ms = sum(plain_serial_step_ms(sp["grid"], sp["params"], sp["emitters"], 3) for sp in specs)
    result["plain_fp32_serial"] = {"voxel_steps_per_s": cells / (plain_ms * 1e-3), "ms_per_step": plain_ms}
    result["speedup_vs_plain_fp32"] = {m: result[m]["voxel_steps_per_s"] / result["plain_fp32_serial"]["voxel_steps_per_s"] for m in ("fp32", "bf16")}
    return result


def bench_tall(seed: int, steps: int = 10):
    """The tall single scene (256x384x256, 8 emitters, SOR 120): the last
    `steps` of its 100-step run on the captured-graph step path, then the same
    number of steps with HIP events between passes for the pressure kernel's
    algorithmic rate (the configuration the reference's bandwidth is quoted on)."""
    cfg, params, emitters = workload("c3", seed)
    nx, ny, nz = cfg["grid"]
    cells = nx * ny * nz
    scene = firesim.Scene(nx, ny, nz, 64, 64)
    scene.set_params(params)
    scene.set_emitters(emitters)
    scene.reset()
    scene.step(cfg["steps"] - steps)
    scene.synchronize()
    scene.autotune(reps=1)
    scene.step_latency(0)  # captures the chosen kernel's step graphs outside the timed steps
    t0 = time.perf_counter()
    scene.step_latency(steps)
    wall = time.perf_counter() - t0
    times, _ = scene.step_timed(steps)
    ms = times.stage_ms[firesim.STAGE_PRESSURE] / steps
    launches = times.pressure_launches
    gbs = PASS_BYTES["pressure"] * params.sor_iterations * cells / (ms * 1e-3) / 1e9
    out = {"workload": cfg["desc"], "grid": list(cfg["grid"]), "sor_iterations": params.sor_iterations, "timed_steps": f"{cfg['steps'] - steps + 1}-{cfg['steps']}",
           "voxel_steps_per_s": cells * steps / wall, "ms_per_step": wall * 1e3 / steps, "sor_kernel": firesim.SOR_KERNELS[scene.sor_kernel()],
           "pressure": {"ms": round(ms, 4), "launches": launches, "launch_ms": round(ms / launches, 5), "alg_GBps": round(gbs, 1),
                        "frac": round(gbs / HBM_PEAK_GBS, 4)}}
    scene.close()
    return out


def _serve(reqs, budget_gb: float, storage: int, use_graph: bool, max_round_voxels: int = 0, sor_kernel: str = "tb_t24_z28_k4") -> dict:
    """One pass of the request stream through fs_server; its statistics."""
    # A fixed SOR tile (default 24x24x28, the fastest batched tile on this
    # request mix of mostly large scenes, whose edges are multiples of 16): timing
    # every kernel on the resident mix inside the stream would stall its rounds.
    with firesim.Server(hbm_budget_bytes=int(budget_gb * (1 << 30)), max_active=256, sor_kernel=firesim.SOR_KERNELS.index(sor_kernel),
                        use_graph=use_graph, max_round_voxels=max_round_voxels) as sv:
        for r in reqs:
            sv.submit(firesim.make_request(*r["grid"], r["params"], r["emitters"], steps=r["steps"], render_every=r["render_every"],
                                           frame=r["frame"], storage=storage, arrival_s=r["arrival"]))
        t0 = time.perf_counter()
        sv.run()
        wall = time.perf_counter() - t0
        st = sv.stats()
    return {"voxel_steps": st["voxel_steps"], "busy_s": round(st["busy_s"], 3), "wall_s": round(wall, 3),
            "voxel_steps_per_s": st["voxel_steps"] / st["busy_s"] if st["busy_s"] > 0 else None,
            "step_ms": {"p50": round(st["step_ms_p50"], 3), "p99": round(st["step_ms_p99"], 3), "mean": round(st["step_ms_mean"], 3),
                        "max": round(st["step_ms_max"], 3)},
            "queue_ms": {"p50": round(st["queue_ms_p50"], 3), "p99": round(st["queue_ms_p99"], 3)},
            "max_active": st["max_active_seen"], "rounds": st["rounds"], "frames": st["frames"],
            "peak_resident_gb": round(st["peak_resident_bytes"] / (1 << 30), 2), "sor_kernel": firesim.SOR_KERNELS[st["sor_kernel"]],
            "failed": st["requests_failed"]}


def bench_served(seed: int, count: int, rate: float, budget_gb: float, storage: int, use_graph: bool = True, target_ms: float = 10.0,
                 sor_kernel: str = "tb_t24_z28_k4"):
    """Served stream: Poisson requests through fs_server (continuous batching,
    HBM-budget admission, a frame every 10 steps), once without a round cap
    and once with the cap that holds a round to `target_ms` at the uncapped
    rate (the p99 step latency is reported against that target), then the same
    requests serially through the plain FP32 path."""
    reqs = workloads.served_requests(seed, count, rate)
    total_voxel_steps = sum(math.prod(r["grid"]) * r["steps"] for r in reqs)
    served = {"requests": count, "arrival_rate_per_s": rate, "hbm_budget_gb": budget_gb,
              "storage": "bf16" if storage == firesim.STORAGE_BF16 else "fp32"}
    served.update(_serve(reqs, budget_gb, storage, use_graph, sor_kernel=sor_kernel))
    st = {"voxel_steps": served["voxel_steps"], "requests_failed": served["failed"]}
    # 40% of the target left for the frames rendered between steps, graph
    # captures and rounds whose members run more SOR iterations than the mean
    cap = int(served["voxel_steps_per_s"] * target_ms * 1e-3 * 0.6)
    capped = _serve(reqs, budget_gb, storage, use_graph, max_round_voxels=cap, sor_kernel=sor_kernel)
    assert capped["voxel_steps"] == total_voxel_steps and capped["failed"] == 0
    served["latency_capped"] = dict(capped, target_step_ms=target_ms, max_round_voxels=cap,
                                    p99_within_target=capped["step_ms"]["p99"] <= target_ms)
    assert st["voxel_steps"] == total_voxel_steps and st["requests_failed"] == 0
    # serially, plain FP32 path, same requests back to back (arrivals ignored: the GPU never waits)
    lat = []
    t0 = time.perf_counter()
    for r in reqs:
        s = make_scene(r["grid"], r["params"], r["emitters"], firesim.STORAGE_FP32, r["frame"])
        s.set_tuning(sor_kernel=0, use_graph=False)
        for k in range(1, r["steps"] + 1):
            a = time.perf_counter()
            s.step(1)
            if k % r["render_every"] == 0 or k == r["steps"]:
                s.render_bgra8()
            else:
                s.synchronize()
            lat.append((time.perf_counter() - a) * 1e3)
        s.close()
    serial_wall = time.perf_counter() - t0
    served["plain_fp32_serial"] = {"voxel_steps_per_s": total_voxel_steps / serial_wall, "wall_s": round(serial_wall, 3), "step_ms": pcts(lat)}
    served["speedup_vs_plain_fp32"] = served["voxel_steps_per_s"] / served["plain_fp32_serial"]["voxel_steps_per_s"]
    served["latency_capped"]["speedup_vs_plain_fp32"] = capped["voxel_steps_per_s"] / served["plain_fp32_serial"]["voxel_steps_per_s"]
    return served


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20,
                    help="untimed steps before the timed ones; raised so the timed steps end at the configuration's last step")
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--sor-kernel", type=int, default=None, help="enum fs_sor_kernel (default: library default)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the library's default SOR kernel")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of timed oracle work for cpu_baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-render", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true", help="run only the 256 MiB copy used to scale PMC counters")
    ap.add_argument("--no-served", action="store_true", help="skip the served request stream")
    ap.add_argument("--served-requests", type=int, default=80)
    ap.add_argument("--served-rate", type=float, default=40.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--served-budget-gb", type=float, default=200.0)
    ap.add_argument("--served-fp32", action="store_true", help="serve FP32 fields instead of BF16 heat/fuel/soot")
    ap.add_argument("--served-no-graph", action="store_true", help="direct launches instead of captured batch step graphs")
    ap.add_argument("--served-target-ms", type=float, default=10.0, help="step latency target of the capped served pass")
    ap.add_argument("--served-sor-kernel", default="tb_t24_z28_k4", help="fixed SOR tile of the served stream (a name of firesim.SOR_KERNELS)")
    ap.add_argument("--no-batched", action="store_true", help="skip the 48-scene uneven batch")
    ap.add_argument("--no-tall", action="store_true", help="skip the 256x384x256 scene's pressure rate")
    ap.add_argument("--batched-scenes", type=int, default=48)
    ap.add_argument("--batched-steps", type=int, default=10)
    ap.add_argument("--batched-no-graph", action="store_true", help="direct launches instead of captured batch step graphs")
    ap.add_argument("--only", choices=["single", "batched", "served"], default=None, help="run one section only")
    ap.add_argument("--from-reset", action="store_true",
                    help="time the steps right after --warmup untimed steps from reset (short profiled runs), "
                         "not the last steps of the configuration's run")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: runs on one GPU only; --gpus {args.gpus} is not supported "
              "(scenes shard one per GPU with no exchange, see the design notes).", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.only == "batched":
        print(json.dumps({"batched": bench_batched(args.seed, args.batched_steps, args.batched_scenes, not args.batched_no_graph)}))
        return
    if args.only == "served":
        storage = firesim.STORAGE_FP32 if args.served_fp32 else firesim.STORAGE_BF16
        print(json.dumps({"served": bench_served(args.seed, args.served_requests, args.served_rate, args.served_budget_gb, storage,
                                                 not args.served_no_graph, args.served_target_ms, args.served_sor_kernel)}))
        return
    if args.only == "single":
        args.no_batched = args.no_served = args.no_tall = True

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
    # The timed steps are the last --steps of the configuration's run (a young
    # fire is cheaper: cold cells take the reaction's early-out), or follow
    # --warmup untimed steps when that is later.
    start = args.warmup if args.from_reset else max(args.warmup, cfg["steps"] - args.steps)
    if start:
        scene.step(start)
    scene.synchronize()
    tuned = None
    if args.sor_kernel is None and not args.no_autotune:
        tuned = scene.autotune(reps=2)  # picks the fastest bit-identical SOR kernel for this grid
    # the same developed state seeds the CPU baseline
    start_state = ({n: scene.read(n) for n in firesim.STATE_FIELDS}, *scene.get_time()[:2])
    # Zero steps: captures the step graphs of the chosen kernel (the autotuner
    # dropped them) before the timed region, which then holds only steps.
    scene.step_latency(0)

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
    progress(f"single scene: {args.steps} timed steps, {value / 1e9:.3f} G voxel-steps/s")
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
    if dominant == "pressure":
        kernel_names = [profile_kernel_name(firesim.SOR_KERNELS[scene.sor_kernel()])]
    else:
        kernel_names = [PASS_KERNELS[dominant][0], PASS_KERNELS[dominant][0] + "<false>"]
    traffic = read_traffic(kernel_names, cfg["grid"])
    copy_gbs = firesim.copy_bandwidth(1 << 30, 20)  # achievable: streaming float4 copy of 1 GiB
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": kernel_names[0], "pass": dominant, "launch_ms": round(launch_ms, 5),
                "alg_bytes_per_launch": bytes_per_launch,
                "copy_GBps": round(copy_gbs, 1), "frac_of_copy": round(achieved / copy_gbs, 4)}

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

    progress("per-pass timing and render done")
    cpu = None if args.no_cpu_baseline else cpu_baseline(args.config, args.seed, args.cpu_budget, start_state, start)
    progress("cpu baseline done")
    tall = None if args.no_tall or args.config != "c2" else bench_tall(args.seed)
    progress("tall scene done")
    batched = None if args.no_batched else bench_batched(args.seed, args.batched_steps, args.batched_scenes, not args.batched_no_graph)
    progress("batched done")
    served = None if args.no_served else bench_served(args.seed, args.served_requests, args.served_rate, args.served_budget_gb,
                                                      firesim.STORAGE_FP32 if args.served_fp32 else firesim.STORAGE_BF16,
                                                      not args.served_no_graph, args.served_target_ms, args.served_sor_kernel)
    out = {
        "metric": "voxel-steps/s at 128³ + served p50/p99 step latency vs plain FP32 path",
        "value": value,
        "unit": "voxel-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": start,
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
                   "timed_steps": f"{start + 1}-{start + args.steps} of the {cfg['steps']}-step run ({start} untimed steps first)",
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stages": stages,
        "step_latency_ms": {"p50": float(np.percentile(per_step, 50)), "p99": float(np.percentile(per_step, 99)),
                            "mean": float(np.mean(per_step))},
        "render": render,
        "sanity": {"invalid_cells": metrics["invalid_cells"], "heat_release_total": metrics["heat_release_total"],
                   "divergence_after_l2": metrics["divergence_after_l2"]},
        "tall": tall,
        "served": served,
        "batched": batched,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
