"""Synthetic workloads of the benchmark: seeded scene, batch and request-stream
generators in the step's own units.

Voxel-unit inputs are converted the way the step expects them: emitter
spheres stay in voxels (fs_emitter), wind speeds in voxels/s become domain
fractions per second per axis (the reference's velocity unit, normalized
coordinates src/fire_cuda.cu:880-932), temperatures in kelvin become
heat = (T - 293)/335 inside the library.
"""
from __future__ import annotations

import math

import numpy as np

import firesim

# steps: the length of the run the configuration names; extra: fs_params extensions
SINGLE = {
    "c1": dict(grid=(64, 64, 64), sor=40, turbulence=0.0, wind=False, emitters="single", render=(256, 256, 104), steps=100,
               extra=dict(closed_top=1, noise_forcing=1, curl_noise=0.0, dir_light=1, dir_light_x=0.35, dir_light_y=-1.0,
                          dir_light_z=0.25, dir_light_r=1.6, dir_light_g=1.5, dir_light_b=1.35),
               desc="64^3 FP32, emitter r=6 at (32,6,32), closed walls, no noise forcing, SOR 40 x omega 1.7, "
                    "render with a directional light"),
    "c2": dict(grid=(128, 128, 128), sor=80, turbulence=0.5, wind=True, emitters="single", render=(512, 512, 128), steps=200,
               extra=dict(noise_forcing=1, curl_noise=0.5),
               desc="128^3 FP32, emitter r=12 at (64,12,64), seeded wind 0-2 voxel/s, divergence-free curl-noise forcing "
                    "of amplitude 0.5, SOR 80 x omega 1.7"),
    "c3": dict(grid=(256, 384, 256), sor=120, turbulence=0.5, wind=True, emitters="eight", render=(1280, 720, 104), steps=100,
               extra={}, desc="256x384x256 FP32, 8 seeded floor emitters, SOR 120 x omega 1.7"),
}

BATCH_SHAPES = [(48, 48, 48), (64, 64, 64), (96, 96, 96), (128, 128, 128), (64, 128, 64)]


def seeded_wind(rng, nx, nz, max_speed=2.0):
    """Constant wind of U(0, max_speed) voxels/s in a U(0, 2pi) xz direction."""
    speed, theta = rng.uniform(0.0, max_speed), rng.uniform(0.0, 2.0 * math.pi)
    return speed * math.cos(theta) / nx, speed * math.sin(theta) / nz


def floor_emitters(rng, nx, nz, count, r_lo, r_hi):
    """Spheres resting on the floor at uniform positions: (cx, cy, cz, r, fuel rate, kelvin)."""
    out = []
    for _ in range(count):
        r = rng.uniform(r_lo, r_hi)
        out.append((rng.uniform(r, nx - r), r, rng.uniform(r, nz - r), r, rng.uniform(0.5, 1.5), rng.uniform(1200.0, 1800.0)))
    return out


def params_for(seed, sor, turbulence, wind=(0.0, 0.0), raymarch=64, **extra):
    return firesim.default_params(scene_id=firesim.SCENE_EMITTERS_ONLY, turbulence=turbulence, wind_x=wind[0], wind_z=wind[1],
                                  sor_omega=1.7, sor_iterations=sor, raymarch_steps=raymarch, noise_seed=seed,
                                  camera_distance=3.0, cinematic_mode=1, **extra)


def single_scene(name: str, seed: int):
    """(config, fs_params, emitters) of one of the single-scene workloads."""
    cfg = SINGLE[name]
    nx, ny, nz = cfg["grid"]
    rng = np.random.default_rng(seed)
    wind = seeded_wind(rng, nx, nz) if cfg["wind"] else (0.0, 0.0)
    if cfg["emitters"] == "single":
        r = 6.0 * nx / 64.0
        emitters = [(nx / 2, r, nz / 2, r, 1.0, 1500.0)]
    else:
        emitters = floor_emitters(rng, nx, nz, 8, 4.0, 10.0)
    return cfg, params_for(seed, cfg["sor"], cfg["turbulence"], wind, cfg["render"][2], **cfg["extra"]), emitters


def batch_scenes(seed: int, count: int = 48):
    """Uneven batch: grids drawn from BATCH_SHAPES, 1-4 floor emitters of radius
    U(3, 0.08 n), SOR iterations U{30..100}, seeded wind and turbulence 0.5."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        nx, ny, nz = BATCH_SHAPES[int(rng.integers(len(BATCH_SHAPES)))]
        ems = floor_emitters(rng, nx, nz, int(rng.integers(1, 5)), 3.0, max(3.5, 0.08 * min(nx, nz)))
        sor = int(rng.integers(30, 101))
        wind = seeded_wind(rng, nx, nz)
        out.append(dict(grid=(nx, ny, nz), params=params_for(seed * 1000 + i, sor, 0.5, wind), emitters=ems))
    return out


def served_requests(seed: int, count: int, rate_per_s: float, sor: int = 80, frame=(256, 256), render_every: int = 10):
    """Poisson stream: cubic edge U{32, 48, ..., 192}, lognormal step count
    (median 60, sigma 0.75, clipped to [10, 500]), emitter and wind as the 128^3
    scene (r = 6n/64 at the floor centre, wind U(0, 2) voxels/s), a frame every
    `render_every` steps."""
    rng = np.random.default_rng(seed)
    out, t = [], 0.0
    edges = list(range(32, 193, 16))
    for i in range(count):
        t += rng.exponential(1.0 / rate_per_s)
        n = int(edges[int(rng.integers(len(edges)))])
        steps = int(np.clip(round(rng.lognormal(math.log(60.0), 0.75)), 10, 500))
        r = 6.0 * n / 64.0
        wind = seeded_wind(rng, n, n)
        out.append(dict(grid=(n, n, n), params=params_for(seed * 1000 + i, sor, 0.5, wind), emitters=[(n / 2, r, n / 2, r, 1.0, 1500.0)],
                        steps=steps, render_every=render_every, frame=frame, arrival=t))
    return out
