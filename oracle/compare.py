"""Parity helpers (test infrastructure): teacher-forced comparison of the HIP
path against the FP64 oracle.

Protocol: a warm state is produced by the oracle, rounded to FP32, loaded into
both the oracle and the HIP scene together with the same clock, one pass (or
step) runs on both, and every field is compared by relative L2.
"""
from __future__ import annotations

import numpy as np

import firesim
import oracle as oracle_mod

# Fields a pass writes (src/fire_cuda.cu launch order :3196-3318).
STAGE_OUTPUTS = {
    firesim.STAGE_ADVECT_VELOCITY: ["u", "v", "w"],
    firesim.STAGE_ADVECT_REACT: list(firesim.SCALAR_FIELDS),
    firesim.STAGE_FORCE: ["u", "v", "w", "divergence", "pressure"],
    firesim.STAGE_PRESSURE: ["pressure"],
    firesim.STAGE_PROJECT: ["u", "v", "w"],
    firesim.STAGE_LIGHT: ["light", "shadow"],
}
STEP_FIELDS = list(firesim.SCALAR_FIELDS) + ["u", "v", "w", "pressure", "divergence", "light", "shadow"]


def rel_l2(a: np.ndarray, b: np.ndarray) -> float:
    """||a-b|| / ||b||; RMS difference when the reference is (near) zero."""
    a = np.asarray(a, dtype=np.float64).ravel()
    b = np.asarray(b, dtype=np.float64).ravel()
    diff = np.linalg.norm(a - b)
    ref = np.linalg.norm(b)
    if ref > 1e-12 * np.sqrt(b.size):
        return float(diff / ref)
    return float(diff / np.sqrt(max(1, b.size)))


def emitter_params(scene_id=3, **kw):
    base = dict(scene_id=scene_id, turbulence=0.5, sor_omega=1.7, sor_iterations=40, raymarch_steps=64, noise_seed=7)
    base.update(kw)
    return firesim.default_params(**base)


def warm_state(nx, ny, nz, params, emitters, steps, frame=(32, 32)):
    """Oracle state after `steps` steps from reset, rounded to FP32, plus its clocks."""
    o = oracle_mod.OracleScene(nx, ny, nz, *frame)
    o.set_params(params)
    o.set_emitters(emitters)
    o.reset()
    o.step(steps)
    state = {k: v.astype(np.float32) for k, v in o.state().items()}
    t, rt, _ = o.get_time()
    o.close()
    return state, t, rt


def load_both(scene, orc, state, t, rt):
    for name, arr in state.items():
        scene.write(name, arr)
        orc.write(name, arr.astype(np.float64))
    scene.set_time(t, rt)
    orc.set_time(t, rt)


def compare_fields(scene, orc, names) -> dict:
    return {n: rel_l2(scene.read(n), orc.read(n)) for n in names}
