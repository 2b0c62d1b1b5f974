"""Writes the regression fixtures in this directory from the FP64 oracle.

Each case is a tiny scene (a few hundred cells) stepped a few times from
reset by the oracle, stored as inputs (fs_params bytes, emitters, grid, step
count) and expected outputs (every field after the steps, FP64, plus the
metrics and the HDR frame). The reference holds no field-level vectors of its
own, so these fixtures pin later changes of the oracle and of the HIP path to
today's oracle output; they are data written by this script, nothing else.

Run from the repository root after building the oracle:
    python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(REPO, "native-firesim_amd"), os.path.join(REPO, "oracle")]

import firesim  # noqa: E402
import oracle  # noqa: E402

CASES = {
    # emitters only, wind in x and z, turbulence forcing, SOR 12 at omega 1.7
    "emitter_6x7x5": dict(grid=(6, 7, 5), frame=(12, 9), steps=4,
                          params=dict(scene_id=3, turbulence=0.6, wind_x=0.25, wind_z=-0.15, sor_omega=1.7, sor_iterations=12,
                                      noise_seed=11, raymarch_steps=24),
                          emitters=[(3.0, 1.5, 2.5, 1.6, 1.0, 1500.0)]),
    # the reference campfire bed (scene 1) with the mouse source held down
    "campfire_8x8x8": dict(grid=(8, 8, 8), frame=(10, 10), steps=3,
                           params=dict(scene_id=1, left_down=1, mouse_x=0.5, mouse_y=0.2, sor_iterations=40, noise_seed=3,
                                       raymarch_steps=32),
                           emitters=[]),
    # gas burner bed, ragged grid, right-button soot source, 2 steps
    "burner_9x5x7": dict(grid=(9, 5, 7), frame=(8, 6), steps=2,
                         params=dict(scene_id=2, right_down=1, mouse_x=0.3, mouse_y=0.4, turbulence=1.2, sor_iterations=20,
                                     noise_seed=5, raymarch_steps=24),
                         emitters=[(2.0, 1.0, 3.0, 1.2, 0.8, 1300.0), (6.0, 1.0, 3.0, 1.0, 1.2, 1700.0)]),
}


def make(name, c):
    p = firesim.default_params(**c["params"])
    o = oracle.OracleScene(*c["grid"], *c["frame"])
    o.set_params(p)
    o.set_emitters(c["emitters"])
    o.reset()
    o.step(c["steps"])
    out = {f"field_{k}": v.astype(np.float64) for k, v in o.state().items()}
    out["hdr"] = o.render_hdr().astype(np.float64)
    m = o.metrics()
    o.close()
    meta = dict(grid=list(c["grid"]), frame=list(c["frame"]), steps=c["steps"], params=c["params"], emitters=c["emitters"],
                metrics={k: v for k, v in m.items()})
    np.savez_compressed(os.path.join(HERE, f"{name}.npz"), params_bytes=np.frombuffer(bytes(p), dtype=np.uint8), **out)
    with open(os.path.join(HERE, f"{name}.json"), "w") as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print("wrote", name)


if __name__ == "__main__":
    oracle.set_threads(1)
    for name, c in CASES.items():
        make(name, c)
