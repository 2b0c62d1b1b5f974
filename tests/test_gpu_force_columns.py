"""The force pass face by face against the FP64 oracle on grids whose edges
cut the pass's 64 x 4 x 4 cell columns: partial columns in x, y and z, a
row wider than one wavefront (the u face shared across lanes 63 and 0 of the
next column), rows and planes that end inside a column. The per-field
relative L2 of the stage parity tests could let one wrong face through; here
every face and every divergence value is held to 1e-5 of the field's largest
magnitude, with the fbm noise, with curl noise, and with a closed top.
"""
import numpy as np
import pytest

import compare
import firesim
import oracle

pytestmark = pytest.mark.gpu

GRIDS = [(130, 9, 7), (64, 4, 4), (65, 13, 10), (47, 6, 5)]
MODES = {
    "fbm": dict(),
    "curl": dict(noise_forcing=1, curl_noise=0.7),
    "curl_closed_top": dict(noise_forcing=1, curl_noise=0.7, closed_top=1),
}


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("grid", GRIDS, ids=lambda g: "x".join(map(str, g)))
def test_force_faces_elementwise(hip_lib, grid, mode):
    nx, ny, nz = grid
    params = compare.emitter_params(wind_x=0.25, wind_z=-0.15, turbulence=0.8, **MODES[mode])
    em = [(nx * 0.4, 2.0, nz * 0.5, max(2.0, min(nx, nz) * 0.3), 1.0, 1500.0)]
    state, t, rt = compare.warm_state(nx, ny, nz, params, em, steps=8)
    s = firesim.Scene(nx, ny, nz, 32, 32)
    o = oracle.OracleScene(nx, ny, nz, 32, 32)
    for x in (s, o):
        x.set_params(params)
        x.set_emitters(em)
        x.reset()
    compare.load_both(s, o, state, t, rt)
    s.write("pressure", np.full((nz, ny, nx), 3.0, dtype=np.float32))
    o.write("pressure", np.full((nz, ny, nx), 3.0))
    s.run_stage(firesim.STAGE_FORCE)
    o.run_stage(firesim.STAGE_FORCE)
    print(f"force {grid} {mode}", flush=True)
    for name in ("u", "v", "w", "divergence"):
        g = s.read(name).astype(np.float64)
        r = o.read(name)
        scale = max(np.abs(r).max(), 1e-30)
        worst = np.abs(g - r).max() / scale
        where = np.unravel_index(np.argmax(np.abs(g - r)), g.shape)
        assert worst <= 1e-5, f"{name}: max |gpu - oracle| = {worst:.3e} of the field's max, at (z, y, x) = {where}"
    assert np.all(s.read("pressure") == 0.0)
    s.close()
    o.close()
