"""The captured-graph step gives the same bits as direct launches, on every
replay, with the program's own eager work (field reads, metrics) between
replays, and stays valid across parameter changes, resets and clock edits."""
import numpy as np
import pytest

import compare
import firesim

pytestmark = pytest.mark.gpu

GRID = (40, 36, 28)
EMITTERS = [(20.0, 4.0, 14.0, 4.0, 1.0, 1500.0)]


def _scene(use_graph, sor_kernel=None):
    s = firesim.Scene(*GRID, 24, 16)
    s.set_params(compare.emitter_params(wind_x=0.2, turbulence=0.5, sor_iterations=13))
    s.set_emitters(EMITTERS)
    if sor_kernel is not None:
        s.set_sor_kernel(sor_kernel)
    s.set_tuning(use_graph=use_graph)
    s.reset()
    return s


@pytest.mark.parametrize("sor_kernel", [0, 2])
def test_graph_replays_match_direct_launches(hip_lib, sor_kernel):
    g, d = _scene(True, sor_kernel), _scene(False, sor_kernel)
    for step in range(7):
        g.step(1)
        d.step(1)
        for name in compare.STEP_FIELDS:
            assert np.array_equal(g.read(name), d.read(name)), (step, name)
        assert g.get_time() == d.get_time()
        assert g.metrics()["heat_release_total"] == d.metrics()["heat_release_total"]
    # parameter change invalidates and recaptures; both paths continue in step
    p = compare.emitter_params(wind_x=-0.3, turbulence=0.8, sor_iterations=9)
    for s in (g, d):
        s.set_params(p)
    g.step(3)
    d.step(3)
    for name in compare.STEP_FIELDS:
        assert np.array_equal(g.read(name), d.read(name)), name
    # direct single-pass runs and clock edits in between keep the graph path consistent
    for s in (g, d):
        s.run_stage(firesim.STAGE_ADVECT_VELOCITY)
        s.set_time(1.25, 1.5)
    g.step(2)
    d.step(2)
    for name in compare.STEP_FIELDS:
        assert np.array_equal(g.read(name), d.read(name)), name
    assert g.get_time() == d.get_time()


def test_step_latency_path(hip_lib):
    s = _scene(True)
    per = s.step_latency(6)
    assert per.shape == (6,) and np.all(per > 0)
    t, rt, frame = s.get_time()
    assert frame == 6
