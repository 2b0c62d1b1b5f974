"""GPU parity of the HIP step against the FP64 oracle, through the C ABI.

Tolerances (from the step's accuracy target): per pass and per step, the
relative L2 error of every field is <= 1e-5 for the FP32 path; after 100
free-running steps the total reaction heat release and the soot mass agree
within 1%. The HDR frame, whose per-pixel branches (empty-space skip,
early termination) can flip between FP32 and FP64, is held to relative L2
<= 1e-3; the tonemap to at most 1 LSB per channel on identical HDR input.
"""
import numpy as np
import pytest

import compare
import firesim
import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-5

GRID = (32, 40, 24)
EMITTERS = [(12.0, 4.0, 11.0, 4.0, 1.0, 1500.0), (22.0, 5.0, 14.0, 3.0, 0.8, 1350.0)]


@pytest.fixture(scope="module")
def warm():
    params = compare.emitter_params(wind_x=0.3, wind_z=-0.2, turbulence=0.6)
    state, t, rt = compare.warm_state(*GRID, params, EMITTERS, steps=30)
    return params, state, t, rt


def _pair(params, emitters, grid=GRID, frame=(48, 32), storage=firesim.STORAGE_FP32):
    s = firesim.Scene(*grid, frame[0], frame[1], storage=storage)
    o = oracle.OracleScene(*grid, frame[0], frame[1])
    for x in (s, o):
        x.set_params(params)
        x.set_emitters(emitters)
        x.reset()
    return s, o


@pytest.mark.parametrize("stage", range(6), ids=firesim.STAGE_NAMES)
def test_stage_parity(hip_lib, warm, stage):
    params, state, t, rt = warm
    s, o = _pair(params, EMITTERS)
    compare.load_both(s, o, state, t, rt)
    s.run_stage(stage)
    o.run_stage(stage)
    errs = compare.compare_fields(s, o, compare.STAGE_OUTPUTS[stage])
    print(firesim.STAGE_NAMES[stage], errs, flush=True)
    bad = {k: v for k, v in errs.items() if not v <= TOL}
    assert not bad, f"stage {firesim.STAGE_NAMES[stage]} exceeds {TOL}: {bad}"


def test_step_parity(hip_lib, warm):
    params, state, t, rt = warm
    s, o = _pair(params, EMITTERS)
    compare.load_both(s, o, state, t, rt)
    s.step(1)
    o.step(1)
    assert s.get_time() == o.get_time()
    errs = compare.compare_fields(s, o, compare.STEP_FIELDS)
    print("step", errs, flush=True)
    bad = {k: v for k, v in errs.items() if not v <= TOL}
    assert not bad, f"step exceeds {TOL}: {bad}"
    ms, mo = s.metrics(), o.metrics()
    assert ms["heat_release_step"] == pytest.approx(mo["heat_release_step"], rel=TOL)


def test_free_running_conservation(hip_lib):
    """64^3, one emitter at (32,6,32) r=6, 1500 K, fuel 1/s, SOR 40 x omega 1.7, 100 steps."""
    params = compare.emitter_params(turbulence=0.0, noise_seed=0)
    em = [(32.0, 6.0, 32.0, 6.0, 1.0, 1500.0)]
    s, o = _pair(params, em, grid=(64, 64, 64))
    s.step(100)
    o.step(100)
    ms, mo = s.metrics(), o.metrics()
    print({k: (ms[k], mo[k]) for k in ("heat_release_total", "soot_sum", "heat_sum", "fuel_sum")}, flush=True)
    assert mo["heat_release_total"] > 100.0
    assert ms["heat_release_total"] == pytest.approx(mo["heat_release_total"], rel=0.01)
    assert ms["soot_sum"] == pytest.approx(mo["soot_sum"], rel=0.01)
    assert ms["invalid_cells"] == 0
    # reference validation thresholds, src/main.cpp:4334-4336
    assert ms["divergence_after_l2"] < 1.25 and ms["divergence_after_max"] < 8.0


def test_metrics_parity(hip_lib, warm):
    params, state, t, rt = warm
    s, o = _pair(params, EMITTERS)
    compare.load_both(s, o, state, t, rt)
    ms, mo = s.metrics(), o.metrics()
    for k in ("heat_sum", "fuel_sum", "oxygen_sum", "soot_sum", "progress_sum", "turbulence_energy_sum",
              "soot_optical_depth_sum", "max_heat", "divergence_before_l2", "divergence_after_l2", "sor_residual_l2",
              "heat_release_proxy", "mean_scene_light", "mean_scene_shadow", "flame_mass_proxy", "smoke_mass_proxy"):
        assert ms[k] == pytest.approx(mo[k], rel=1e-5, abs=1e-9), k
    assert ms["invalid_cells"] == mo["invalid_cells"] == 0


def test_render_parity(hip_lib, warm):
    params, state, t, rt = warm
    p = compare.emitter_params(wind_x=0.3, wind_z=-0.2, turbulence=0.6, camera_distance=2.2, camera_pitch=0.1)
    s, o = _pair(p, EMITTERS, frame=(64, 48))
    compare.load_both(s, o, state, t, rt)
    hs, ho = s.render_hdr(), o.render_hdr()
    err = compare.rel_l2(hs[..., :3], ho[..., :3])
    print("hdr rel l2", err, flush=True)
    assert err <= 1e-3
    assert np.all(hs[..., 3] == 1.0)


@pytest.mark.parametrize("scene_id", [0, 1, 2, 3])
def test_ember_parity(hip_lib, scene_id):
    """The spark increment render(192 sparks) - render(0) against the oracle's (rel L2 <= 1e-3)."""
    p = compare.emitter_params(scene_id=scene_id, burner_center_count=0, ember_count=192, camera_distance=2.4)
    em = [(20.0, 3.0, 16.0, 5.0, 1.0, 1500.0)]
    state, t, rt = compare.warm_state(40, 48, 32, p, em, steps=20)
    s, o = _pair(p, em, grid=(40, 48, 32), frame=(96, 64))
    compare.load_both(s, o, state, t, rt)
    hs1, ho1 = s.render_hdr(), o.render_hdr()
    p.ember_count = 0
    s.set_params(p)
    o.set_params(p)
    hs0, ho0 = s.render_hdr(), o.render_hdr()
    inc_s, inc_o = hs1 - hs0.astype(np.float64), ho1 - ho0
    print(scene_id, "spark pixels", int((inc_o[..., :3].sum(-1) > 0).sum()), flush=True)
    if scene_id != 2:
        assert inc_o.max() > 0.0
    assert compare.rel_l2(inc_s[..., :3], inc_o[..., :3]) <= 1e-3
    assert compare.rel_l2(hs1[..., :3], ho1[..., :3]) <= 1e-3
    assert np.all(hs1[..., 3] == 1.0)


def test_fp16_pack(hip_lib, warm):
    """FP16 RGBA frame == numpy's round-to-nearest-even of max(0, HDR), alpha 1, bit for bit."""
    params, state, t, rt = warm
    s, o = _pair(params, EMITTERS, frame=(64, 48))
    compare.load_both(s, o, state, t, rt)
    hdr = s.render_hdr()
    packed = s.render_rgba16f()
    want = np.maximum(hdr, 0.0).astype(np.float16)
    want[..., 3] = 1.0
    assert packed.shape == (48, 64, 4)
    assert np.array_equal(packed.view(np.uint16), want.view(np.uint16))
    assert hdr[..., :3].max() > 0.0


def test_tonemap_parity(hip_lib, warm):
    params, state, t, rt = warm
    s, o = _pair(params, EMITTERS, frame=(64, 48))
    compare.load_both(s, o, state, t, rt)
    hdr = o.render_hdr()
    bs, bo = s.tonemap(hdr.astype(np.float32)), o.tonemap(hdr.astype(np.float32).astype(np.float64))
    cs = np.stack([(bs >> k) & 0xFF for k in (0, 8, 16, 24)], -1).astype(int)
    co = np.stack([(bo >> k) & 0xFF for k in (0, 8, 16, 24)], -1).astype(int)
    assert np.abs(cs - co).max() <= 1
    assert (cs == co).mean() >= 0.99


def test_reset_bed_scenes(hip_lib):
    """Procedural beds (scenes 0-2) exercise the integer-hash fbm bit for bit."""
    for scene_id in (0, 1, 2):
        p = compare.emitter_params(scene_id=scene_id, burner_center_count=0)
        s, o = _pair(p, [], grid=(40, 32, 28))
        errs = compare.compare_fields(s, o, firesim.SCALAR_FIELDS)
        assert all(v <= TOL for v in errs.values()), (scene_id, errs)
        s.step(2)
        o.step(2)
        errs = compare.compare_fields(s, o, compare.STEP_FIELDS)
        assert all(v <= TOL for v in errs.values()), (scene_id, errs)


def test_determinism(hip_lib, warm):
    params, state, t, rt = warm
    outs = []
    for _ in range(2):
        s = firesim.Scene(*GRID, 32, 32)
        s.set_params(params)
        s.set_emitters(EMITTERS)
        for name, arr in state.items():
            s.write(name, arr)
        s.set_time(t, rt)
        s.step(3)
        outs.append({n: s.read(n) for n in compare.STEP_FIELDS})
        s.close()
    for n in compare.STEP_FIELDS:
        assert np.array_equal(outs[0][n], outs[1][n]), n


@pytest.mark.parametrize("grid", [(4, 4, 4), (5, 9, 4), (17, 6, 11)])
def test_small_and_ragged_grids(hip_lib, grid):
    p = compare.emitter_params(sor_iterations=3)
    em = [(grid[0] / 2, 1.0, grid[2] / 2, 1.5, 1.0, 1500.0)]
    s, o = _pair(p, em, grid=grid, frame=(7, 5))
    s.step(3)
    o.step(3)
    errs = compare.compare_fields(s, o, compare.STEP_FIELDS)
    bad = {k: v for k, v in errs.items() if not v <= TOL}
    assert not bad, bad


def test_bf16_storage_error(hip_lib):
    """BF16 heat/fuel/soot storage: per-field error against the FP64 oracle is reported;
    bounded loosely here (bf16 keeps 8 mantissa bits)."""
    params = compare.emitter_params(turbulence=0.0)
    em = [(16.0, 3.0, 16.0, 3.0, 1.0, 1500.0)]
    s, o = _pair(params, em, grid=(32, 32, 32), storage=firesim.STORAGE_BF16)
    s.step(20)
    o.step(20)
    errs = compare.compare_fields(s, o, compare.STEP_FIELDS)
    print("bf16 errors", errs, flush=True)
    ms, mo = s.metrics(), o.metrics()
    assert ms["heat_release_total"] == pytest.approx(mo["heat_release_total"], rel=0.05)
    assert errs["heat"] < 0.05 and errs["u"] < 0.05


@pytest.mark.parametrize("grid", [(256, 12, 8), (48, 20, 16)])
def test_force_parity_across_noise_sharing(hip_lib, grid):
    """Force pass against the oracle where a wavefront's 64 cells share the
    lattice values of every noise octave (256 wide) and where the upper
    octaves span more than 64 lattice points and use per-lane hashes (48 wide)."""
    params = compare.emitter_params(wind_x=0.3, wind_z=-0.2, turbulence=0.6)
    em = [(grid[0] / 2.0, 3.0, grid[2] / 2.0, 3.0, 1.0, 1500.0)]
    state, t, rt = compare.warm_state(*grid, params, em, steps=10)
    s, o = _pair(params, em, grid=grid)
    compare.load_both(s, o, state, t, rt)
    s.run_stage(firesim.STAGE_FORCE)
    o.run_stage(firesim.STAGE_FORCE)
    errs = compare.compare_fields(s, o, compare.STAGE_OUTPUTS[firesim.STAGE_FORCE])
    print(grid, errs, flush=True)
    bad = {k: v for k, v in errs.items() if not v <= TOL}
    assert not bad, f"force pass exceeds {TOL} on {grid}: {bad}"
