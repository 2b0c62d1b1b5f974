"""GPU parity of the extensions against the FP64 oracle, through the C ABI:
divergence-free curl-noise forcing, the closed top, the directional-light
transmittance and render, the gizmo overlay on the BGRA frame, and the
sin-hash visual noise mode (not a parity mode: checked for determinism and
plausibility only, since its values follow the platform's sinf).

Tolerances as in test_gpu_parity: relative L2 <= 1e-5 per pass and step for
the FP32 path, HDR frame <= 1e-3, BGRA within 1 LSB per channel.
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
EXT = dict(noise_forcing=1, curl_noise=0.5, dir_light=1, dir_light_x=0.35, dir_light_y=-1.0, dir_light_z=0.2, dir_light_r=1.4,
           dir_light_g=1.2, dir_light_b=1.0)


def _params(**kw):
    base = dict(wind_x=0.3, wind_z=-0.2, turbulence=0.6, camera_distance=2.2, camera_pitch=0.1)
    base.update(kw)
    return compare.emitter_params(**base)


@pytest.fixture(scope="module", params=[0, 1], ids=["open_top", "closed_top"])
def warm(request):
    params = _params(closed_top=request.param, **EXT)
    state, t, rt = compare.warm_state(*GRID, params, EMITTERS, steps=30)
    return params, state, t, rt


def _pair(params, emitters=EMITTERS, grid=GRID, frame=(48, 32)):
    s = firesim.Scene(*grid, *frame)
    o = oracle.OracleScene(*grid, *frame)
    for x in (s, o):
        x.set_params(params)
        x.set_emitters(emitters)
        x.reset()
    return s, o


OUTPUTS = dict(compare.STAGE_OUTPUTS)
OUTPUTS[firesim.STAGE_LIGHT] = ["light", "shadow", "dir_transmittance"]


@pytest.mark.parametrize("stage", range(6), ids=firesim.STAGE_NAMES)
def test_stage_parity_with_extensions(hip_lib, warm, stage):
    params, state, t, rt = warm
    s, o = _pair(params)
    compare.load_both(s, o, state, t, rt)
    s.run_stage(stage)
    o.run_stage(stage)
    errs = compare.compare_fields(s, o, OUTPUTS[stage])
    print(firesim.STAGE_NAMES[stage], params.closed_top, errs, flush=True)
    bad = {k: v for k, v in errs.items() if not v <= TOL}
    assert not bad, bad
    if params.closed_top and stage in (firesim.STAGE_ADVECT_VELOCITY, firesim.STAGE_FORCE, firesim.STAGE_PROJECT):
        assert np.all(s.read("v")[:, -1, :] == 0)
    s.close()
    o.close()


def test_curl_noise_divergence_free_on_the_gpu(hip_lib):
    """A zero state forced by curl noise alone: the FP32 faces' divergence is at
    FP32 roundoff of the face values, and the faces match the oracle's."""
    p = _params(wind_x=0.0, wind_z=0.0, noise_forcing=1, curl_noise=0.8)
    s, o = _pair(p, emitters=[], grid=(48, 40, 36))
    for x in (s, o):
        for name in firesim.STATE_FIELDS:
            x.write(name, np.zeros(x.shape(name), np.float32))
        x.set_time(1.7, 1.7)
        x.run_stage(firesim.STAGE_FORCE)
    u, v, w = (s.read(n).astype(np.float64) for n in ("u", "v", "w"))
    scale = max(np.abs(u).max(), np.abs(v).max(), np.abs(w).max())
    div = (u[:, :, 1:] - u[:, :, :-1]) + (v[:, 1:, :] - v[:, :-1, :]) + (w[1:] - w[:-1])
    assert scale > 1e-3 and np.abs(div).max() <= 1e-6 * scale
    errs = compare.compare_fields(s, o, ["u", "v", "w"])
    assert all(e <= TOL for e in errs.values()), errs
    s.close()
    o.close()


def test_step_parity_with_extensions(hip_lib, warm):
    params, state, t, rt = warm
    s, o = _pair(params)
    compare.load_both(s, o, state, t, rt)
    s.step(1)
    o.step(1)
    errs = compare.compare_fields(s, o, compare.STEP_FIELDS + ["dir_transmittance"])
    print("step", params.closed_top, errs, flush=True)
    bad = {k: v for k, v in errs.items() if not v <= TOL}
    assert not bad, bad
    s.close()
    o.close()


def test_render_parity_with_directional_light(hip_lib, warm):
    params, state, t, rt = warm
    s, o = _pair(params, frame=(64, 48))
    compare.load_both(s, o, state, t, rt)
    for x in (s, o):
        x.run_stage(firesim.STAGE_LIGHT)
    hs, ho = s.render_hdr(), o.render_hdr()
    err = compare.rel_l2(hs[..., :3], ho[..., :3])
    # the light's contribution itself, against the oracle's
    p0 = _params(closed_top=params.closed_top, **dict(EXT, dir_light=0))
    for x in (s, o):
        x.set_params(p0)
    inc_s, inc_o = hs - s.render_hdr().astype(np.float64), ho - o.render_hdr()
    print("hdr rel l2", err, "increment", compare.rel_l2(inc_s[..., :3], inc_o[..., :3]), flush=True)
    assert err <= 1e-3
    assert inc_o[..., :3].max() > 1e-3
    assert compare.rel_l2(inc_s[..., :3], inc_o[..., :3]) <= 1e-3
    s.close()
    o.close()


@pytest.mark.parametrize("scene_id,active", [(3, 1), (2, 2), (0, 1)])
def test_gizmo_overlay_parity(hip_lib, scene_id, active):
    """Tonemap + gizmos on identical HDR input: within 1 LSB of the oracle, and
    the gizmo pixels (those that change with show_gizmos) agree too."""
    p = _params(scene_id=scene_id, show_gizmos=1, active_gizmo=active, camera_yaw=0.3)
    s, o = _pair(p, frame=(160, 120))
    rng = np.random.default_rng(scene_id)
    hdr = rng.uniform(0.0, 1.5, (120, 160, 4)).astype(np.float32) ** 2
    bs, bo = s.tonemap(hdr), o.tonemap(hdr.astype(np.float64))
    p.show_gizmos = 0
    s.set_params(p)
    plain = s.tonemap(hdr)
    cs = np.stack([(bs >> k) & 0xFF for k in (0, 8, 16, 24)], -1).astype(int)
    co = np.stack([(bo >> k) & 0xFF for k in (0, 8, 16, 24)], -1).astype(int)
    drawn = bs != plain
    print("gizmo pixels", int(drawn.sum()), flush=True)
    assert drawn.sum() > 50
    assert np.abs(cs - co).max() <= 1
    assert (cs == co).mean() >= 0.99
    s.close()
    o.close()


def test_gizmos_in_the_reference_interface_frame(hip_lib):
    """fireCudaStepAndRender with the default showGizmos = 1 draws the gizmos
    over the frame that showGizmos = 0 gives; nothing else changes."""
    import fire_api
    from fire_api import FireSettings
    api = fire_api.FireApi()
    frames = {}
    for g in (1, 0):
        assert api.fireCudaInitialize(96, 64, 200, 140), api.last_error()
        try:
            frames[g] = np.zeros(96 * 64, np.uint32)
            assert api.step_and_render(frames[g], FireSettings(width=96, height=64, showGizmos=g), None), api.last_error()
        finally:
            api.fireCudaShutdown()
    drawn = frames[1] != frames[0]
    print("gizmo pixels", int(drawn.sum()), flush=True)
    assert 20 < drawn.sum() < 0.1 * drawn.size


@pytest.mark.parametrize("scene_id", [0, 1, 2])
def test_sin_hash_visual_mode(hip_lib, scene_id):
    """The reference's sin-hash noise (visual mode) on the three reference
    scenes: deterministic, finite, different from the integer hash, and a fire
    of the same order (heat sum within a factor 3 after 40 steps)."""
    p = compare.emitter_params(scene_id=scene_id, burner_center_count=0, noise_mode=1, turbulence=0.9)
    runs = []
    for _ in range(2):
        s = firesim.Scene(48, 56, 40, 64, 48)
        s.set_params(p)
        s.reset()
        s.step(40)
        runs.append((s.read("heat"), s.read("u"), s.metrics(), s.render_bgra8()))
        s.close()
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
    assert np.array_equal(runs[0][3], runs[1][3])
    assert np.all(np.isfinite(runs[0][0])) and runs[0][2]["invalid_cells"] == 0
    p.noise_mode = 0
    s = firesim.Scene(48, 56, 40, 64, 48)
    s.set_params(p)
    s.reset()
    s.step(40)
    heat_int, m_int = s.read("heat"), s.metrics()
    s.close()
    assert not np.array_equal(runs[0][0], heat_int)
    ratio = runs[0][2]["heat_sum"] / max(m_int["heat_sum"], 1e-9)
    print(scene_id, "heat sum sin/int", ratio, flush=True)
    assert 1 / 3 <= ratio <= 3


def test_batch_with_extensions_equals_alone(hip_lib):
    """Batch members with curl noise, a closed top and the directional light step
    bit-identically to the same scenes alone."""
    def make(grid, seed, **kw):
        nx, ny, nz = grid
        s = firesim.Scene(nx, ny, nz, 24, 16)
        p = firesim.default_params(scene_id=3, turbulence=0.5, sor_omega=1.7, sor_iterations=9, noise_seed=seed, raymarch_steps=24, **kw)
        s.set_params(p)
        s.set_emitters([(nx / 2, 3.0, nz / 2, 3.0, 1.0, 1500.0)])
        s.reset()
        return s
    cases = [((24, 28, 20), dict(noise_forcing=1, curl_noise=0.6)), ((32, 24, 28), dict(closed_top=1, dir_light=1, dir_light_r=1.0)),
             ((20, 32, 24), dict(noise_forcing=1, curl_noise=0.3, closed_top=1, dir_light=1, dir_light_x=0.5, dir_light_g=2.0)),
             ((16, 16, 16), dict())]
    alone = [make(g, 5 + i, **kw) for i, (g, kw) in enumerate(cases)]
    mixed = [make(g, 5 + i, **kw) for i, (g, kw) in enumerate(cases)]
    fields = firesim.SCALAR_FIELDS + ["u", "v", "w", "pressure", "divergence", "light", "shadow"]
    with firesim.Batch() as batch:
        for s in mixed:
            batch.add(s)
        batch.step(4)
        for s in alone:
            s.step(4)
        for i, (a, m) in enumerate(zip(alone, mixed)):
            for f in fields:
                assert np.array_equal(a.read(f), m.read(f)), (i, f)
            if cases[i][1].get("dir_light"):
                assert np.array_equal(a.read("dir_transmittance"), m.read("dir_transmittance")), i
    for s in alone + mixed:
        s.close()
