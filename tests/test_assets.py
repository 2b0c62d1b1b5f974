"""Host-side reference scene placement (fs_scene_emitter_params, the
reference's loadSceneEmitterParams src/main.cpp:1397-1476) and frame colour
statistics (fs_image_stats_of, computeImageStats :2635-2690). No GPU."""
import os

import numpy as np
import pytest

import firesim

ASSETS = "/root/reference/assets/fire-scenes"  # present only where the reference is checked out
F32 = lambda x: float(np.float32(x))  # noqa: E731


def test_built_in_placements_without_scene_files(tmp_path):
    room = firesim.scene_emitter_params(0, str(tmp_path))
    assert (room.emitter_radius, room.emitter_height_band_norm, room.burner_center_count) == pytest.approx((0.48, 0.06, 0))
    camp = firesim.scene_emitter_params(1, str(tmp_path))
    assert camp.emitter_radius == pytest.approx(0.32)
    assert camp.emitter_height_norm == pytest.approx((0.045 - 0.02) / 2.03, rel=1e-6)
    assert camp.emitter_height_band_norm == pytest.approx(0.20 / 2.03, rel=1e-6)
    burner = firesim.scene_emitter_params(2, str(tmp_path))
    assert burner.emitter_radius == pytest.approx(0.24)
    with pytest.raises(firesim.FireSimError):
        firesim.scene_emitter_params(3, str(tmp_path))


def test_first_occurrence_lookup_and_mesh_translation(tmp_path):
    d = tmp_path / "gas-burner-aver1"
    d.mkdir()
    (d / "scene.json").write_text('{"meshTransform": {"translationMeters": [0.0, -0.5, 0.0]},'
                                  ' "emitter": {"centerMeters": [0.1, 0.3, 0.2], "radiusMeters": 0.2, "heightBandMeters": 0.1}}')
    (d / "emitter-mask.json").write_text('{"hint": {"heightBandMeters": 0.3}, "burnerCentersMeters": [[0.5, 1.0, 0.25], [0, 0, 0]],'
                                         ' "derived": {"centerXMeters": 0.4, "heightBandMeters": 0.9}}')
    p = firesim.scene_emitter_params(2, str(tmp_path))
    assert p.emitter_center_x == pytest.approx(0.4)          # mask overrides the contract
    assert p.emitter_center_z == pytest.approx(0.2)          # from the contract
    assert p.emitter_height_band_norm == pytest.approx(0.3 / 2.03, rel=1e-6)  # first heightBandMeters in the text
    assert p.emitter_radius == pytest.approx(0.2)
    assert p.burner_center_count == 2
    assert list(p.burner_center_y)[:2] == pytest.approx([0.5, -0.5])        # translated by the mesh transform
    assert p.emitter_height_norm == pytest.approx((F32(1.0) - 0.5 - 0.02) / 2.03, rel=1e-6)


@pytest.mark.skipif(not os.path.isdir(ASSETS), reason="reference scene files not present")
def test_reference_scene_files():
    camp = firesim.scene_emitter_params(1, ASSETS)
    assert camp.emitter_center_x == pytest.approx(0.011655047535896301, rel=1e-6)
    assert camp.emitter_center_z == pytest.approx(-0.00791538693010807, rel=1e-6)
    assert camp.emitter_height_norm == pytest.approx((0.02828231453895569 - 0.02) / 2.03, rel=1e-5)
    assert camp.emitter_height_band_norm == pytest.approx(0.2 / 2.03, rel=1e-6)
    assert camp.emitter_radius == pytest.approx(0.3531242311000824, rel=1e-6)
    assert camp.burner_center_count == 0
    gas = firesim.scene_emitter_params(2, ASSETS)
    assert gas.burner_center_count == 1
    assert (gas.burner_center_x[0], gas.burner_center_y[0], gas.burner_center_z[0]) == pytest.approx(
        (0.1966859996318817, 0.37620601058006287 - 0.24, 0.23663799464702606), rel=1e-5)
    assert gas.emitter_height_norm == pytest.approx((0.37620601058006287 - 0.24 - 0.02) / 2.03, rel=1e-5)
    assert gas.emitter_height_band_norm == pytest.approx(0.105 / 2.03, rel=1e-6)
    assert gas.emitter_radius == pytest.approx(0.13822224736213684, rel=1e-6)
    room = firesim.scene_emitter_params(0, ASSETS)
    assert room.emitter_height_band_norm == pytest.approx(0.12 / 2.03, rel=1e-6)


def test_image_stats_closed_forms():
    w = 64
    frame = np.full(w * w, 0xFF000000, dtype=np.uint32)           # black
    s = firesim.image_stats(frame)
    assert s["mean_luma"] == 0 and s["black_smoke_fraction"] == 1.0 and s["bright_pixels"] == 0
    frame[: w * w // 4] = 0xFFFFFFFF                                   # a quarter white
    s = firesim.image_stats(frame)
    assert s["mean_luma"] == pytest.approx(0.25, rel=1e-6) and s["max_luma"] == pytest.approx(1.0, rel=1e-6)
    assert s["bright_pixels"] == w * w // 4 and s["bright_fraction"] == pytest.approx(0.25)
    assert s["warm_fire_fraction"] == 0                                # white is not warm
    frame[:] = 0xFFFF8020                                              # r=255 g=128 b=32: warm, saturated
    s = firesim.image_stats(frame)
    luma = 1.0 * 0.2126 + 128 / 255 * 0.7152 + 32 / 255 * 0.0722
    assert s["mean_luma"] == pytest.approx(luma, rel=1e-5)
    assert s["mean_saturation"] == pytest.approx(1 - 32 / 255, rel=1e-5)
    assert s["warm_fire_fraction"] == 1.0 and s["white_core_fraction"] == 0
