"""The reference's process-wide interface (fire_api.h, src/fire_cuda.h:1-119)
on libfiresim.so: exports, error behaviour, grid derivation and clamps, and on
the GPU that a frame through it equals the same frame through the per-scene C
ABI, bit for bit, and the reference smoke sequence (src/main.cpp:4170-4218)."""
import json
import os
import re
import struct
import subprocess
from ctypes import byref

import numpy as np
import pytest

import fire_api
import firesim
from fire_api import FireApi, FireCudaDiagnostics, FireCudaFrameMetrics, FireSettings

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(os.path.dirname(HERE), "native-firesim_amd")
API = FireApi()


def test_every_declared_function_is_exported():
    text = open(os.path.join(PKG, "fire_api.h")).read()
    declared = re.findall(r"^(?:bool|void|const char\*|fs_scene\*)\s+(fireCuda\w+)\(", text, re.M)
    assert len(declared) == 18
    assert sorted(declared) == sorted(fire_api.SYMBOLS)


def test_struct_layouts_match_the_header():
    # FireSettings: 32 four-byte members + three float[4]; metrics: 6 ints + 40 floats
    assert len(bytes(FireSettings())) == (32 + 12) * 4
    assert len(bytes(FireCudaFrameMetrics())) == 46 * 4
    assert len(bytes(FireCudaDiagnostics())) == 6 * 4 + 8 + 128


def test_calls_before_initialize_fail_with_the_reference_message():
    API.fireCudaShutdown()
    frame = np.zeros(4, np.uint32)
    assert not API.step_and_render(frame, FireSettings())
    assert API.last_error() == "Renderer is not initialized."
    assert not API.fireCudaSynchronize()
    assert not API.fireCudaReset()
    assert API.fireCudaScene() is None


def test_bad_arguments_are_rejected_before_any_device_work():
    assert not API.fireCudaInitialize(0, 540, 384, 240)
    assert API.last_error() == "Frame dimensions must be positive."
    assert not API.fireCudaStepAndRenderMeasured(None, byref(FireSettings()), None)
    assert API.last_error() == "Metrics output pointer was null."
    assert not API.fireCudaGetDiagnostics(None)
    assert not API.fireCudaRegisterD3D11Texture(None)
    assert "D3D11" in API.last_error()
    API.fireCudaUnregisterD3D11Texture()


@pytest.mark.parametrize("req,grid", [((384, 240), (176, 208, 128)), ((200, 140), (100, 108, 66)),
                                      ((40, 40), (72, 80, 48)), ((1000, 1000), (176, 208, 132))])
def test_reference_grid_derivation(req, grid):
    assert fire_api.reference_grid(*req) == grid


def test_settings_clamps_match_the_reference():
    s = FireSettings(turbulence=0.0, raymarchSteps=500, sceneId=9, wind=-3.0, dt=1.0, exposure=9.0)
    p = fire_api.settings_to_params(s)
    clamped = firesim.Params()
    assert firesim.load_library().fs_clamp_params(byref(p), byref(clamped)) == 0
    assert clamped.turbulence == pytest.approx(0.05)   # reference floor (:3152)
    assert clamped.raymarch_steps == 120               # kMaxRaymarchSteps
    assert clamped.scene_id == 2 and clamped.wind_x == -1.0
    assert clamped.dt == pytest.approx(1 / 30) and clamped.exposure == pytest.approx(2.4)
    assert clamped.sor_iterations == 40 and clamped.sor_omega == pytest.approx(1.52)


def _frames(n):
    """Per-frame settings of the reference smoke sequence, shortened."""
    out = []
    for i in range(n):
        out.append(FireSettings(width=96, height=64, mouseX=0.5 + 0.08 * np.sin(i * 0.08), mouseY=0.13,
                                wind=0.08, detail=0.92, cinematicMode=1, raymarchSteps=48, turbulence=1.34,
                                intensity=1.42, smoke=1.05, cameraYaw=0.18, cameraPitch=0.08, cameraDistance=2.62,
                                sceneId=1 if i < 3 else 2, reset=1 if i == 3 else 0))
    return out


@pytest.mark.gpu
def test_frames_equal_the_per_scene_abi(hip_lib):
    """fireCudaStepAndRender == fs_set_params + fs_step + fs_render_bgra8 with the clamped settings, including
    a scene switch and an in-call reset."""
    assert API.fireCudaInitialize(96, 64, 200, 140), API.last_error()
    try:
        frames = _frames(6)
        a = np.zeros(96 * 64, np.uint32)
        b = np.zeros((64, 96), np.uint32)
        scene = firesim.Scene(*fire_api.reference_grid(200, 140), frame_w=96, frame_h=64)
        scene.reset()
        for i, s in enumerate(frames):
            m = FireCudaFrameMetrics()
            assert API.step_and_render(a, s, m if i == 5 else None), API.last_error()
            scene.set_params(fire_api.settings_to_params(s))
            if s.reset:
                # the reference steps the reset frame at the pre-reset clock, then zeroes the clocks
                t0, rt0, _ = scene.get_time()
                scene.reset()
                scene.set_time(t0, rt0)
            scene.step(1)
            b[:] = scene.render_bgra8()
            if s.reset:
                scene.set_time(0.0, 0.0)
            assert np.array_equal(a.reshape(64, 96), b), f"frame {i}"
        ref = scene.metrics()
        api_scene = firesim.Scene.attach(API.fireCudaScene())
        assert (api_scene.nx, api_scene.ny, api_scene.nz, api_scene.frame_w) == (100, 108, 66, 96)
        assert (m.gridX, m.gridY, m.gridZ) == (100, 108, 66)
        assert m.pressureIterations == 40 and m.invalidCells == 0
        assert m.frameIndex == 2  # counted from the reset at frame 3
        assert m.timeSeconds == pytest.approx(2 / 60, rel=1e-6)  # clocks zeroed by the reset frame
        assert m.heatSum == pytest.approx(ref["heat_sum"], rel=1e-6)
        assert m.divergenceAfterL2 == pytest.approx(ref["divergence_after_l2"], rel=1e-5)
        assert m.gpuSolveMs > 0 and m.gpuProjectionMs > 0
        for name in ("heat", "soot", "u", "pressure"):
            assert np.array_equal(api_scene.read(name), scene.read(name)), name
        api_scene.close()  # a borrowed handle: stays alive
        assert API.fireCudaSynchronize()
        scene.close()
        d = FireCudaDiagnostics()
        assert API.fireCudaGetDiagnostics(byref(d)) and d.deviceCount >= 1 and d.totalGlobalMem > 0
        assert API.fireCudaSynchronize()
    finally:
        API.fireCudaShutdown()


@pytest.mark.gpu
def test_reference_interface_frame_matches_the_oracle(hip_lib):
    """A frame through fireCudaStepAndRender against the FP64 oracle: after
    three frames through the interface (campfire bed, mouse source, gizmos on by
    default), the oracle takes the interface's state and clock, the clamped
    settings, and steps and renders once beside it. Every field of the step is
    held to relative L2 1e-5 and the BGRA frame (raymarch, tonemap and the
    gizmo overlay) to the render's FP32 rounding: at most 3 LSB per channel,
    and 99% of channels within 1 LSB."""
    import compare
    import oracle

    assert API.fireCudaInitialize(96, 64, 200, 140), API.last_error()
    try:
        frames = _frames(3)
        a = np.zeros(96 * 64, np.uint32)
        for s in frames:
            assert API.step_and_render(a, s, None), API.last_error()
        api_scene = firesim.Scene.attach(API.fireCudaScene())
        state = {n: api_scene.read(n) for n in firesim.STATE_FIELDS}
        t, rt, _ = api_scene.get_time()
        s = frames[2]
        assert s.showGizmos == 1 and s.reset == 0
        o = oracle.OracleScene(*fire_api.reference_grid(200, 140), 96, 64)
        o.set_params(fire_api.settings_to_params(s))
        o.reset()
        for n, arr in state.items():
            o.write(n, arr.astype(np.float64))
        o.set_time(t, rt)
        assert API.step_and_render(a, s, None), API.last_error()
        o.step(1)
        errs = {n: compare.rel_l2(api_scene.read(n), o.read(n)) for n in compare.STEP_FIELDS}
        print("fire_api step vs oracle", errs, flush=True)
        assert all(v <= 1e-5 for v in errs.values()), errs
        fo = o.render_bgra8()
        ca = np.stack([(a.reshape(64, 96) >> k) & 0xFF for k in (0, 8, 16, 24)], -1).astype(int)
        co = np.stack([(fo >> k) & 0xFF for k in (0, 8, 16, 24)], -1).astype(int)
        d = np.abs(ca - co)
        print("fire_api frame vs oracle: max LSB", d.max(), "within 1 LSB", (d <= 1).mean(), flush=True)
        assert d.max() <= 3 and (d <= 1).mean() >= 0.99
        api_scene.close()
        o.close()
    finally:
        API.fireCudaShutdown()


@pytest.mark.gpu
def test_reference_smoke_sequence(hip_lib, tmp_path):
    """runCudaSmokeTest: 960x540 frame, 384x240 grid request, 64 frames, BMP of the last one."""
    exe = os.path.join(PKG, "fire_smoke")
    out = str(tmp_path / "frame.bmp")
    r = subprocess.run([exe, out], capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stderr
    m = json.loads(r.stdout.strip().splitlines()[-1])
    assert m["grid"] == [176, 208, 128] and m["invalidCells"] == 0
    assert m["maxHeat"] > 0.1 and m["flameHeightMeters"] > 0
    data = open(out, "rb").read()
    assert struct.unpack_from("<HI", data, 0) == (0x4D42, 54 + 960 * 540 * 4)
    pix = np.frombuffer(data, np.uint32, offset=54)
    assert (pix >> 24 == 255).all() and len(np.unique(pix)) > 1000
