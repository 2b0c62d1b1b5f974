"""ctypes binding of fire_api.h: the reference's process-wide step interface
(src/fire_cuda.h:1-119) as exported by libfiresim.so.

The functions have C++ linkage, exactly as in the reference, so they are bound
here by their Itanium-mangled names; `const FireSettings&` is passed as a
pointer, which is how the C++ ABI passes a reference. Struct layouts mirror
fire_api.h member for member (all members 4 bytes wide except the 64-bit
memory size and the name array of FireCudaDiagnostics).
"""
from __future__ import annotations

from ctypes import (POINTER, Structure, byref, c_bool, c_char, c_char_p, c_float, c_int, c_uint32, c_uint64,
                    c_void_p)

import numpy as np

import firesim


class FireSettings(Structure):
    """src/fire_cuda.h:65-101, same members and defaults."""
    _fields_ = [
        ("width", c_int), ("height", c_int), ("dt", c_float),
        ("mouseX", c_float), ("mouseY", c_float),
        ("leftDown", c_int), ("rightDown", c_int), ("reset", c_int),
        ("showGizmos", c_int), ("activeGizmo", c_int), ("sceneId", c_int),
        ("wind", c_float), ("turbulence", c_float), ("detail", c_float), ("smoke", c_float),
        ("intensity", c_float), ("cameraYaw", c_float), ("cameraPitch", c_float), ("cameraDistance", c_float),
        ("cinematicMode", c_int), ("raymarchSteps", c_int), ("emberCount", c_int), ("renderDebugMode", c_int),
        ("exposure", c_float), ("reflectionGain", c_float), ("smokeDarkness", c_float),
        ("emitterCenterX", c_float), ("emitterCenterZ", c_float), ("emitterHeightNorm", c_float),
        ("emitterHeightBandNorm", c_float), ("emitterRadius", c_float),
        ("burnerCenterCount", c_int),
        ("burnerCenterX", c_float * 4), ("burnerCenterY", c_float * 4), ("burnerCenterZ", c_float * 4),
    ]

    _defaults = dict(width=1280, height=720, dt=1.0 / 60.0, mouseX=0.5, mouseY=0.12, showGizmos=1, activeGizmo=1,
                     turbulence=0.72, detail=0.88, smoke=0.92, intensity=1.0, cameraPitch=0.18, cameraDistance=3.35,
                     raymarchSteps=56, emberCount=96, exposure=1.0, reflectionGain=1.0, smokeDarkness=1.0,
                     emitterHeightBandNorm=0.06, emitterRadius=0.48)

    def __init__(self, **kw):
        super().__init__()
        for k, v in {**self._defaults, **kw}.items():
            setattr(self, k, v)


_METRIC_INTS = ["frameIndex", "gridX", "gridY", "gridZ", "pressureIterations", "invalidCells"]
_METRIC_FLOATS = [
    "timeSeconds", "gpuSolveMs", "gpuRenderMs", "gpuVelocityMs", "gpuReactionMs", "gpuProjectionMs",
    "gpuLightingMs", "gpuRaymarchMs", "gpuPackMs", "heatSum", "fuelSum", "oxygenSum", "sootSum", "charSum",
    "ashSum", "pyrolysisSum", "progressSum", "turbulenceEnergySum", "sootOpticalDepthSum", "maxHeat", "maxFuel",
    "maxSoot", "maxPyrolysis", "maxProgress", "maxTurbulenceEnergy", "flameHeightMeters", "meanOpticalDepth",
    "meanSceneLight", "meanSceneShadow", "flameMassProxy", "smokeMassProxy", "flameSmokeOverlapProxy",
    "meanVolumetricShadow", "meanRoomIrradiance", "heatReleaseProxy", "divergenceBeforeL2", "divergenceBeforeMax",
    "divergenceAfterL2", "divergenceAfterMax", "divergenceReduction",
]


class FireCudaFrameMetrics(Structure):
    """src/fire_cuda.h:16-63."""
    _fields_ = [(n, c_int) for n in _METRIC_INTS] + [(n, c_float) for n in _METRIC_FLOATS]

    def as_dict(self) -> dict:
        return {n: getattr(self, n) for n in _METRIC_INTS + _METRIC_FLOATS}


class FireCudaDiagnostics(Structure):
    """src/fire_cuda.h:5-14."""
    _fields_ = [
        ("driverVersion", c_int), ("runtimeVersion", c_int), ("deviceCount", c_int), ("activeDevice", c_int),
        ("computeMajor", c_int), ("computeMinor", c_int), ("totalGlobalMem", c_uint64), ("deviceName", c_char * 128),
    ]


# name -> (mangled symbol, restype, argtypes)
_SETTINGS = POINTER(FireSettings)
SYMBOLS = {
    "fireCudaInitialize": ("_Z18fireCudaInitializeiiii", c_bool, [c_int, c_int, c_int, c_int]),
    "fireCudaStepAndRender": ("_Z21fireCudaStepAndRenderPjRK12FireSettings", c_bool, [c_void_p, _SETTINGS]),
    "fireCudaStepAndRenderMeasured": ("_Z29fireCudaStepAndRenderMeasuredPjRK12FireSettingsP20FireCudaFrameMetrics",
                                      c_bool, [c_void_p, _SETTINGS, POINTER(FireCudaFrameMetrics)]),
    "fireCudaSelectDeviceForD3D11": ("_Z28fireCudaSelectDeviceForD3D11Pv", c_bool, [c_void_p]),
    "fireCudaRegisterD3D11Texture": ("_Z28fireCudaRegisterD3D11TexturePv", c_bool, [c_void_p]),
    "fireCudaRegisterD3D11TextureSlot": ("_Z32fireCudaRegisterD3D11TextureSlotiPv", c_bool, [c_int, c_void_p]),
    "fireCudaSetD3D11TextureSlot": ("_Z27fireCudaSetD3D11TextureSloti", c_bool, [c_int]),
    "fireCudaStepD3D11": ("_Z17fireCudaStepD3D11RK12FireSettings", c_bool, [_SETTINGS]),
    "fireCudaStepAndRenderD3D11": ("_Z26fireCudaStepAndRenderD3D11RK12FireSettings", c_bool, [_SETTINGS]),
    "fireCudaStepAndRenderD3D11Measured": ("_Z34fireCudaStepAndRenderD3D11MeasuredRK12FireSettingsP20FireCudaFrameMetrics",
                                           c_bool, [_SETTINGS, POINTER(FireCudaFrameMetrics)]),
    "fireCudaRenderD3D11": ("_Z19fireCudaRenderD3D11RK12FireSettings", c_bool, [_SETTINGS]),
    "fireCudaUnregisterD3D11Texture": ("_Z30fireCudaUnregisterD3D11Texturev", None, []),
    "fireCudaSynchronize": ("_Z19fireCudaSynchronizev", c_bool, []),
    "fireCudaReset": ("_Z13fireCudaResetv", c_bool, []),
    "fireCudaGetDiagnostics": ("_Z22fireCudaGetDiagnosticsP19FireCudaDiagnostics", c_bool,
                               [POINTER(FireCudaDiagnostics)]),
    "fireCudaShutdown": ("_Z16fireCudaShutdownv", None, []),
    "fireCudaLastError": ("_Z17fireCudaLastErrorv", c_char_p, []),
    "fireCudaScene": ("_Z13fireCudaScenev", c_void_p, []),
}


class FireApi:
    """The bound functions, callable by their reference names."""

    def __init__(self, lib=None):
        self.lib = lib if lib is not None else firesim.load_library()
        for name, (sym, res, args) in SYMBOLS.items():
            fn = getattr(self.lib, sym)
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)

    def last_error(self) -> str:
        return self.fireCudaLastError().decode()

    def step_and_render(self, frame: np.ndarray, settings: FireSettings, metrics: FireCudaFrameMetrics | None = None):
        assert frame.dtype == np.uint32 and frame.flags.c_contiguous
        ptr = frame.ctypes.data_as(c_void_p)
        if metrics is None:
            return bool(self.fireCudaStepAndRender(ptr, byref(settings)))
        return bool(self.fireCudaStepAndRenderMeasured(ptr, byref(settings), byref(metrics)))


def reference_grid(grid_width: int, grid_height: int) -> tuple[int, int, int]:
    """Grid derived by fireCudaInitialize from the requested size (src/fire_cuda.cu:2916-2919)."""
    hq = grid_width >= 320 or grid_height >= 220
    clamp = lambda v, lo, hi: max(lo, min(hi, v))
    return (clamp(grid_width // 2, 72, 176 if hq else 104), clamp(grid_height - 32, 80, 208 if hq else 112),
            clamp(grid_width // 3, 48, 132 if hq else 72))


def settings_to_params(s: FireSettings) -> firesim.Params:
    """The per-call clamps of the reference (src/fire_cuda.cu:3135-3179) as fs_params, pressure solve at the
    reference constants (omega 1.52, 40 iterations, :104-105)."""
    c = lambda v, lo, hi: max(lo, min(hi, v))
    p = firesim.default_params()
    p.dt = c(s.dt, 0.001, 1.0 / 30.0)
    p.mouse_x, p.mouse_y = c(s.mouseX, 0.0, 1.0), c(s.mouseY, 0.0, 1.0)
    p.left_down, p.right_down = s.leftDown, s.rightDown
    p.scene_id = c(s.sceneId, 0, 2)
    p.wind_x, p.wind_z = c(s.wind, -1.0, 1.0), 0.0
    p.turbulence = c(s.turbulence, 0.05, 1.8)
    p.detail, p.smoke, p.intensity = c(s.detail, 0.1, 1.5), c(s.smoke, 0.0, 1.6), c(s.intensity, 0.2, 2.4)
    p.camera_yaw, p.camera_pitch = s.cameraYaw, c(s.cameraPitch, -0.75, 0.75)
    p.camera_distance = c(s.cameraDistance, 1.5, 6.0)
    p.cinematic_mode = 1 if s.cinematicMode else 0
    p.raymarch_steps = c(s.raymarchSteps, 24, 120)
    p.ember_count = c(s.emberCount, 0, 192)
    p.render_debug_mode = c(s.renderDebugMode, 0, 6)
    p.show_gizmos = 1 if s.showGizmos != 0 else 0
    p.active_gizmo = s.activeGizmo
    p.exposure, p.reflection_gain = c(s.exposure, 0.45, 2.4), c(s.reflectionGain, 0.0, 2.2)
    p.smoke_darkness = c(s.smokeDarkness, 0.35, 2.4)
    p.emitter_center_x, p.emitter_center_z = s.emitterCenterX, s.emitterCenterZ
    p.emitter_height_norm, p.emitter_height_band_norm = s.emitterHeightNorm, s.emitterHeightBandNorm
    p.emitter_radius, p.burner_center_count = s.emitterRadius, s.burnerCenterCount
    for i in range(4):
        p.burner_center_x[i], p.burner_center_y[i], p.burner_center_z[i] = \
            s.burnerCenterX[i], s.burnerCenterY[i], s.burnerCenterZ[i]
    p.sor_omega, p.sor_iterations = 1.52, 40
    return p
