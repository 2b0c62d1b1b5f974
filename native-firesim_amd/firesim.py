"""ctypes bindings of include/firesim.h (the C ABI of libfiresim.so).

The library is the MI355X HIP implementation of the reference's per-frame
step (src/fire_cuda.h:103-119 is the interface it replaces). Loading fails
loudly when the built library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, Structure, byref, c_char_p, c_double, c_float, c_int32, c_int64, c_uint16, c_uint32, c_uint64,
                    c_void_p)

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfiresim.so")

FS_OK = 0

# enum fs_field_id
FIELD_NAMES = [
    "heat", "fuel", "oxygen", "soot", "char", "ash", "pyrolysis", "progress",
    "turbulence", "soot_optics", "u", "v", "w", "pressure", "divergence", "light", "shadow",
    "dir_transmittance",
]
FIELD = {name: i for i, name in enumerate(FIELD_NAMES)}
SCALAR_FIELDS = FIELD_NAMES[:10]
# the simulation state and its light volumes; dir_transmittance is written only while dir_light is on
STATE_FIELDS = FIELD_NAMES[:17]

# enum fs_stage
STAGE_ADVECT_VELOCITY, STAGE_ADVECT_REACT, STAGE_FORCE, STAGE_PRESSURE, STAGE_PROJECT, STAGE_LIGHT = range(6)
STAGE_NAMES = ["advect_velocity", "advect_react", "force", "pressure", "project", "light"]

STORAGE_FP32, STORAGE_BF16 = 0, 1
SOR_KERNELS = ["half_sweep", "tb_t16_k4", "tb_t24_k4", "tb_t24_z8_k4", "tb_t24_z16_k4", "tb_t24_z28_k4", "tb_t24_z10_k4", "tb_t24_z12_k4",
               "tb_t24_z6_k4", "tb_t26_z13_k3", "tb_t26_z9_k3",
               "tb_t22_z8_k5"]
SCENE_ROOM_TRAY, SCENE_CAMPFIRE, SCENE_GAS_BURNER, SCENE_EMITTERS_ONLY = 0, 1, 2, 3
MAX_EMITTERS = 16


class SceneDesc(Structure):
    _fields_ = [
        ("nx", c_int32), ("ny", c_int32), ("nz", c_int32),
        ("frame_w", c_int32), ("frame_h", c_int32),
        ("storage", c_int32), ("device", c_int32),
    ]


class Params(Structure):
    _fields_ = [
        ("dt", c_float),
        ("mouse_x", c_float), ("mouse_y", c_float),
        ("left_down", c_int32), ("right_down", c_int32),
        ("scene_id", c_int32),
        ("wind_x", c_float), ("wind_z", c_float),
        ("turbulence", c_float), ("detail", c_float), ("smoke", c_float), ("intensity", c_float),
        ("camera_yaw", c_float), ("camera_pitch", c_float), ("camera_distance", c_float),
        ("cinematic_mode", c_int32), ("raymarch_steps", c_int32), ("render_debug_mode", c_int32),
        ("exposure", c_float), ("reflection_gain", c_float), ("smoke_darkness", c_float),
        ("emitter_center_x", c_float), ("emitter_center_z", c_float),
        ("emitter_height_norm", c_float), ("emitter_height_band_norm", c_float), ("emitter_radius", c_float),
        ("burner_center_count", c_int32),
        ("burner_center_x", c_float * 4), ("burner_center_y", c_float * 4), ("burner_center_z", c_float * 4),
        ("sor_omega", c_float), ("sor_iterations", c_int32),
        ("noise_seed", c_uint32),
        ("ember_count", c_int32),
        ("closed_top", c_int32), ("noise_forcing", c_int32), ("curl_noise", c_float), ("noise_mode", c_int32),
        ("show_gizmos", c_int32), ("active_gizmo", c_int32),
        ("dir_light", c_int32), ("dir_light_x", c_float), ("dir_light_y", c_float), ("dir_light_z", c_float),
        ("dir_light_r", c_float), ("dir_light_g", c_float), ("dir_light_b", c_float),
    ]


class Emitter(Structure):
    _fields_ = [
        ("cx", c_float), ("cy", c_float), ("cz", c_float), ("radius", c_float),
        ("fuel_rate", c_float), ("temperature_k", c_float),
    ]


_METRIC_DOUBLES = [
    "time_seconds",
    "heat_sum", "fuel_sum", "oxygen_sum", "soot_sum", "char_sum", "ash_sum",
    "pyrolysis_sum", "progress_sum", "turbulence_energy_sum", "soot_optical_depth_sum",
    "max_heat", "max_fuel", "max_soot", "max_pyrolysis", "max_progress", "max_turbulence_energy",
    "flame_height_meters",
    "mean_optical_depth", "mean_scene_light", "mean_scene_shadow",
    "flame_mass_proxy", "smoke_mass_proxy", "flame_smoke_overlap_proxy",
    "mean_volumetric_shadow", "mean_room_irradiance", "heat_release_proxy",
    "divergence_before_l2", "divergence_before_max",
    "divergence_after_l2", "divergence_after_max", "divergence_reduction",
    "heat_release_step", "heat_release_total", "sor_residual_l2",
]


class Metrics(Structure):
    _fields_ = [
        ("frame_index", c_int32),
        ("grid_x", c_int32), ("grid_y", c_int32), ("grid_z", c_int32),
        ("pressure_iterations", c_int32), ("invalid_cells", c_int32),
    ] + [(name, c_double) for name in _METRIC_DOUBLES]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServerDesc(Structure):
    _fields_ = [("device", c_int32), ("max_active", c_int32), ("hbm_budget_bytes", c_uint64),
                ("sor_kernel", c_int32), ("use_graph", c_int32), ("max_round_voxels", c_uint64)]


class Request(Structure):
    _fields_ = [("desc", SceneDesc), ("params", Params), ("emitters", Emitter * MAX_EMITTERS), ("emitter_count", c_int32),
                ("steps", c_int32), ("render_every", c_int32), ("keep_scene", c_int32), ("arrival_s", c_double)]


class RequestStatus(Structure):
    _fields_ = [("state", c_int32), ("steps_done", c_int32), ("frames", c_int32), ("frame_w", c_int32), ("frame_h", c_int32),
                ("arrival_s", c_double), ("admitted_s", c_double), ("finished_s", c_double),
                ("step_ms_p50", c_double), ("step_ms_p99", c_double), ("step_ms_max", c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServerStats(Structure):
    _fields_ = [("requests_submitted", c_int64), ("requests_done", c_int64), ("requests_failed", c_int64),
                ("steps", c_int64), ("voxel_steps", c_int64), ("rounds", c_int64), ("frames", c_int64),
                ("max_active_seen", c_int32), ("sor_kernel", c_int32), ("peak_resident_bytes", c_uint64),
                ("busy_s", c_double), ("step_ms_p50", c_double), ("step_ms_p99", c_double), ("step_ms_mean", c_double),
                ("step_ms_max", c_double), ("queue_ms_p50", c_double), ("queue_ms_p99", c_double)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


REQ_PENDING, REQ_WAITING, REQ_RUNNING, REQ_DONE, REQ_FAILED = range(5)


class ImageStats(Structure):
    _fields_ = [("mean_luma", c_float), ("max_luma", c_float), ("mean_saturation", c_float),
                ("warm_fire_fraction", c_float), ("tan_smoke_fraction", c_float), ("black_smoke_fraction", c_float),
                ("white_core_fraction", c_float), ("bright_fraction", c_float), ("bright_pixels", c_int32)]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Tuning(Structure):
    _fields_ = [("sor_kernel", c_int32), ("use_graph", c_int32)]


class AutotuneReport(Structure):
    _fields_ = [("sor_ms", c_double * len(SOR_KERNELS)), ("chosen", c_int32)]


class StageTimes(Structure):
    _fields_ = [("stage_ms", c_double * 6), ("total_ms", c_double), ("steps", c_int32), ("pressure_launches", c_int32)]


def default_params(**overrides) -> Params:
    """FireSettings defaults (src/fire_cuda.h:65-101) plus the SOR constants
    (src/fire_cuda.cu:104-105); keyword overrides set fields by name."""
    p = Params()
    p.dt = 1.0 / 60.0
    p.mouse_x, p.mouse_y = 0.5, 0.12
    p.left_down = p.right_down = 0
    p.scene_id = 0
    p.wind_x = p.wind_z = 0.0
    p.turbulence, p.detail, p.smoke, p.intensity = 0.72, 0.88, 0.92, 1.0
    p.camera_yaw, p.camera_pitch, p.camera_distance = 0.0, 0.18, 3.35
    p.cinematic_mode, p.raymarch_steps, p.render_debug_mode = 0, 56, 0
    p.exposure, p.reflection_gain, p.smoke_darkness = 1.0, 1.0, 1.0
    p.emitter_center_x = p.emitter_center_z = 0.0
    p.emitter_height_norm, p.emitter_height_band_norm, p.emitter_radius = 0.0, 0.06, 0.48
    p.burner_center_count = 0
    p.sor_omega, p.sor_iterations = 1.52, 40
    p.noise_seed = 0
    p.ember_count = 96
    for k, v in overrides.items():
        if k in ("burner_center_x", "burner_center_y", "burner_center_z"):
            arr = getattr(p, k)
            for i, val in enumerate(v):
                arr[i] = val
        else:
            if not hasattr(p, k):
                raise AttributeError(f"fs_params has no field {k!r}")
            setattr(p, k, v)
    return p


def params_to_dict(p: Params) -> dict:
    out = {}
    for name, _ in p._fields_:
        val = getattr(p, name)
        out[name] = list(val) if name.startswith("burner_center_") and name != "burner_center_count" else val
    return out


def light_dims(nx: int, ny: int, nz: int) -> tuple[int, int, int]:
    """clamp((n+3)/4, lo, hi), src/fire_cuda.cu:2920-2922."""
    c = lambda n, lo, hi: max(lo, min(hi, (n + 3) // 4))  # noqa: E731
    return c(nx, 18, 48), c(ny, 20, 56), c(nz, 16, 36)


def copy_bandwidth(nbytes: int = 1 << 30, reps: int = 20, device: int = 0) -> float:
    """Achievable device-memory bandwidth in GB/s (read + write bytes of
    streaming float4 copies, fs_copy_bandwidth)."""
    lib = load_library()
    out = c_double()
    if lib.fs_copy_bandwidth(device, nbytes, reps, byref(out)) != 0:
        raise FireSimError(lib.fs_last_global_error().decode())
    return out.value


def field_shape(field: int, nx: int, ny: int, nz: int) -> tuple:
    """Numpy shape (z-major, x fastest) of a field (id or name)."""
    if isinstance(field, str):
        field = FIELD[field]
    if field < 10 or field in (FIELD["pressure"], FIELD["divergence"]):
        return (nz, ny, nx)
    if field == FIELD["u"]:
        return (nz, ny, nx + 1)
    if field == FIELD["v"]:
        return (nz, ny + 1, nx)
    if field == FIELD["w"]:
        return (nz + 1, ny, nx)
    lnx, lny, lnz = light_dims(nx, ny, nz)
    if field == FIELD["light"]:
        return (lnz, lny, lnx, 4)
    if field in (FIELD["shadow"], FIELD["dir_transmittance"]):
        return (lnz, lny, lnx)
    raise ValueError(field)


class FireSimError(RuntimeError):
    pass


_lib = None


def load_library(path: str = LIB_PATH):
    """Loads libfiresim.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FireSimError(f"{path} is not built; run __graft_entry__.build() first")
    lib = ctypes.CDLL(path)
    P = POINTER
    sigs = {
        "fs_abi_version": (c_int32, []),
        "fs_default_params": (None, [P(Params)]),
        "fs_clamp_params": (c_int32, [P(Params), P(Params)]),
        "fs_light_dims": (c_int32, [c_int32, c_int32, c_int32, P(c_int32), P(c_int32), P(c_int32)]),
        "fs_scene_bytes": (c_int32, [P(SceneDesc), P(c_uint64)]),
        "fs_write_bmp": (c_int32, [c_char_p, P(c_uint32), c_int32, c_int32]),
        "fs_last_global_error": (c_char_p, []),
        "fs_copy_bandwidth": (c_int32, [c_int32, c_uint64, c_int32, P(c_double)]),
        "fs_create": (c_int32, [P(SceneDesc), P(c_void_p)]),
        "fs_destroy": (c_int32, [c_void_p]),
        "fs_reset": (c_int32, [c_void_p]),
        "fs_set_params": (c_int32, [c_void_p, P(Params)]),
        "fs_get_params": (c_int32, [c_void_p, P(Params)]),
        "fs_get_desc": (c_int32, [c_void_p, P(SceneDesc)]),
        "fs_set_emitters": (c_int32, [c_void_p, P(Emitter), c_int32]),
        "fs_last_error": (c_char_p, [c_void_p]),
        "fs_step": (c_int32, [c_void_p, c_int32]),
        "fs_run_stage": (c_int32, [c_void_p, c_int32]),
        "fs_step_timed": (c_int32, [c_void_p, c_int32, P(StageTimes), P(c_double)]),
        "fs_set_tuning": (c_int32, [c_void_p, P(Tuning)]),
        "fs_get_tuning": (c_int32, [c_void_p, P(Tuning)]),
        "fs_step_latency": (c_int32, [c_void_p, c_int32, P(c_double)]),
        "fs_autotune": (c_int32, [c_void_p, c_int32, P(AutotuneReport)]),
        "fs_synchronize": (c_int32, [c_void_p]),
        "fs_set_time": (c_int32, [c_void_p, c_float, c_float]),
        "fs_get_time": (c_int32, [c_void_p, P(c_float), P(c_float), P(c_int32)]),
        "fs_render_hdr": (c_int32, [c_void_p, P(c_float)]),
        "fs_render_bgra8": (c_int32, [c_void_p, P(c_uint32)]),
        "fs_tonemap": (c_int32, [c_void_p, P(c_float), P(c_uint32)]),
        "fs_render_rgba16f": (c_int32, [c_void_p, P(c_uint16)]),
        "fs_field_count": (c_int32, [c_void_p, c_int32, P(c_uint64)]),
        "fs_read_field": (c_int32, [c_void_p, c_int32, P(c_float)]),
        "fs_write_field": (c_int32, [c_void_p, c_int32, P(c_float)]),
        "fs_metrics_collect": (c_int32, [c_void_p, P(Metrics)]),
        "fs_stream": (c_void_p, [c_void_p]),
        "fs_batch_create": (c_int32, [c_int32, P(c_void_p)]),
        "fs_batch_destroy": (c_int32, [c_void_p]),
        "fs_batch_add": (c_int32, [c_void_p, c_void_p]),
        "fs_batch_remove": (c_int32, [c_void_p, c_void_p]),
        "fs_batch_count": (c_int32, [c_void_p, P(c_int32)]),
        "fs_batch_set_tuning": (c_int32, [c_void_p, P(Tuning)]),
        "fs_batch_get_tuning": (c_int32, [c_void_p, P(Tuning)]),
        "fs_batch_autotune": (c_int32, [c_void_p, c_int32, P(AutotuneReport)]),
        "fs_batch_step": (c_int32, [c_void_p, c_int32]),
        "fs_batch_step_latency": (c_int32, [c_void_p, c_int32, P(c_double)]),
        "fs_batch_synchronize": (c_int32, [c_void_p]),
        "fs_batch_last_error": (c_char_p, [c_void_p]),
        "fs_batch_stream": (c_void_p, [c_void_p]),
        "fs_create_in": (c_int32, [P(SceneDesc), c_void_p, c_uint64, P(c_void_p)]),
        "fs_scene_emitter_params": (c_int32, [c_int32, c_char_p, P(Params)]),
        "fs_image_stats_of": (c_int32, [P(c_uint32), c_int64, P(ImageStats)]),
        "fs_render_bgra8_async": (c_int32, [c_void_p, P(c_uint32)]),
        "fs_server_create": (c_int32, [P(ServerDesc), P(c_void_p)]),
        "fs_server_destroy": (c_int32, [c_void_p]),
        "fs_server_submit": (c_int32, [c_void_p, P(Request), P(c_uint64)]),
        "fs_server_tick": (c_int32, [c_void_p, P(c_int32)]),
        "fs_server_run": (c_int32, [c_void_p]),
        "fs_server_status": (c_int32, [c_void_p, c_uint64, P(RequestStatus)]),
        "fs_server_frame": (c_int32, [c_void_p, c_uint64, P(c_uint32), P(c_int32)]),
        "fs_server_take_scene": (c_int32, [c_void_p, c_uint64, P(c_void_p)]),
        "fs_server_stats_get": (c_int32, [c_void_p, P(ServerStats)]),
        "fs_server_step_latencies": (c_int32, [c_void_p, P(c_double), c_int64, P(c_int64)]),
        "fs_server_clock": (c_double, [c_void_p]),
        "fs_server_last_error": (c_char_p, [c_void_p]),
        "fs_pressure_solve_noalloc": (c_int32, [c_void_p, c_void_p, c_int32, c_int32, c_int32, c_float, c_int32, c_int32, c_int32,
                                                c_void_p]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def exported_symbols() -> list:
    """Names declared in include/firesim.h, in declaration order."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "firesim.h")
    text = open(hdr).read()
    return re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?\s*(fs_[a-z0-9_]+)\s*\(", text, re.M)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32)


class Scene:
    """One simulation handle (fs_scene) with numpy field access."""

    def __init__(self, nx: int, ny: int, nz: int, frame_w: int = 64, frame_h: int = 64,
                 storage: int = STORAGE_FP32, device: int = 0):
        self.lib = load_library()
        self.nx, self.ny, self.nz = nx, ny, nz
        self.frame_w, self.frame_h = frame_w, frame_h
        desc = SceneDesc(nx, ny, nz, frame_w, frame_h, storage, device)
        h = c_void_p()
        st = self.lib.fs_create(byref(desc), byref(h))
        if st != FS_OK:
            raise FireSimError(f"fs_create failed ({st}): {self.lib.fs_last_global_error().decode()}")
        self.h = h

    @classmethod
    def attach(cls, handle) -> "Scene":
        """Non-owning view of an existing fs_scene handle (for example the one
        behind the process-wide fireCuda* interface); close() leaves it alive."""
        if not handle:
            raise FireSimError("attach: null scene handle")
        self = cls.__new__(cls)
        self.lib = load_library()
        self.h = c_void_p(handle)
        self._owned = False
        desc = SceneDesc()
        self._check(self.lib.fs_get_desc(self.h, byref(desc)), "fs_get_desc")
        self.nx, self.ny, self.nz = desc.nx, desc.ny, desc.nz
        self.frame_w, self.frame_h = desc.frame_w, desc.frame_h
        return self

    def _check(self, st: int, what: str):
        if st != FS_OK:
            msg = self.lib.fs_last_error(self.h)
            raise FireSimError(f"{what} failed ({st}): {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            if getattr(self, "_owned", True):
                self.lib.fs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def reset(self):
        self._check(self.lib.fs_reset(self.h), "fs_reset")

    def set_params(self, p: Params):
        self._check(self.lib.fs_set_params(self.h, byref(p)), "fs_set_params")

    def get_params(self) -> Params:
        p = Params()
        self._check(self.lib.fs_get_params(self.h, byref(p)), "fs_get_params")
        return p

    def set_emitters(self, emitters):
        arr = (Emitter * max(1, len(emitters)))(*[Emitter(*e) if not isinstance(e, Emitter) else e for e in emitters])
        self._check(self.lib.fs_set_emitters(self.h, arr, len(emitters)), "fs_set_emitters")

    def step(self, n: int = 1):
        self._check(self.lib.fs_step(self.h, n), "fs_step")

    def tuning(self) -> Tuning:
        t = Tuning()
        self._check(self.lib.fs_get_tuning(self.h, byref(t)), "fs_get_tuning")
        return t

    def set_tuning(self, sor_kernel=None, use_graph=None):
        t = self.tuning()
        if sor_kernel is not None:
            t.sor_kernel = sor_kernel
        if use_graph is not None:
            t.use_graph = int(bool(use_graph))
        self._check(self.lib.fs_set_tuning(self.h, byref(t)), "fs_set_tuning")

    def set_sor_kernel(self, kernel: int):
        """Selects the pressure-solve kernel (enum fs_sor_kernel); results are bit-identical."""
        self.set_tuning(sor_kernel=kernel)

    def sor_kernel(self) -> int:
        return self.tuning().sor_kernel

    def autotune(self, reps: int = 2) -> dict:
        """Times every SOR kernel on this scene and keeps the fastest."""
        r = AutotuneReport()
        self._check(self.lib.fs_autotune(self.h, reps, byref(r)), "fs_autotune")
        return {"sor_ms": {SOR_KERNELS[i]: r.sor_ms[i] for i in range(len(SOR_KERNELS))}, "chosen": SOR_KERNELS[r.chosen]}

    def step_latency(self, n: int) -> np.ndarray:
        """Runs n steps on the production path; returns per-step GPU ms."""
        out = np.zeros(max(1, n), dtype=np.float64)
        self._check(self.lib.fs_step_latency(self.h, n, out.ctypes.data_as(POINTER(c_double))), "fs_step_latency")
        return out[:n]

    def step_timed(self, n: int):
        """Runs n steps with per-pass HIP event timing; returns (StageTimes, per-step ms)."""
        st = StageTimes()
        per = np.zeros(max(1, n), dtype=np.float64)
        self._check(self.lib.fs_step_timed(self.h, n, byref(st), per.ctypes.data_as(POINTER(c_double))), "fs_step_timed")
        return st, per[:n]

    def run_stage(self, stage: int):
        self._check(self.lib.fs_run_stage(self.h, stage), "fs_run_stage")

    def synchronize(self):
        self._check(self.lib.fs_synchronize(self.h), "fs_synchronize")

    def set_time(self, t: float, render_t: float):
        self._check(self.lib.fs_set_time(self.h, t, render_t), "fs_set_time")

    def get_time(self):
        t, rt, fi = c_float(), c_float(), c_int32()
        self._check(self.lib.fs_get_time(self.h, byref(t), byref(rt), byref(fi)), "fs_get_time")
        return t.value, rt.value, fi.value

    def shape(self, field: int):
        return field_shape(field, self.nx, self.ny, self.nz)

    def read(self, field) -> np.ndarray:
        f = FIELD[field] if isinstance(field, str) else field
        out = np.empty(self.shape(f), dtype=np.float32)
        self._check(self.lib.fs_read_field(self.h, f, out.ctypes.data_as(POINTER(c_float))), "fs_read_field")
        return out

    def write(self, field, data):
        f = FIELD[field] if isinstance(field, str) else field
        arr = _f32(data).reshape(self.shape(f))
        self._check(self.lib.fs_write_field(self.h, f, arr.ctypes.data_as(POINTER(c_float))), "fs_write_field")

    def render_hdr(self) -> np.ndarray:
        out = np.empty((self.frame_h, self.frame_w, 4), dtype=np.float32)
        self._check(self.lib.fs_render_hdr(self.h, out.ctypes.data_as(POINTER(c_float))), "fs_render_hdr")
        return out

    def render_bgra8(self) -> np.ndarray:
        out = np.empty((self.frame_h, self.frame_w), dtype=np.uint32)
        self._check(self.lib.fs_render_bgra8(self.h, out.ctypes.data_as(POINTER(c_uint32))), "fs_render_bgra8")
        return out

    def render_rgba16f(self) -> np.ndarray:
        """The HDR frame packed as FP16 RGBA (alpha 1), as a float16 array."""
        out = np.empty((self.frame_h, self.frame_w, 4), dtype=np.uint16)
        self._check(self.lib.fs_render_rgba16f(self.h, out.ctypes.data_as(POINTER(c_uint16))), "fs_render_rgba16f")
        return out.view(np.float16)

    def tonemap(self, hdr: np.ndarray) -> np.ndarray:
        src = _f32(hdr).reshape(self.frame_h, self.frame_w, 4)
        out = np.empty((self.frame_h, self.frame_w), dtype=np.uint32)
        self._check(self.lib.fs_tonemap(self.h, src.ctypes.data_as(POINTER(c_float)), out.ctypes.data_as(POINTER(c_uint32))),
                    "fs_tonemap")
        return out

    def metrics(self) -> dict:
        m = Metrics()
        self._check(self.lib.fs_metrics_collect(self.h, byref(m)), "fs_metrics_collect")
        return m.as_dict()

    def stream(self) -> int:
        return self.lib.fs_stream(self.h)


def write_bmp(path: str, bgra: np.ndarray) -> None:
    lib = load_library()
    arr = np.ascontiguousarray(bgra, dtype=np.uint32)
    h, w = arr.shape
    st = lib.fs_write_bmp(path.encode(), arr.ctypes.data_as(POINTER(c_uint32)), w, h)
    if st != FS_OK:
        raise FireSimError(f"fs_write_bmp failed ({st}): {lib.fs_last_global_error().decode()}")


class Batch:
    """Many scenes stepped together, one launch per pass (fs_batch)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = c_void_p()
        st = self.lib.fs_batch_create(device, byref(h))
        if st != FS_OK:
            raise FireSimError(f"fs_batch_create failed ({st}): {self.lib.fs_last_global_error().decode()}")
        self.h = h
        self.members: list = []

    def _check(self, st: int, what: str):
        if st != FS_OK:
            msg = self.lib.fs_batch_last_error(self.h)
            raise FireSimError(f"{what} failed ({st}): {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.fs_batch_destroy(self.h)
            self.h = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def add(self, scene: Scene):
        self._check(self.lib.fs_batch_add(self.h, scene.h), "fs_batch_add")
        self.members.append(scene)

    def remove(self, scene: Scene):
        self._check(self.lib.fs_batch_remove(self.h, scene.h), "fs_batch_remove")
        self.members.remove(scene)

    def count(self) -> int:
        n = c_int32()
        self._check(self.lib.fs_batch_count(self.h, byref(n)), "fs_batch_count")
        return n.value

    def tuning(self) -> Tuning:
        t = Tuning()
        self._check(self.lib.fs_batch_get_tuning(self.h, byref(t)), "fs_batch_get_tuning")
        return t

    def set_tuning(self, sor_kernel=None, use_graph=None):
        t = self.tuning()
        if sor_kernel is not None:
            t.sor_kernel = SOR_KERNELS.index(sor_kernel) if isinstance(sor_kernel, str) else sor_kernel
        if use_graph is not None:
            t.use_graph = int(bool(use_graph))
        self._check(self.lib.fs_batch_set_tuning(self.h, byref(t)), "fs_batch_set_tuning")

    def autotune(self, reps: int = 2) -> dict:
        r = AutotuneReport()
        self._check(self.lib.fs_batch_autotune(self.h, reps, byref(r)), "fs_batch_autotune")
        return {"sor_ms": {SOR_KERNELS[i]: r.sor_ms[i] for i in range(len(SOR_KERNELS))}, "chosen": SOR_KERNELS[r.chosen]}

    def step(self, n: int = 1):
        self._check(self.lib.fs_batch_step(self.h, n), "fs_batch_step")

    def step_latency(self, n: int) -> np.ndarray:
        out = np.zeros(max(1, n), dtype=np.float64)
        self._check(self.lib.fs_batch_step_latency(self.h, n, out.ctypes.data_as(POINTER(c_double))), "fs_batch_step_latency")
        return out[:n]

    def synchronize(self):
        self._check(self.lib.fs_batch_synchronize(self.h), "fs_batch_synchronize")

    def stream(self) -> int:
        return self.lib.fs_batch_stream(self.h)


def make_request(nx, ny, nz, params: Params, emitters=(), steps=60, render_every=0, frame=(256, 256),
                 storage=STORAGE_FP32, keep_scene=False, arrival_s=-1.0) -> Request:
    """An fs_request for Server.submit."""
    r = Request()
    r.desc = SceneDesc(nx, ny, nz, frame[0], frame[1], storage, 0)
    r.params = params
    for i, e in enumerate(emitters):
        r.emitters[i] = e if isinstance(e, Emitter) else Emitter(*e)
    r.emitter_count = len(emitters)
    r.steps, r.render_every, r.keep_scene, r.arrival_s = steps, render_every, int(bool(keep_scene)), arrival_s
    return r


class Server:
    """Continuous batching of simulation requests (fs_server)."""

    def __init__(self, hbm_budget_bytes: int, max_active: int = 256, sor_kernel: int = -1, use_graph: bool = True,
                 device: int = 0, max_round_voxels: int = 0):
        self.lib = load_library()
        d = ServerDesc(device, max_active, hbm_budget_bytes, sor_kernel, int(use_graph), max_round_voxels)
        h = c_void_p()
        st = self.lib.fs_server_create(byref(d), byref(h))
        if st != FS_OK:
            raise FireSimError(f"fs_server_create failed ({st}): {self.lib.fs_last_global_error().decode()}")
        self.h = h

    def _check(self, st: int, what: str):
        if st != FS_OK:
            msg = self.lib.fs_server_last_error(self.h)
            raise FireSimError(f"{what} failed ({st}): {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.fs_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, req: Request) -> int:
        rid = c_uint64()
        self._check(self.lib.fs_server_submit(self.h, byref(req), byref(rid)), "fs_server_submit")
        return rid.value

    def tick(self) -> int:
        n = c_int32()
        self._check(self.lib.fs_server_tick(self.h, byref(n)), "fs_server_tick")
        return n.value

    def run(self):
        self._check(self.lib.fs_server_run(self.h), "fs_server_run")

    def status(self, rid: int) -> dict:
        s = RequestStatus()
        self._check(self.lib.fs_server_status(self.h, rid, byref(s)), "fs_server_status")
        return s.as_dict()

    def frame(self, rid: int):
        """Latest frame of a request, sized from the request's own descriptor."""
        st = self.status(rid)
        out = np.empty((st["frame_h"], st["frame_w"]), dtype=np.uint32)
        step = c_int32()
        self._check(self.lib.fs_server_frame(self.h, rid, out.ctypes.data_as(POINTER(c_uint32)), byref(step)), "fs_server_frame")
        return out, step.value

    def take_scene(self, rid: int) -> Scene:
        h = c_void_p()
        self._check(self.lib.fs_server_take_scene(self.h, rid, byref(h)), "fs_server_take_scene")
        sc = Scene.attach(h.value)
        sc._owned = True  # the caller now owns the scene
        return sc

    def stats(self) -> dict:
        s = ServerStats()
        self._check(self.lib.fs_server_stats_get(self.h, byref(s)), "fs_server_stats_get")
        return s.as_dict()

    def latencies(self) -> np.ndarray:
        n = c_int64()
        self._check(self.lib.fs_server_step_latencies(self.h, None, 0, byref(n)), "fs_server_step_latencies")
        out = np.zeros(max(1, n.value), dtype=np.float64)
        self._check(self.lib.fs_server_step_latencies(self.h, out.ctypes.data_as(POINTER(c_double)), n.value, byref(n)),
                    "fs_server_step_latencies")
        return out[:n.value]

    def clock(self) -> float:
        return self.lib.fs_server_clock(self.h)


def scene_emitter_params(scene_id: int, assets_dir: str | None = None, params: Params | None = None) -> Params:
    """Emitter/burner placement of a reference scene (fs_scene_emitter_params) applied to params."""
    p = params if params is not None else default_params(scene_id=scene_id)
    st = load_library().fs_scene_emitter_params(scene_id, assets_dir.encode() if assets_dir else None, byref(p))
    if st != FS_OK:
        raise FireSimError(f"fs_scene_emitter_params({scene_id}) failed ({st})")
    return p


def image_stats(bgra: np.ndarray) -> dict:
    """Colour statistics of a BGRA frame (fs_image_stats_of)."""
    a = np.ascontiguousarray(bgra, dtype=np.uint32).ravel()
    out = ImageStats()
    st = load_library().fs_image_stats_of(a.ctypes.data_as(POINTER(c_uint32)), a.size, byref(out))
    if st != FS_OK:
        raise FireSimError(f"fs_image_stats_of failed ({st})")
    return out.as_dict()
