"""ctypes wrapper of liboracle.so, the FP64 CPU restatement of the step.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product. The method names mirror
firesim.Scene so parity tests read the same on both sides.
"""
from __future__ import annotations

import ctypes
import os
import sys
from ctypes import POINTER, byref, c_double, c_float, c_int, c_long, c_uint32, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(REPO, "native-firesim_amd"))

import firesim  # noqa: E402  (boundary structs: fs_params, fs_emitter, fs_metrics)

# ORACLE_LIB points the tests at another build (the mutation sweep's mutants)
LIB_PATH = os.environ.get("ORACLE_LIB", os.path.join(HERE, "liboracle.so"))

SALT_NOISE2, SALT_NOISE3, SALT_JITTER, SALT_DITHER, SALT_GRAIN = 1, 2, 3, 4, 5

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing; build it with `make -C oracle`")
    lib = ctypes.CDLL(LIB_PATH)
    D = POINTER(c_double)
    sigs = {
        "or_create": (c_void_p, [c_int, c_int, c_int, c_int, c_int]),
        "or_destroy": (None, [c_void_p]),
        "or_clamp_params": (c_int, [POINTER(firesim.Params), POINTER(firesim.Params)]),
        "or_set_params": (None, [c_void_p, POINTER(firesim.Params)]),
        "or_get_params": (None, [c_void_p, POINTER(firesim.Params)]),
        "or_set_emitters": (None, [c_void_p, POINTER(firesim.Emitter), c_int]),
        "or_reset": (None, [c_void_p]),
        "or_set_time": (None, [c_void_p, c_float, c_float]),
        "or_get_time": (None, [c_void_p, POINTER(c_float), POINTER(c_float), POINTER(c_int)]),
        "or_light_dims": (None, [c_void_p, POINTER(c_int), POINTER(c_int), POINTER(c_int)]),
        "or_field_count": (c_long, [c_void_p, c_int]),
        "or_set_field": (c_int, [c_void_p, c_int, D]),
        "or_get_field": (c_int, [c_void_p, c_int, D]),
        "or_run_stage": (None, [c_void_p, c_int]),
        "or_step": (None, [c_void_p, c_int]),
        "or_render_hdr": (None, [c_void_p, D]),
        "or_tonemap": (None, [c_void_p, D, POINTER(c_uint32)]),
        "or_metrics": (None, [c_void_p, POINTER(firesim.Metrics)]),
        "or_threads": (c_int, []),
        "or_set_threads": (None, [c_int]),
        "or_sample_grid": (c_double, [D, c_int, c_int, c_int, c_double, c_double, c_double, c_double, c_double, c_double]),
        "or_smoothstep": (c_double, [c_double, c_double, c_double]),
        "or_pcg": (c_uint32, [c_uint32]),
        "or_hash2": (c_double, [c_int, c_int, c_uint32]),
        "or_hash3": (c_double, [c_int, c_int, c_int, c_uint32]),
        "or_value_noise2": (c_double, [c_double, c_double, c_uint32]),
        "or_value_noise3": (c_double, [c_double, c_double, c_double, c_uint32]),
        "or_fbm2": (c_double, [c_double, c_double, c_uint32]),
        "or_fbm3": (c_double, [c_double, c_double, c_double, c_uint32]),
        "or_salt": (c_uint32, [c_uint32, c_uint32]),
        "or_blackbody": (None, [c_double, D]),
        "or_aces": (c_double, [c_double]),
        "or_camera_response": (None, [D, c_double, D]),
        "or_domain_fade": (c_double, [c_double, c_double, c_double]),
        "or_maccormack": (c_double, [D, c_int, c_int, c_int, c_double, c_double, D, c_double, c_double, c_double, c_double]),
        "or_capsule_mask": (c_double, [c_double] * 7),
        "or_bed_material": (c_double, [c_void_p, c_double, c_double]),
        "or_bed_source": (c_double, [c_void_p, c_double, c_double, c_double, c_double]),
        "or_light_build": (None, [c_void_p]),
        "or_light_propagate": (None, [c_void_p, c_double, c_double]),
        "or_dir_transmittance": (None, [c_void_p]),
        "or_room_background": (None, [c_void_p, c_double, c_double, D, c_double, D]),
        "or_camera_eye": (None, [c_void_p, D]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(POINTER(c_double))


class OracleScene:
    """FP64 scene with the same method names as firesim.Scene."""

    def __init__(self, nx, ny, nz, frame_w=64, frame_h=64):
        self.lib = load()
        self.nx, self.ny, self.nz = nx, ny, nz
        self.frame_w, self.frame_h = frame_w, frame_h
        self.h = self.lib.or_create(nx, ny, nz, frame_w, frame_h)
        if not self.h:
            raise ValueError("invalid oracle grid")

    def close(self):
        if self.h:
            self.lib.or_destroy(self.h)
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

    def set_params(self, p):
        self.lib.or_set_params(self.h, byref(p))

    def get_params(self):
        p = firesim.Params()
        self.lib.or_get_params(self.h, byref(p))
        return p

    def set_emitters(self, emitters):
        arr = (firesim.Emitter * max(1, len(emitters)))(
            *[firesim.Emitter(*e) if not isinstance(e, firesim.Emitter) else e for e in emitters])
        self.lib.or_set_emitters(self.h, arr, len(emitters))

    def reset(self):
        self.lib.or_reset(self.h)

    def set_time(self, t, rt):
        self.lib.or_set_time(self.h, t, rt)

    def get_time(self):
        t, rt, fi = c_float(), c_float(), c_int()
        self.lib.or_get_time(self.h, byref(t), byref(rt), byref(fi))
        return t.value, rt.value, fi.value

    def shape(self, field):
        return firesim.field_shape(field, self.nx, self.ny, self.nz)

    def read(self, field) -> np.ndarray:
        f = firesim.FIELD[field] if isinstance(field, str) else field
        out = np.empty(self.shape(f), dtype=np.float64)
        if self.lib.or_get_field(self.h, f, _dptr(out)) != 0:
            raise ValueError(field)
        return out

    def write(self, field, data):
        f = firesim.FIELD[field] if isinstance(field, str) else field
        arr = np.ascontiguousarray(data, dtype=np.float64).reshape(self.shape(f))
        if self.lib.or_set_field(self.h, f, _dptr(arr)) != 0:
            raise ValueError(field)

    def run_stage(self, stage):
        self.lib.or_run_stage(self.h, stage)

    def step(self, n=1):
        self.lib.or_step(self.h, n)

    def render_hdr(self) -> np.ndarray:
        out = np.empty((self.frame_h, self.frame_w, 4), dtype=np.float64)
        self.lib.or_render_hdr(self.h, _dptr(out))
        return out

    def tonemap(self, hdr) -> np.ndarray:
        src = np.ascontiguousarray(hdr, dtype=np.float64).reshape(self.frame_h, self.frame_w, 4)
        out = np.empty((self.frame_h, self.frame_w), dtype=np.uint32)
        self.lib.or_tonemap(self.h, _dptr(src), out.ctypes.data_as(POINTER(c_uint32)))
        return out

    def render_bgra8(self) -> np.ndarray:
        return self.tonemap(self.render_hdr())

    def metrics(self) -> dict:
        m = firesim.Metrics()
        self.lib.or_metrics(self.h, byref(m))
        return m.as_dict()

    def state(self) -> dict:
        return {name: self.read(name) for name in firesim.STATE_FIELDS}

    def load_state(self, state: dict):
        for name, arr in state.items():
            self.write(name, arr)


def set_threads(n: int):
    load().or_set_threads(n)


def threads() -> int:
    return load().or_threads()


def clamp_params(p):
    out = firesim.Params()
    load().or_clamp_params(byref(p), byref(out))
    return out


def sample_grid(f: np.ndarray, scale, u, v, w) -> float:
    """Trilinear sample of a (dz, dy, dx) array with per-axis coordinate scales."""
    arr = np.ascontiguousarray(f, dtype=np.float64)
    dz, dy, dx = arr.shape
    return load().or_sample_grid(_dptr(arr), dx, dy, dz, scale[0], scale[1], scale[2], u, v, w)
