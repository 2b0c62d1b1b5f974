"""Host-side checks of the C ABI (no GPU): the library loads, exports every
entry point declared in include/firesim.h, and its host-only helpers and
argument validation behave as documented."""
import ctypes
import os
import struct

import numpy as np
import pytest

import firesim

LIB = firesim.load_library()


def test_exports_every_declared_symbol():
    declared = firesim.exported_symbols()
    assert len(declared) >= 30
    missing = [name for name in declared if not hasattr(LIB, name)]
    assert not missing, missing
    assert LIB.fs_abi_version() == 2


def test_default_params_are_reference_defaults():
    p = firesim.Params()
    LIB.fs_default_params(ctypes.byref(p))
    # FireSettings defaults, src/fire_cuda.h:65-101; SOR constants :104-105
    assert p.dt == pytest.approx(1 / 60)
    assert (p.mouse_x, p.turbulence, p.detail, p.smoke, p.intensity) == pytest.approx((0.5, 0.72, 0.88, 0.92, 1.0))
    assert p.raymarch_steps == 56 and p.sor_iterations == 40
    assert p.sor_omega == pytest.approx(1.52)
    assert bytes(p) == bytes(firesim.default_params())


def test_light_dims_formula():
    out = [ctypes.c_int32() for _ in range(3)]
    for n in [(4, 4, 4), (64, 64, 64), (176, 208, 128), (256, 384, 256), (1000, 1000, 1000)]:
        assert LIB.fs_light_dims(*n, *[ctypes.byref(o) for o in out]) == 0
        assert tuple(o.value for o in out) == firesim.light_dims(*n)
    assert firesim.light_dims(176, 208, 128) == (44, 52, 32)  # reference canonical grid


def test_scene_bytes_accounting():
    b = ctypes.c_uint64()
    d = firesim.SceneDesc(128, 128, 128, 512, 512, firesim.STORAGE_FP32, 0)
    assert LIB.fs_scene_bytes(ctypes.byref(d), ctypes.byref(b)) == 0
    fp32 = b.value
    cells = 128 ** 3
    # 20 scalar + 6 face + 6 cell-sized scratch/pressure + 3 curl-potential edge arrays, plus light/frame/partials
    assert 35 * 4 * cells < fp32 < 37 * 4 * cells + 16 * 1024 * 1024
    d.storage = firesim.STORAGE_BF16
    assert LIB.fs_scene_bytes(ctypes.byref(d), ctypes.byref(b)) == 0
    assert fp32 - b.value == 6 * 2 * cells  # heat, fuel, soot in both buffers at 2 B


@pytest.mark.parametrize("desc", [(3, 8, 8, 16, 16, 0, 0), (8, 8, 8, 0, 16, 0, 0), (8, 8, 8, 16, 16, 7, 0),
                                  (1024, 1024, 1023, 16, 16, 0, 0)])
def test_create_rejects_bad_descriptors_without_touching_the_device(desc):
    d = firesim.SceneDesc(*desc)
    h = ctypes.c_void_p()
    assert LIB.fs_create(ctypes.byref(d), ctypes.byref(h)) == 1  # FS_ERR_INVALID_ARGUMENT
    assert not h.value
    assert LIB.fs_last_global_error()


def test_grid_cap_keeps_every_field_under_2_30_values():
    # kernels address fields with 32-bit byte offsets: (nx+1)(ny+1)(nz+1) <= 2^30
    h = ctypes.c_void_p()
    over = firesim.SceneDesc(1024, 1024, 1023, 16, 16, 0, 0)  # 1025 * 1025 * 1024 > 2^30
    assert LIB.fs_create(ctypes.byref(over), ctypes.byref(h)) == 1
    assert b"2^30" in LIB.fs_last_global_error()


def test_null_handles_are_errors():
    assert LIB.fs_step(None, 1) == 2
    assert LIB.fs_reset(None) == 2
    assert LIB.fs_destroy(None) == 0
    p = firesim.Params()
    assert LIB.fs_clamp_params(None, ctypes.byref(p)) == 1


def test_bmp_matches_reference_layout(tmp_path):
    h, w = 3, 5
    pix = (np.arange(h * w, dtype=np.uint32) * 0x010203) | 0xFF000000
    path = str(tmp_path / "f.bmp")
    firesim.write_bmp(path, pix.reshape(h, w))
    data = open(path, "rb").read()
    # BITMAPFILEHEADER + BITMAPINFOHEADER, top-down 32-bit BI_RGB (src/main.cpp:2511-2536)
    bfType, bfSize, _, _, bfOffBits = struct.unpack_from("<HIHHI", data, 0)
    biSize, biWidth, biHeight, biPlanes, biBitCount, biCompression, biSizeImage = struct.unpack_from("<IiiHHII", data, 14)
    assert (bfType, bfOffBits, biSize) == (0x4D42, 54, 40)
    assert (biWidth, biHeight, biPlanes, biBitCount, biCompression) == (w, -h, 1, 32, 0)
    assert biSizeImage == h * w * 4 and bfSize == 54 + h * w * 4 == len(data)
    assert np.array_equal(np.frombuffer(data[54:], dtype="<u4"), pix)
    assert LIB.fs_write_bmp(path.encode(), None, w, h) == 1
    assert LIB.fs_write_bmp(b"/nonexistent-dir/x.bmp", pix.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)), w, h) == 6


def test_copy_bandwidth_rejects_bad_sizes_before_touching_a_device():
    out = ctypes.c_double(-1.0)
    for nbytes, reps in ((0, 1), ((1 << 20) + 8, 1), (1 << 20, 0), (1 << 20, 1001), (65 << 30, 1)):
        assert LIB.fs_copy_bandwidth(0, nbytes, reps, ctypes.byref(out)) == 1, (nbytes, reps)
        assert out.value == 0.0
    assert LIB.fs_copy_bandwidth(0, 1 << 20, 1, None) == 1


def test_no_packed_fp32_in_the_device_code():
    """The gfx950 code objects hold no v_pk_{add,mul,fma}_f32 (the build switches
    the packed-fp32-ops feature off; the checker disassembles each object)."""
    import glob
    import subprocess
    pkg = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "native-firesim_amd")
    objs = sorted(glob.glob(os.path.join(pkg, "_build", "*.hip.o")))
    if not objs:
        pytest.skip("objects not built")
    r = subprocess.run([os.path.join(pkg, "check_packed_fp32.sh"), *objs], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
