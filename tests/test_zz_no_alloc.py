"""fs_pressure_solve_noalloc: the pressure pass on caller device buffers with
the library's static scratch array in place of a second pressure buffer. Its
results equal the scene path's pressure pass (fs_run_stage, same kernels and
red/black order) bit for bit, on two calls in a row with the second on a
smaller grid (the scratch keeps the first call's contents), for kernels with
odd and even launch counts; oversized grids are rejected before any launch."""
import ctypes

import numpy as np
import pytest

import firesim

pytestmark = pytest.mark.gpu

HIP_H2D, HIP_D2H = 1, 2


def hip():
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipDeviceSynchronize.argtypes = []
    return lib


class DevBuf:
    def __init__(self, h, arr):
        self.h, self.n = h, arr.nbytes
        self.ptr = ctypes.c_void_p()
        assert h.hipMalloc(ctypes.byref(self.ptr), self.n) == 0
        assert h.hipMemcpy(self.ptr, arr.ctypes.data, self.n, HIP_H2D) == 0

    def read(self, shape):
        out = np.empty(shape, dtype=np.float32)
        assert self.h.hipMemcpy(out.ctypes.data, self.ptr, self.n, HIP_D2H) == 0
        return out

    def free(self):
        self.h.hipFree(self.ptr)


def scene_pressure(grid, d, p0, kernel, omega, iters):
    """The pressure pass of a scene (the entry point the no-allocation one varies)."""
    nx, ny, nz = grid
    s = firesim.Scene(nx, ny, nz, 16, 16)
    s.set_params(firesim.default_params(scene_id=3, sor_omega=omega, sor_iterations=iters))
    s.set_tuning(sor_kernel=firesim.SOR_KERNELS.index(kernel), use_graph=False)
    s.write("divergence", d)
    s.write("pressure", p0)
    s.run_stage(firesim.STAGE_PRESSURE)
    out = s.read("pressure")
    s.close()
    return out


@pytest.mark.parametrize("kernel", ["tb_t26_z13_k3", "tb_t24_z10_k4", "half_sweep", "tb_t22_z8_k5"])
def test_noalloc_pressure_equals_scene_pass(hip_lib, kernel):
    h = hip()
    lib = hip_lib
    kid = firesim.SOR_KERNELS.index(kernel)
    rng = np.random.default_rng(kid)
    for grid, iters in (((72, 60, 66), 41), ((30, 24, 28), 13)):  # second call smaller
        nx, ny, nz = grid
        d = rng.normal(size=(nz, ny, nx)).astype(np.float32)
        p0 = (0.1 * rng.normal(size=(nz, ny, nx))).astype(np.float32)
        want = scene_pressure(grid, d, p0, kernel, 1.7, iters)
        print(f"noalloc: kernel {kernel} grid {grid} iterations {iters}", flush=True)
        pb, db = DevBuf(h, p0), DevBuf(h, d)
        st = lib.fs_pressure_solve_noalloc(pb.ptr, db.ptr, nx, ny, nz, 1.7, iters, kid, 0, None)
        assert st == 0, lib.fs_last_global_error()
        assert h.hipDeviceSynchronize() == 0
        got = pb.read((nz, ny, nx))
        pb.free()
        db.free()
        assert np.array_equal(got, want), f"{kernel} {grid}: max |d| {np.abs(got - want).max()}"


def test_noalloc_rejects_oversized_and_bad_arguments(hip_lib):
    lib = hip_lib
    fake = ctypes.c_void_p(0x1000)  # never dereferenced: every case fails its host check first
    assert lib.fs_pressure_solve_noalloc(fake, fake, 256, 256, 257, 1.7, 10, 9, 0, None) == 1
    assert b"FS_NOALLOC_MAX_CELLS" in lib.fs_last_global_error()
    assert lib.fs_pressure_solve_noalloc(None, fake, 8, 8, 8, 1.7, 10, 9, 0, None) == 1
    assert lib.fs_pressure_solve_noalloc(fake, fake, 3, 8, 8, 1.7, 10, 9, 0, None) == 1
    assert lib.fs_pressure_solve_noalloc(fake, fake, 8, 8, 8, 2.5, 10, 9, 0, None) == 1
    assert lib.fs_pressure_solve_noalloc(fake, fake, 8, 8, 8, 1.7, 0, 9, 0, None) == 1
    assert lib.fs_pressure_solve_noalloc(fake, fake, 8, 8, 8, 1.7, 10, 99, 0, None) == 1
