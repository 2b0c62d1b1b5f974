"""Every pressure-solve kernel variant gives bit-identical pressure to the
plain half-sweep kernel (same per-cell FP32 expression and red/black order),
including grids smaller than a tile and grids that are not tile multiples.

The larger grids have many more tiles than the GPU holds at once, so a tile
kernel that read halos from the buffer it writes would see neighbours' later
values and fail here (the tile launches read one pressure buffer and write
the other)."""
import numpy as np
import pytest

import compare
import firesim

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("grid", [(32, 40, 24), (4, 4, 4), (19, 7, 33), (48, 48, 48), (160, 96, 144)])
@pytest.mark.parametrize("iters", [1, 3, 40])
def test_sor_variants_bit_identical(hip_lib, grid, iters):
    rng = np.random.default_rng(grid[0] * 100 + iters)
    d = rng.normal(size=(grid[2], grid[1], grid[0])).astype(np.float32)
    p0 = rng.normal(size=d.shape).astype(np.float32) * 0.1
    params = compare.emitter_params(sor_iterations=iters, sor_omega=1.7)
    results = []
    for kernel in range(len(firesim.SOR_KERNELS)):
        s = firesim.Scene(*grid, 8, 8)
        s.set_params(params)
        s.set_sor_kernel(kernel)
        s.write("divergence", d)
        s.write("pressure", p0)
        s.run_stage(firesim.STAGE_PRESSURE)
        results.append(s.read("pressure"))
        s.close()
    for kernel, r in enumerate(results[1:], start=1):
        assert np.array_equal(r, results[0]), firesim.SOR_KERNELS[kernel]


def test_autotune_restores_pressure_and_picks_a_kernel(hip_lib):
    grid = (64, 48, 40)
    rng = np.random.default_rng(9)
    s = firesim.Scene(*grid, 8, 8)
    s.set_params(compare.emitter_params(sor_iterations=20, sor_omega=1.7))
    d = rng.normal(size=(grid[2], grid[1], grid[0])).astype(np.float32)
    p0 = rng.normal(size=d.shape).astype(np.float32)
    s.write("divergence", d)
    s.write("pressure", p0)
    rep = s.autotune(reps=1)
    assert rep["chosen"] in firesim.SOR_KERNELS
    assert all(ms > 0 for ms in rep["sor_ms"].values())
    assert np.array_equal(s.read("pressure"), p0)
    assert np.array_equal(s.read("divergence"), d)
    assert firesim.SOR_KERNELS[s.sor_kernel()] == rep["chosen"]
    s.close()


def test_copy_bandwidth_is_a_plausible_hbm_rate(hip_lib):
    """Streaming copy of 256 MiB: read + write bytes per second below the
    nominal 8 TB/s HBM peak and above a tenth of it."""
    gbps = firesim.copy_bandwidth(256 << 20, 5)
    print("copy GB/s", gbps, flush=True)
    assert 800.0 < gbps < 8000.0
