"""Batches (fs_batch_*): scenes of uneven grids and SOR iteration counts stepped
together, one persistent launch per pass, give exactly the fields of the same
scenes stepped alone. Covers every pressure kernel family, graph and direct
launches, BF16 storage, membership changes between steps, per-scene calls on
members, autotuning, and the argument checks."""
import numpy as np
import pytest

import firesim

pytestmark = pytest.mark.gpu

# (nx, ny, nz, SOR iterations, scene id)
GRIDS = [(24, 28, 20, 7, 3), (32, 32, 32, 13, 3), (40, 24, 36, 30, 1), (20, 48, 24, 4, 3), (16, 16, 16, 1, 2)]
FIELDS = firesim.SCALAR_FIELDS + ["u", "v", "w", "pressure", "divergence", "light", "shadow"]


def make(grid, seed, storage=firesim.STORAGE_FP32):
    nx, ny, nz, iters, sid = grid
    s = firesim.Scene(nx, ny, nz, 24, 16, storage)
    rng = np.random.default_rng(seed)
    p = firesim.default_params(scene_id=sid, turbulence=0.5, sor_omega=1.7, sor_iterations=iters, noise_seed=seed,
                               wind_x=float(rng.uniform(-0.5, 0.5)), wind_z=float(rng.uniform(-0.5, 0.5)), raymarch_steps=24)
    s.set_params(p)
    r = max(2.0, min(nx, nz) / 5)
    s.set_emitters([(nx / 2 + rng.uniform(-2, 2), r, nz / 2 + rng.uniform(-2, 2), r, 1.0, 1500.0)])
    s.reset()
    return s


def assert_same(a, b, what):
    for f in FIELDS:
        x, y = a.read(f), b.read(f)
        assert np.array_equal(x, y), f"{what}: field {f} differs (max |d| {np.abs(x - y).max()})"
    ma, mb = a.metrics(), b.metrics()
    assert ma["heat_release_total"] == mb["heat_release_total"], what
    assert a.get_time() == b.get_time(), what


@pytest.mark.parametrize("sor_kernel", ["half_sweep", "tb_t16_k4", "tb_t24_k4", "tb_t24_z8_k4", "tb_t24_z10_k4", "tb_t24_z6_k4", "tb_t26_z13_k3", "tb_t22_z8_k5"])
@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_equals_alone(hip_lib, sor_kernel, use_graph):
    alone = [make(g, 11 + i) for i, g in enumerate(GRIDS)]
    mixed = [make(g, 11 + i) for i, g in enumerate(GRIDS)]
    with firesim.Batch() as batch:
        batch.set_tuning(sor_kernel=sor_kernel, use_graph=use_graph)
        for s in mixed:
            batch.add(s)
        assert batch.count() == len(GRIDS)
        for _ in range(5):
            batch.step(1)
        for s in alone:
            s.step(5)
        for i, (a, m) in enumerate(zip(alone, mixed)):
            assert_same(a, m, f"scene {i} after 5 steps")
        # membership changes between steps: one leaves, a new one joins
        batch.remove(mixed[2])
        extra_a, extra_m = make((28, 20, 28, 9, 3), 99), make((28, 20, 28, 9, 3), 99)
        batch.add(extra_m)
        batch.step(3)
        for s in alone[:2] + alone[3:] + [extra_a]:
            s.step(3)
        mixed[2].step(2)  # the removed scene continues on its own stream
        alone[2].step(2)
        for i, (a, m) in enumerate(zip(alone + [extra_a], mixed + [extra_m])):
            assert_same(a, m, f"scene {i} after membership change")
    for s in alone + mixed + [extra_a, extra_m]:
        s.close()


def test_batch_bf16_equals_alone(hip_lib):
    alone = [make(g, 31 + i, firesim.STORAGE_BF16) for i, g in enumerate(GRIDS[:3])]
    mixed = [make(g, 31 + i, firesim.STORAGE_BF16) for i, g in enumerate(GRIDS[:3])]
    with firesim.Batch() as batch:
        for s in mixed:
            batch.add(s)
        batch.step(6)
        for s in alone:
            s.step(6)
        for i, (a, m) in enumerate(zip(alone, mixed)):
            assert_same(a, m, f"bf16 scene {i}")


def test_member_calls_interleave_with_batch_steps(hip_lib):
    """Per-scene calls on a member (step, parameter change, render) run in order on the batch stream."""
    a, m = make(GRIDS[1], 5), make(GRIDS[1], 5)
    other = make(GRIDS[0], 6)
    p = firesim.default_params(scene_id=3, turbulence=0.9, sor_omega=1.6, sor_iterations=11, noise_seed=5, raymarch_steps=24)
    with firesim.Batch() as batch:
        batch.add(m)
        batch.add(other)
        batch.step(2)
        m.step(1)
        m.set_params(p)
        fm = m.render_bgra8()
        batch.step(2)
        a.step(3)
        a.set_params(p)
        fa = a.render_bgra8()
        a.step(2)
        assert np.array_equal(fa, fm)
        assert_same(a, m, "interleaved")


def test_batch_autotune_keeps_state(hip_lib):
    alone = [make(g, 41 + i) for i, g in enumerate(GRIDS)]
    mixed = [make(g, 41 + i) for i, g in enumerate(GRIDS)]
    with firesim.Batch() as batch:
        for s in mixed:
            batch.add(s)
        batch.step(2)
        for s in alone:
            s.step(2)
        before = [s.read("pressure") for s in mixed]
        rep = batch.autotune(reps=1)
        assert rep["chosen"] in firesim.SOR_KERNELS
        assert all(v > 0 for v in rep["sor_ms"].values())
        for b, s in zip(before, mixed):
            assert np.array_equal(b, s.read("pressure"))
        batch.step(2)
        for s in alone:
            s.step(2)
        for i, (a, m) in enumerate(zip(alone, mixed)):
            assert_same(a, m, f"after autotune, scene {i}")
        ms = batch.step_latency(3)
        assert ms.shape == (3,) and (ms > 0).all()


def test_batch_argument_checks(hip_lib):
    s32 = make(GRIDS[0], 1)
    s16 = make(GRIDS[0], 1, firesim.STORAGE_BF16)
    b1, b2 = firesim.Batch(), firesim.Batch()
    b1.add(s32)
    with pytest.raises(firesim.FireSimError, match="already in this batch"):
        b1.add(s32)
    with pytest.raises(firesim.FireSimError, match="another batch"):
        b2.add(s32)
    with pytest.raises(firesim.FireSimError, match="storage"):
        b1.add(s16)
    with pytest.raises(firesim.FireSimError, match="not in this batch"):
        b2.remove(s32)
    with pytest.raises(firesim.FireSimError, match="SOR kernel"):
        b1.set_tuning(sor_kernel=99)
    b1.close()  # members leave and stay usable
    s32.step(1)
    s32.synchronize()
    assert np.isfinite(s32.read("heat")).all()
    b2.close()
    s32.close()
    s16.close()


@pytest.mark.parametrize("first,second", [(firesim.STORAGE_FP32, firesim.STORAGE_BF16), (firesim.STORAGE_BF16, firesim.STORAGE_FP32)])
def test_batch_storage_switch_after_empty(hip_lib, first, second):
    """A batch that stepped scenes of one storage mode, emptied, and then takes a
    scene of the other mode steps it with that mode's passes (the captured step
    graphs are keyed by storage mode): bit for bit the scene stepped alone."""
    grid = (24, 28, 20, 7, 3)
    a = make(grid, 41, first)
    b_alone, b_batched = make(grid, 42, second), make(grid, 42, second)
    with firesim.Batch() as batch:
        batch.set_tuning(sor_kernel="tb_t24_z10_k4", use_graph=True)
        batch.add(a)
        batch.step(2)
        batch.remove(a)
        batch.add(b_batched)
        batch.step(3)
        batch.synchronize()
        b_alone.step(3)
        assert_same(b_alone, b_batched, "second storage mode after the switch")
        a_alone = make(grid, 41, first)
        a_alone.step(2)
        assert_same(a_alone, a, "first storage mode before the switch")
    for s in (a, a_alone, b_alone, b_batched):
        s.close()
