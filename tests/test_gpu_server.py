"""Serving (fs_server_*): requests admitted by HBM budget into a continuously
batched loop. A request's fields and frames are the same alone and in a mix of
other requests arriving while it runs (bit for bit, the tolerance of a batched
step), admission respects the budget and FIFO order, and latency bookkeeping
covers every step."""
import numpy as np
import pytest

import firesim

pytestmark = pytest.mark.gpu

FIELDS = firesim.SCALAR_FIELDS + ["u", "v", "w", "pressure", "divergence", "light", "shadow"]


def params(seed, iters, turbulence=0.5):
    rng = np.random.default_rng(seed)
    ang = rng.uniform(0, 2 * np.pi)
    speed = rng.uniform(0, 0.5)
    return firesim.default_params(scene_id=3, turbulence=turbulence, sor_omega=1.7, sor_iterations=iters, noise_seed=seed,
                                  wind_x=float(speed * np.cos(ang)), wind_z=float(speed * np.sin(ang)), raymarch_steps=32)


def request(n, seed, steps, iters=20, storage=firesim.STORAGE_FP32, keep=False, arrival=-1.0, render_every=5):
    nx, ny, nz = n
    r = max(2.0, min(nx, nz) * 6 / 64)
    em = [(nx / 2, r, nz / 2, r, 1.0, 1500.0)]
    return firesim.make_request(nx, ny, nz, params(seed, iters), em, steps=steps, render_every=render_every, frame=(48, 32),
                                storage=storage, keep_scene=keep, arrival_s=arrival)


def alone_scene(n, seed, steps, iters=20, storage=firesim.STORAGE_FP32):
    nx, ny, nz = n
    s = firesim.Scene(nx, ny, nz, 48, 32, storage)
    s.set_params(params(seed, iters))
    r = max(2.0, min(nx, nz) * 6 / 64)
    s.set_emitters([(nx / 2, r, nz / 2, r, 1.0, 1500.0)])
    s.reset()
    s.step(steps)
    return s


GIB = 1 << 30


@pytest.mark.parametrize("storage", [firesim.STORAGE_FP32, firesim.STORAGE_BF16])
def test_request_alone_equals_in_mix(hip_lib, storage):
    target = ((40, 48, 36), 7, 25)
    with firesim.Server(hbm_budget_bytes=8 * GIB, sor_kernel=2) as sv:
        rid = sv.submit(request(*target, storage=storage, keep=True))
        sv.run()
        st = sv.status(rid)
        assert st["state"] == firesim.REQ_DONE and st["steps_done"] == 25 and st["frames"] == 5
        frame_alone, fstep = sv.frame(rid)
        assert fstep == 25
        a = sv.take_scene(rid)
    mix = [((24, 24, 24), 1, 40, 13), ((64, 32, 48), 2, 12, 30), ((32, 64, 32), 3, 30, 7), ((16, 16, 16), 4, 60, 1),
           ((48, 48, 48), 5, 20, 45), ((20, 40, 28), 6, 35, 20)]
    with firesim.Server(hbm_budget_bytes=8 * GIB, sor_kernel=-1) as sv:
        ids = [sv.submit(request(n, seed, steps, iters, storage=storage, arrival=0.003 * i)) for i, (n, seed, steps, iters) in enumerate(mix)]
        rid = sv.submit(request(*target, storage=storage, keep=True, arrival=0.004))
        sv.run()
        stats = sv.stats()
        assert stats["requests_done"] == len(mix) + 1 and stats["requests_failed"] == 0
        assert stats["max_active_seen"] >= 2  # it really ran in a mix
        frame_mix, _ = sv.frame(rid)
        m = sv.take_scene(rid)
        for i in ids:
            assert sv.status(i)["steps_done"] == [x[2] for x in mix][ids.index(i)]
    plain = alone_scene(*target, storage=storage)
    for f in FIELDS:
        assert np.array_equal(a.read(f), m.read(f)), f"{f}: alone vs mix"
        assert np.array_equal(a.read(f), plain.read(f)), f"{f}: server vs plain scene"
    assert np.array_equal(frame_alone, frame_mix)
    assert np.array_equal(frame_alone, plain.render_bgra8())
    for s in (a, m, plain):
        s.close()


def test_admission_respects_budget_and_fifo(hip_lib):
    n = (48, 48, 48)
    one = firesim.load_library()
    desc = firesim.SceneDesc(*n, 48, 32, firesim.STORAGE_FP32, 0)
    b = firesim.c_uint64()
    assert one.fs_scene_bytes(firesim.byref(desc), firesim.byref(b)) == 0
    with firesim.Server(hbm_budget_bytes=int(b.value * 1.5), sor_kernel=0) as sv:
        ids = [sv.submit(request(n, 10 + i, 6, iters=5)) for i in range(4)]
        with pytest.raises(firesim.FireSimError, match="budget"):
            sv.submit(request((96, 96, 96), 1, 5))
        sv.run()
        st = [sv.status(i) for i in ids]
        stats = sv.stats()
    assert stats["max_active_seen"] == 1
    assert stats["peak_resident_bytes"] <= b.value * 1.5
    assert all(s["state"] == firesim.REQ_DONE for s in st)
    fin = [s["finished_s"] for s in st]
    adm = [s["admitted_s"] for s in st]
    assert fin == sorted(fin) and adm == sorted(adm)
    for earlier, later in zip(st, st[1:]):
        assert later["admitted_s"] >= earlier["finished_s"] - 1e-9  # one at a time


def test_latency_bookkeeping(hip_lib):
    reqs = [((32, 32, 32), 21, 15), ((24, 40, 24), 22, 9), ((40, 24, 32), 23, 12)]
    with firesim.Server(hbm_budget_bytes=4 * GIB, sor_kernel=2, use_graph=False) as sv:
        for i, (n, seed, steps) in enumerate(reqs):
            sv.submit(request(n, seed, steps, arrival=0.002 * i, render_every=0))
        sv.run()
        stats = sv.stats()
        lat = sv.latencies()
    total_steps = sum(r[2] for r in reqs)
    assert stats["steps"] == total_steps == lat.size
    assert stats["voxel_steps"] == sum(int(np.prod(r[0])) * r[2] for r in reqs)
    assert (lat > 0).all()
    assert 0 < stats["step_ms_p50"] <= stats["step_ms_p99"] <= stats["step_ms_max"] + 1e-9
    assert stats["queue_ms_p50"] >= 0 and stats["busy_s"] > 0
    assert stats["frames"] == 0


def test_round_voxel_cap_bounds_the_round(hip_lib):
    """With max_round_voxels, no round steps more admitted cells than the cap
    (a request larger than the cap runs alone), every request completes, and a
    request's fields are those of the uncapped server."""
    sizes = [((32, 32, 32), 11, 12), ((32, 32, 32), 12, 9), ((40, 32, 32), 13, 14), ((32, 32, 32), 14, 7), ((64, 64, 64), 15, 6),
             ((32, 32, 32), 16, 10)]
    cap = 3 * 32 ** 3
    out = {}
    for limit in (0, cap):
        with firesim.Server(hbm_budget_bytes=8 * GIB, sor_kernel=2, max_round_voxels=limit) as sv:
            ids = [sv.submit(request(n, seed, steps, keep=(i == 2))) for i, (n, seed, steps) in enumerate(sizes)]
            sv.run()
            assert all(sv.status(i)["state"] == firesim.REQ_DONE for i in ids)
            if limit:
                assert sv.stats()["max_active_seen"] <= 3
            s = sv.take_scene(ids[2])
            out[limit] = {f: s.read(f) for f in FIELDS}
            s.close()
    for f in FIELDS:
        assert np.array_equal(out[0][f], out[cap][f]), f
