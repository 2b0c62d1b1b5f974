"""Pins the FP64 oracle (CPU only).

The reference holds no field-level golden vectors for the step, so each oracle
function is pinned by a check that a plausible mistake would fail:

  sampler (sample_grid, src/fire_cuda.cu:489-600)   exact on affine fields; matches
        scipy.ndimage.map_coordinates(order=1); node/clamp cases; the (n-1) quirk
  smoothstep (:227-234)                              textbook Hermite; inverted edges
  PCG hash / value noise / fbm (:381-443 shape)      lattice values, continuity, range,
        octave sum against a direct evaluation; uniformity of the hash
  blackbody (:1420-1443)                             Helland's published table points
  ACES fit (:2246-2253)                              Narkowicz's closed form
  velocity advection (:880-932)                      constant fields -> c*damping; zero
        velocity on faces is the identity (face mapping uses n, not n-1); walls 0
  advect-react (:934-1126)                           cold-air early-out algebra;
        uniform-field MacCormack is exact; Arrhenius burn of a 1500 K cell in closed
        form; clamps; no fuel -> no heat release
  force (:1128-1183, gather)                         zero inputs -> zero force;
        buoyancy-only closed form on v faces; wind profile on u and w faces
  divergence (:1185-1199)                            numpy np.diff of random faces
  SOR (:1201-1227)                                   dense-matrix red-black iteration on
        a tiny grid; convergence for compatible rhs; residual == post-projection
        divergence (ties K11-K13 together)
  light (:1445-1565)                                 empty fields -> zero light, shadow 1
  render + tonemap (:1953-2297)                      empty fields -> room only;
        debug mode 3 is the transmittance; dither bounded by 1/2 LSB
  clamps (:3135-3179)                                product fs_clamp_params == oracle
  embers (:2323-2452)                                additive and non-negative; absent in
        debug modes; campfire/burner spawn gates on the right hash (k=2 / k=0);
        each spark is rank-1 in its colour, between the hot and cooled colours

The passes beyond these primitives (advect-react term by term, force, light,
metrics, beds, reset, room colour, raymarch, embers, tonemap) are pinned in
test_oracle_pins.py.

"Parity unpinned" items: none of the above is compared against the reference's
own sin-hash noise (not portable, replaced by the integer hash by design).
"""
import ctypes
import math

import numpy as np
import pytest
from scipy import ndimage

import compare
import firesim
import oracle

L = oracle.load()


def F32(x: float) -> float:
    """The float32 value of a reference literal, as a Python float."""
    return float(np.float32(x))


def test_sampler_affine_exact_and_scipy():
    rng = np.random.default_rng(0)
    dz, dy, dx = 5, 6, 7
    z, y, x = np.meshgrid(np.arange(dz), np.arange(dy), np.arange(dx), indexing="ij")
    a, b, c, d = rng.normal(size=4)
    f = a + b * x + c * y + d * z
    scale = (dx - 1, dy - 1, dz - 1)
    for _ in range(50):
        u, v, w = rng.uniform(-0.2, 1.2, 3)
        cu, cv, cw = (min(max(t, 0.0), 1.0) for t in (u, v, w))
        got = oracle.sample_grid(f, scale, u, v, w)
        want = a + b * cu * scale[0] + c * cv * scale[1] + d * cw * scale[2]
        assert got == pytest.approx(want, abs=1e-12)
    g = rng.normal(size=(dz, dy, dx))
    for _ in range(50):
        u, v, w = rng.uniform(0, 1, 3)
        coords = np.array([[w * scale[2]], [v * scale[1]], [u * scale[0]]])
        want = ndimage.map_coordinates(g, coords, order=1, mode="nearest")[0]
        assert oracle.sample_grid(g, scale, u, v, w) == pytest.approx(want, abs=1e-12)
    # lattice node and edge clamps
    assert oracle.sample_grid(g, scale, 1.0, 1.0, 1.0) == g[-1, -1, -1]
    assert oracle.sample_grid(g, scale, -3.0, 0.0, 0.0) == g[0, 0, 0]
    assert oracle.sample_grid(g, scale, 2 / 6, 3 / 5, 1 / 4) == pytest.approx(g[1, 3, 2])
    # cell-centre sampling is the reference's (n-1) resample, not the cell value
    cc = oracle.sample_grid(g, scale, 2.5 / dx, 0.0, 0.0)
    assert cc != g[0, 0, 2]
    fx = 2.5 / dx * (dx - 1)
    assert cc == pytest.approx(g[0, 0, int(fx)] + (g[0, 0, int(fx) + 1] - g[0, 0, int(fx)]) * (fx - int(fx)))


def test_smoothstep():
    for lo, hi, x in [(0.0, 1.0, 0.3), (2.0, 5.0, 4.1), (0.2, 0.9, -1.0), (0.2, 0.9, 3.0)]:
        t = min(max((x - lo) / (hi - lo), 0.0), 1.0)
        assert L.or_smoothstep(lo, hi, x) == pytest.approx(t * t * (3 - 2 * t))
        assert L.or_smoothstep(hi, lo, x) == pytest.approx(1 - t * t * (3 - 2 * t))


def test_hash_and_value_noise():
    salt = L.or_salt(7, 1)
    vals = np.array([L.or_hash2(i, j, salt) for i in range(-20, 20) for j in range(-20, 20)])
    assert vals.min() >= 0 and vals.max() < 1
    hist, _ = np.histogram(vals, bins=10, range=(0, 1))
    assert hist.min() > 100 and hist.max() < 220  # 1600 samples, ~160/bin
    # 24-bit values: exactly representable multiples of 2^-24
    assert np.all(vals * 2**24 == np.round(vals * 2**24))
    # the hash is the published PCG output-permutation hash
    for v in (0, 1, 12345, 2**31 + 7):
        st = (v * 747796405 + 2891336453) & 0xFFFFFFFF
        word = (((st >> ((st >> 28) + 4)) ^ st) * 277803737) & 0xFFFFFFFF
        assert L.or_pcg(v) == ((word >> 22) ^ word)
    # lattice points return the hash; continuity across cell edges; range
    assert L.or_value_noise2(3.0, -2.0, salt) == L.or_hash2(3, -2, salt)
    assert L.or_value_noise3(1.0, 2.0, -4.0, salt) == L.or_hash3(1, 2, -4, salt)
    for x in (0.0, 5.0, -3.0):
        a = L.or_value_noise2(x - 1e-9, 0.37, salt)
        b = L.or_value_noise2(x + 1e-9, 0.37, salt)
        assert abs(a - b) < 1e-6
    # fbm is the 5-octave sum with the reference's octave map
    px, py = 0.31, 4.7
    want, amp = 0.0, 0.5
    for _ in range(5):
        want += amp * L.or_value_noise2(px, py, salt)
        px, py = px * F32(2.03) + F32(17.7), py * F32(2.01) - F32(9.2)
        amp *= 0.5
    assert L.or_fbm2(0.31, 4.7, salt) == pytest.approx(want, abs=1e-12)
    f3 = [L.or_fbm3(x, 2 * x, -x, salt) for x in np.linspace(-3, 3, 101)]
    assert min(f3) >= 0 and max(f3) <= 1


def test_blackbody_matches_helland_table():
    out = (ctypes.c_double * 3)()
    # (kelvin, published 8-bit rgb of the Helland fit)
    table = [(1000, (255, 68, 0)), (2700, (255, 167, 87)), (6600, (255, 255, 255)), (10000, (202, 218, 255))]
    for k, rgb in table:
        L.or_blackbody(float(k), out)
        got = tuple(round(v * 255) for v in out)
        assert all(abs(g - r) <= 1 for g, r in zip(got, rgb)), (k, got, rgb)


def test_aces_closed_form():
    f = lambda x: x * (2.51 * x + 0.03) / (x * (2.43 * x + 0.59) + 0.14)  # noqa: E731
    for x in (0.0, 0.18, 0.5, 0.9):
        assert L.or_aces(x) == pytest.approx(min(1.0, f(x)), rel=1e-6)
    assert L.or_aces(100.0) == 1.0


def _scene(grid=(8, 9, 7), **kw):
    o = oracle.OracleScene(*grid, 16, 12)
    p = compare.emitter_params(**kw)
    o.set_params(p)
    o.set_emitters([])
    o.reset()
    return o, p


def test_velocity_advection_constant_and_identity():
    o, p = _scene()
    for name, c in (("u", 0.3), ("v", -0.2), ("w", 0.7)):
        o.write(name, np.full(o.shape(firesim.FIELD[name]), c))
    o.run_stage(firesim.STAGE_ADVECT_VELOCITY)
    u, v, w = o.read("u"), o.read("v"), o.read("w")
    assert np.all(u[:, :, 0] == 0) and np.all(u[:, :, -1] == 0)
    assert np.allclose(u[:, :, 1:-1], 0.3 * F32(0.995), atol=1e-15)
    assert np.all(v[:, 0, :] == 0) and np.allclose(v[:, 1:, :], -0.2 * F32(0.996), atol=1e-15)
    assert np.all(w[0] == 0) and np.all(w[-1] == 0) and np.allclose(w[1:-1], 0.7 * F32(0.995), atol=1e-15)
    # shear flow u = f(y): the backtrace moves only along x, where u is constant,
    # so each face receives the y-resample of f at (y+0.5)/ny*(ny-1) (the
    # reference's n-1 mapping on collocated axes), times the damping.
    o2, _ = _scene()
    nz, ny, nxp1 = o2.shape(firesim.FIELD["u"])
    f = np.linspace(-0.4, 0.6, ny) ** 2
    o2.write("u", np.broadcast_to(f[None, :, None], (nz, ny, nxp1)))
    o2.write("v", np.zeros(o2.shape(firesim.FIELD["v"])))
    o2.write("w", np.zeros(o2.shape(firesim.FIELD["w"])))
    o2.run_stage(firesim.STAGE_ADVECT_VELOCITY)
    got = o2.read("u")
    for y in range(ny):
        fy = (y + 0.5) / ny * (ny - 1)
        y0 = int(fy)
        y1 = min(y0 + 1, ny - 1)
        want = (f[y0] + (f[y1] - f[y0]) * (fy - y0)) * F32(0.995)
        assert np.allclose(got[:, y, 1:-1], want, rtol=1e-13), y
    assert np.all(got[:, :, [0, -1]] == 0)
    assert np.all(o2.read("v") == 0) and np.all(o2.read("w") == 0)


def test_react_early_out_algebra():
    o, p = _scene(grid=(6, 12, 6))
    ox = np.full(o.shape(0), 0.5)
    o.write("oxygen", ox)
    k = np.full(o.shape(0), 0.004)  # activity 0.06*k = 0.00024 < 0.00035
    o.write("turbulence", k)
    o.run_stage(firesim.STAGE_ADVECT_REACT)
    dt = F32(p.dt)
    v = (np.arange(12) + 0.5) / 12
    cold = v > 0.18
    got_o = o.read("oxygen")[:, cold, :]
    assert np.allclose(got_o, min(1.0, 0.5 + 0.5 * dt * F32(0.08)), rtol=1e-14)
    assert np.allclose(o.read("turbulence")[:, cold, :], 0.004 * max(0.0, 1 - dt * F32(0.54)), rtol=1e-14)
    assert o.metrics()["heat_release_step"] == 0.0


def test_react_arrhenius_burn_closed_form():
    """A uniform 1500 K fuel-rich cell: MacCormack of uniform fields is exact and
    burnMass = 1.42*O2 * act(1500K)=1 * exp(-3800/1500) * (118+34*I) * dt (O2-limited)."""
    grid = (6, 6, 6)
    o, p = _scene(grid=grid, intensity=1.0)
    heat = (1500.0 - 293.0) / 335.0
    o.write("heat", np.full(o.shape(0), heat))
    o.write("fuel", np.full(o.shape(0), 5.0))
    o.write("oxygen", np.full(o.shape(0), 0.01))
    o.run_stage(firesim.STAGE_ADVECT_REACT)
    dt = F32(p.dt)
    tempK = 293.0 + heat * 335.0
    assert tempK >= 1220.0
    oxy_lim = min(5.0, 0.01 * F32(1.42))
    rate = math.exp(-3800.0 / tempK) * (118.0 + 34.0 * 1.0) * dt
    burn = min(oxy_lim, oxy_lim * rate)
    m = o.metrics()
    assert m["heat_release_step"] == pytest.approx(7.0 * burn * np.prod(grid), rel=1e-6)
    caps = dict(heat=8.8, fuel=5.5, oxygen=1.0, soot=7.4, char=2.8, ash=1.9, pyrolysis=3.6, progress=3.4,
                turbulence=4.2, soot_optics=10.0)
    for name, cap in caps.items():
        arr = o.read(name)
        assert arr.min() >= 0 and arr.max() <= np.float32(cap)


def test_react_no_fuel_no_release():
    o, _ = _scene(grid=(6, 6, 6))
    o.write("heat", np.full(o.shape(0), 4.0))
    o.run_stage(firesim.STAGE_ADVECT_REACT)
    assert o.metrics()["heat_release_step"] == 0.0


def test_force_zero_and_buoyancy_closed_form():
    grid = (5, 8, 4)
    o, p = _scene(grid=grid, turbulence=0.0)
    o.run_stage(firesim.STAGE_FORCE)
    for name in ("u", "v", "w"):
        assert np.all(o.read(name) == 0)
    h = 2.0
    o.write("heat", np.full(o.shape(0), h))
    o.run_stage(firesim.STAGE_FORCE)
    v = o.read("v")
    dt = F32(p.dt)
    for y in range(1, grid[1] + 1):
        vc = (y - 1 + 0.5) / grid[1]
        fade = L.or_smoothstep(0.96, 0.08, vc)
        want = dt * (h * 1.0 * F32(1.62) * fade)
        want = min(6.4, max(-0.9, want))
        assert np.allclose(v[:, y, :], want, rtol=1e-12), y
    assert np.all(v[:, 0, :] == 0)
    assert np.all(o.read("u") == 0) and np.all(o.read("w") == 0)
    # wind: u faces get two cell contributions, w the same with wind_z
    o3, p3 = _scene(grid=grid, turbulence=0.0, wind_x=0.5, wind_z=-0.25)
    o3.run_stage(firesim.STAGE_FORCE)
    u, w = o3.read("u"), o3.read("w")
    for y in range(grid[1]):
        prof = F32(0.34) + (y + 0.5) / grid[1] * F32(0.96)
        cu = dt * 0.5 * prof
        assert np.allclose(u[:, y, 1:-1], 2 * cu, rtol=1e-12)
        assert np.allclose(w[1:-1, y, :], 2 * dt * -0.25 * prof, rtol=1e-12)
    assert np.all(u[:, :, 0] == 0) and np.all(u[:, :, -1] == 0)


def test_divergence_matches_numpy():
    grid = (6, 5, 7)
    o, _ = _scene(grid=grid, turbulence=0.0)
    rng = np.random.default_rng(3)
    faces = {n: rng.uniform(-0.5, 0.5, o.shape(firesim.FIELD[n])) for n in ("u", "v", "w")}
    # interior-only random faces so the boundary clamp leaves values unchanged
    faces["u"][:, :, [0, -1]] = 0
    faces["v"][:, 0, :] = 0
    faces["w"][[0, -1]] = 0
    for n, a in faces.items():
        o.write(n, a)
    o.run_stage(firesim.STAGE_FORCE)
    # with zero heat and turbulence only the vertical damping force acts (on v)
    u, v, w = o.read("u"), o.read("v"), o.read("w")
    assert np.array_equal(u, faces["u"]) and np.array_equal(w, faces["w"])
    div = np.diff(u, axis=2) + np.diff(v, axis=1) + np.diff(w, axis=0)
    assert np.allclose(o.read("divergence"), -div, atol=1e-14)
    assert np.all(o.read("pressure") == 0)


def _dense_red_black(d, omega, iters):
    nz, ny, nx = d.shape
    p = np.zeros_like(d)
    idx = np.arange(d.size).reshape(d.shape)
    nbrs = []
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                nb = [idx[z, y, max(x - 1, 0)], idx[z, y, min(x + 1, nx - 1)], idx[z, max(y - 1, 0), x],
                      idx[z, min(y + 1, ny - 1), x], idx[max(z - 1, 0), y, x], idx[min(z + 1, nz - 1), y, x]]
                nbrs.append(nb)
    A = np.zeros((d.size, d.size))
    for i, nb in enumerate(nbrs):
        for j in nb:
            A[i, j] += 1.0
    colour = (np.indices(d.shape).sum(0) & 1).ravel()
    pf, df = p.ravel(), d.ravel()
    for _ in range(iters):
        for par in (0, 1):
            m = colour == par
            proj = (df + A @ pf) * F32(1.0 / 6.0)
            pf = np.where(m, pf + omega * (proj - pf), pf)
    return pf.reshape(d.shape)


def test_sor_against_dense_red_black_and_residual_identity():
    grid = (4, 5, 3)
    o, p = _scene(grid=grid, sor_omega=1.7, sor_iterations=7)
    rng = np.random.default_rng(4)
    d = rng.normal(size=o.shape(0))
    o.write("divergence", d)
    o.write("pressure", np.zeros_like(d))
    o.run_stage(firesim.STAGE_PRESSURE)
    want = _dense_red_black(d, F32(1.7), 7)
    assert np.allclose(o.read("pressure"), want, rtol=1e-12, atol=1e-14)
    # convergence on a compatible right-hand side (sum d = 0)
    o2, _ = _scene(grid=(6, 6, 6), sor_omega=1.5, sor_iterations=400)
    d2 = rng.normal(size=o2.shape(0))
    d2 -= d2.mean()
    o2.write("divergence", d2)
    o2.run_stage(firesim.STAGE_PRESSURE)
    assert o2.metrics()["sor_residual_l2"] < 1e-6 * np.sqrt((d2 ** 2).mean())


def test_residual_equals_post_projection_divergence():
    o = oracle.OracleScene(16, 20, 12, 16, 16)
    p = compare.emitter_params(sor_iterations=15)
    o.set_params(p)
    o.set_emitters([(8.0, 3.0, 6.0, 3.0, 1.0, 1500.0)])
    o.reset()
    o.step(8)
    m = o.metrics()
    assert m["divergence_after_l2"] > 0
    assert m["sor_residual_l2"] == pytest.approx(m["divergence_after_l2"], rel=1e-9)


def test_light_and_render_of_empty_fields():
    o, p = _scene(grid=(12, 14, 10), turbulence=0.0)
    o.run_stage(firesim.STAGE_LIGHT)
    light = o.read("light")
    assert np.all(light[..., :3] == 0) and np.all(light[..., 3] == 0)
    assert np.all(o.read("shadow") == 1.0)
    hdr = o.render_hdr()
    assert np.all(hdr[..., 3] == 1.0) and hdr[..., :3].min() >= 0
    p3 = compare.emitter_params(turbulence=0.0, render_debug_mode=3)
    o.set_params(p3)
    dbg = o.render_hdr()
    assert np.all(dbg[..., :3] == 1.0)  # transmittance of an empty volume


def test_tonemap_dither_bound():
    o, p = _scene(grid=(6, 6, 6))
    rng = np.random.default_rng(5)
    hdr = np.concatenate([rng.uniform(0, 2, (12, 16, 3)), np.ones((12, 16, 1))], -1)
    out = o.tonemap(hdr)
    resp = (ctypes.c_double * 3)()
    exp = F32(p.exposure)
    for (y, x) in [(0, 0), (3, 7), (11, 15)]:
        L.or_camera_response((ctypes.c_double * 3)(*hdr[y, x, :3]), exp, resp)
        for ch, shift in ((0, 16), (1, 8), (2, 0)):
            byte = (int(out[y, x]) >> shift) & 0xFF
            assert abs(byte - resp[ch] * 255) <= 1.0
        assert (int(out[y, x]) >> 24) == 0xFF


def test_clamps_match_product():
    lib = firesim.load_library()
    rng = np.random.default_rng(6)
    for _ in range(200):
        p = firesim.default_params()
        for name, typ in p._fields_:
            if typ is ctypes.c_float:
                setattr(p, name, float(rng.uniform(-5, 5)))
            elif typ is ctypes.c_int32:
                setattr(p, name, int(rng.integers(-10, 2000)))
        for arr in (p.burner_center_x, p.burner_center_y, p.burner_center_z):
            for i in range(4):
                arr[i] = float(rng.uniform(-3, 3))
        mine = firesim.Params()
        assert lib.fs_clamp_params(ctypes.byref(p), ctypes.byref(mine)) == 0
        theirs = oracle.clamp_params(p)
        assert bytes(mine) == bytes(theirs)


def _pcg(v: int) -> int:
    """PCG hash (Jarzynski & Olano 2020), the integer hash of the noise and the sparks."""
    state = (v * 747796405 + 2891336453) & 0xFFFFFFFF
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & 0xFFFFFFFF
    return (word >> 22) ^ word


def _ember_increments(scene_id, count):
    """HDR increment of each spark k: render(k + 1 sparks) - render(k sparks)."""
    p = compare.emitter_params(scene_id=scene_id, ember_count=0, burner_center_count=0)
    o = oracle.OracleScene(24, 28, 20, 48, 32)
    o.set_params(p)
    o.set_emitters([(12.0, 2.0, 10.0, 4.0, 1.0, 1500.0)])
    o.reset()
    o.step(10)
    prev = o.render_hdr()
    base = prev
    incs = []
    for k in range(1, count + 1):
        p.ember_count = k
        o.set_params(p)
        h = o.render_hdr()
        incs.append(h - prev)
        prev = h
    return o, p, base, prev, incs


def test_embers_additive_and_gated():
    salt = _pcg((7 + 6 * 0x9E3779B9) & 0xFFFFFFFF)  # noise_seed 7, the ember salt
    for scene_id, gate_k, gate_at in ((1, 2, F32(0.34)), (0, None, None), (3, None, None)):
        o, p, base, full, incs = _ember_increments(scene_id, 64)
        assert np.all(full[..., 3] == base[..., 3])
        lit = 0
        for k, inc in enumerate(incs):
            assert inc.min() >= 0.0
            h = [L.or_hash2(k, j, salt) for j in range(4)]
            if gate_k is not None and h[gate_k] > gate_at:
                assert inc.max() == 0.0, (scene_id, k)
                continue
            if inc.max() == 0.0:
                continue
            lit += 1
            # one spark is colour x a scalar footprint; the colour lies between the
            # cooled ember (0.18, 0.026, 0.008) and the hot core (1.0, 0.42, 0.09)
            rgb = inc[..., :3].reshape(-1, 3)
            rgb = rgb[rgb[:, 0] > 1e-9]
            gr, br = rgb[:, 1] / rgb[:, 0], rgb[:, 2] / rgb[:, 0]
            assert np.ptp(gr) <= 1e-9 and np.ptp(br) <= 1e-9
            assert F32(0.026) / F32(0.18) - 1e-9 <= gr[0] <= F32(0.42) + 1e-9
            assert F32(0.008) / F32(0.18) - 1e-9 <= br[0] <= F32(0.09) + 1e-9
        assert lit >= 2, scene_id
        p.render_debug_mode = 3
        p.ember_count = 64
        o.set_params(p)
        with_sparks = o.render_hdr()
        p.ember_count = 0
        o.set_params(p)
        assert np.array_equal(with_sparks, o.render_hdr())
        o.close()


def test_burner_embers_gate_on_first_hash():
    salt = _pcg((7 + 6 * 0x9E3779B9) & 0xFFFFFFFF)
    o, p, base, full, incs = _ember_increments(2, 48)
    for k, inc in enumerate(incs):
        if L.or_hash2(k, 0, salt) > F32(0.14):
            assert inc.max() == 0.0, k
    o.close()
