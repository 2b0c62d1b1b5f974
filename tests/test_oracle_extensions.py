"""Pins of the oracle's extensions (CPU only): divergence-free curl-noise
forcing, the closed top, the directional-light transmittance and the gizmo
overlay (drawGizmos, src/fire_cuda.cu:1909-1951).

The extensions have no reference counterpart, so they are pinned by what the
mathematics fixes: the discrete divergence of a discrete curl is zero and a
potential that vanishes tangentially on a wall lets no flux through it;
amplitude 0 adds nothing and the forcing is linear in the amplitude;
Beer-Lambert through a uniform medium is exp(-kappa * e * length); a zero
light colour changes no pixel; the gizmo colours are the reference's
constants blended at full alpha where the geometry puts alpha at 1.
"""
import ctypes
import math

import numpy as np
import pytest

import compare
import firesim
import oracle

L = oracle.load()


def f(x):
    return float(np.float32(x))


def _zero_scene(grid, frame=(16, 12), **params):
    """Oracle scene with every scalar and face zero (so the reference force terms vanish)."""
    base = dict(wind_x=0.0, wind_z=0.0)
    base.update(params)
    o = oracle.OracleScene(*grid, *frame)
    o.set_params(compare.emitter_params(**base))
    o.reset()
    for name in firesim.STATE_FIELDS:
        o.write(name, np.zeros(o.shape(name)))
    return o


def _faces(o):
    return o.read("u"), o.read("v"), o.read("w")


def _divergence(u, v, w):
    return (u[:, :, 1:] - u[:, :, :-1]) + (v[:, 1:, :] - v[:, :-1, :]) + (w[1:, :, :] - w[:-1, :, :])


# ------------------------------------------------------------ curl noise ----
@pytest.mark.parametrize("closed_top", [0, 1])
def test_curl_noise_is_divergence_free_with_no_wall_flux(closed_top):
    grid = (9, 7, 6)
    o = _zero_scene(grid, noise_forcing=1, curl_noise=0.7, closed_top=closed_top, turbulence=1.2)
    o.set_time(1.3, 1.3)
    o.run_stage(firesim.STAGE_FORCE)
    u, v, w = _faces(o)
    scale = max(np.abs(u).max(), np.abs(v).max(), np.abs(w).max())
    assert scale > 1e-3  # the forcing is not trivially zero
    assert np.abs(_divergence(u, v, w)).max() <= 1e-13 * scale
    assert np.abs(o.read("divergence")).max() <= 1e-13 * scale
    # no flux through the side walls and the floor; the top only when closed
    assert np.all(u[:, :, 0] == 0) and np.all(u[:, :, -1] == 0)
    assert np.all(w[0] == 0) and np.all(w[-1] == 0)
    assert np.all(v[:, 0, :] == 0)
    if closed_top:
        assert np.all(v[:, -1, :] == 0)
    else:
        assert np.abs(v[:, -1, :]).max() > 1e-4  # the open top lets the noise through
    o.close()


def _edge_potential(grid, c, t, seed, closed_top):
    """Edge potential as documented: one octave of the lattice value noise at
    the edge's normalized position, frequency 6, drifting (0.23, -0.17, 0.19) t,
    scaled by (nx + ny + nz) / 18; zero on wall edges."""
    nx, ny, nz = grid
    salt = L.or_salt(seed, 7 + c)
    scale = f(float(np.float32(nx + ny + nz) / np.float32(18.0)))
    out = np.zeros((nz + 1, ny + 1, nx + 1))
    for k in range(nz + 1):
        for j in range(ny + 1):
            for i in range(nx + 1):
                if c == 0 and (j == 0 or (closed_top and j == ny) or k in (0, nz) or i >= nx):
                    continue
                if c == 1 and (i in (0, nx) or k in (0, nz) or j >= ny):
                    continue
                if c == 2 and (i in (0, nx) or j == 0 or (closed_top and j == ny) or k >= nz):
                    continue
                ex = ((i + 0.5) if c == 0 else i) / nx
                ey = ((j + 0.5) if c == 1 else j) / ny
                ez = ((k + 0.5) if c == 2 else k) / nz
                n = L.or_value_noise3(ex * 6.0 + t * f(0.23), ey * 6.0 - t * f(0.17), ez * 6.0 + t * f(0.19), salt)
                out[k, j, i] = scale * (n - 0.5)
    return out


@pytest.mark.parametrize("closed_top", [0, 1])
def test_curl_noise_faces_are_the_curl_of_the_documented_potential(closed_top):
    grid = (7, 6, 5)
    amp, t = 0.9, 2.25
    o = _zero_scene(grid, noise_forcing=1, curl_noise=amp, closed_top=closed_top, noise_seed=41)
    o.set_time(t, t)
    o.run_stage(firesim.STAGE_FORCE)
    u, v, w = _faces(o)
    px, py, pz = (_edge_potential(grid, c, t, 41, closed_top) for c in range(3))
    nx, ny, nz = grid
    kc = f(1.0 / 60.0) * f(amp)
    eu = kc * ((pz[:nz, 1:, :] - pz[:nz, :ny, :]) - (py[1:, :ny, :] - py[:nz, :ny, :]))
    ev = kc * ((px[1:, :, :nx] - px[:nz, :, :nx]) - (pz[:nz, :, 1:] - pz[:nz, :, :nx]))
    ew = kc * ((py[:, :ny, 1:] - py[:, :ny, :nx]) - (px[:, 1:, :nx] - px[:, :ny, :nx]))
    eu[:, :, [0, nx]] = 0.0
    ev[:, 0, :] = 0.0
    if closed_top:
        ev[:, ny, :] = 0.0
    ew[[0, nz], :, :] = 0.0
    for got, want in ((u, eu), (v, ev), (w, ew)):
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)
    o.close()


def test_curl_noise_amplitude_zero_adds_nothing_and_is_linear():
    grid = (8, 6, 7)
    faces = {}
    for amp in (0.0, 0.4, 0.8):
        o = _zero_scene(grid, noise_forcing=1, curl_noise=amp, turbulence=1.5)
        o.set_time(0.7, 0.7)
        o.run_stage(firesim.STAGE_FORCE)
        faces[amp] = _faces(o)
        o.close()
    assert all(np.all(a == 0) for a in faces[0.0])
    for a4, a8 in zip(faces[0.4], faces[0.8]):
        np.testing.assert_allclose(a8, 2.0 * a4, rtol=1e-6, atol=1e-15)  # float32 amplitudes 0.4f and 0.8f
    # the reference's fbm noise forcing (noise_forcing 0) does move the same zero state
    o = _zero_scene(grid, noise_forcing=0, turbulence=1.5)
    o.set_time(0.7, 0.7)
    o.run_stage(firesim.STAGE_FORCE)
    assert np.abs(o.read("u")).max() > 1e-4
    o.close()


def test_curl_noise_replaces_only_the_fbm_terms():
    """On a burning state, noise_forcing 1 at amplitude 0 equals the reference
    force with its two fbm noise terms removed: the difference to noise_forcing
    0 is noise * (0.20 + upperDrag * 0.16) * curlGain * dt on u and w, and v is
    untouched."""
    grid = (10, 12, 8)
    prm = compare.emitter_params(turbulence=0.9)
    state, t, rt = compare.warm_state(*grid, prm, [(5.0, 3.0, 4.0, 2.5, 1.0, 1500.0)], steps=6, frame=(8, 8))
    out = {}
    for mode in (0, 1):
        o = oracle.OracleScene(*grid, 8, 8)
        o.set_params(compare.emitter_params(turbulence=0.9, noise_forcing=mode, curl_noise=0.0))
        o.set_emitters([(5.0, 3.0, 4.0, 2.5, 1.0, 1500.0)])
        o.reset()
        o.load_state(state)
        o.set_time(t, rt)
        o.run_stage(firesim.STAGE_FORCE)
        out[mode] = _faces(o)
        o.close()
    np.testing.assert_array_equal(out[0][1], out[1][1])
    assert np.abs(out[0][0] - out[1][0]).max() > 1e-6


# ------------------------------------------------------------- closed top ----
def test_closed_top_zeroes_the_top_face_through_a_step():
    grid = (10, 10, 8)
    em = [(5.0, 2.0, 4.0, 2.5, 1.0, 1600.0)]
    tops = {}
    for closed in (0, 1):
        o = oracle.OracleScene(*grid, 8, 8)
        o.set_params(compare.emitter_params(closed_top=closed, turbulence=0.8))
        o.set_emitters(em)
        o.reset()
        o.step(25)
        v = o.read("v")
        tops[closed] = np.abs(v[:, -1, :]).max()
        if closed:
            assert np.all(v[:, -1, :] == 0)
            # a closed box: the projected field carries no net flux out of the domain
            u, w = o.read("u"), o.read("w")
            assert np.all(u[:, :, [0, -1]] == 0) and np.all(w[[0, -1]] == 0) and np.all(v[:, 0, :] == 0)
        o.close()
    assert tops[0] > 1e-5  # the reference's open top lets the plume out


def test_closed_top_in_velocity_advection():
    grid = (6, 8, 5)
    rng = np.random.default_rng(4)
    for closed in (0, 1):
        o = _zero_scene(grid, closed_top=closed)
        o.write("u", rng.uniform(-0.5, 0.5, o.shape("u")))
        o.write("v", rng.uniform(0.2, 0.9, o.shape("v")))
        o.write("w", rng.uniform(-0.5, 0.5, o.shape("w")))
        o.run_stage(firesim.STAGE_ADVECT_VELOCITY)
        top = o.read("v")[:, -1, :]
        assert np.all(top == 0) if closed else np.abs(top).max() > 1e-3
        o.close()


# --------------------------------------------------------- directional light ----
def _slab_len(p0, d):
    lo = np.array([f(-1.05), f(0.02), f(-0.82)])
    hi = np.array([f(1.05), f(2.05), f(0.82)])
    best = 1e30
    for a in range(3):
        if d[a] > f(1e-6):
            best = min(best, (hi[a] - p0[a]) / d[a])
        elif d[a] < -f(1e-6):
            best = min(best, (lo[a] - p0[a]) / d[a])
    return best


def _unit32(d):
    d = np.asarray(d, np.float32)
    n = np.sqrt(np.float32(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]))
    return np.array([float(d[0] / n), float(d[1] / n), float(d[2] / n)])


@pytest.mark.parametrize("direction", [(0.0, -1.0, 0.0), (0.3, -1.0, 0.2), (1.0, 0.0, 0.0), (-0.5, -0.4, 0.7), (0.0, 0.0, 0.0)])
@pytest.mark.parametrize("ext", [0.0, 0.37])
def test_dir_transmittance_is_beer_lambert_in_a_uniform_medium(direction, ext):
    """exp(-3 * e * len): the weighted samples sum to the ray's length in the
    volume box exactly, so only the direction, the box and kappa enter."""
    grid = (40, 36, 32)
    o = _zero_scene(grid, dir_light=1, dir_light_x=direction[0], dir_light_y=direction[1], dir_light_z=direction[2])
    lnz, lny, lnx, _ = o.shape("light")
    light = np.zeros((lnz, lny, lnx, 4))
    light[..., 3] = ext
    o.write("light", light)
    L.or_dir_transmittance(o.h)
    got = o.read("dir_transmittance")
    d = _unit32(direction) if any(direction) else np.array([0.0, -1.0, 0.0])
    toward = -d
    want = np.empty_like(got)
    for z in range(lnz):
        for y in range(lny):
            for x in range(lnx):
                p0 = np.array([(x + 0.5) / lnx * f(2.10) - f(1.05), (y + 0.5) / lny * f(2.03) + f(0.02), (z + 0.5) / lnz * f(1.64) - f(0.82)])
                want[z, y, x] = math.exp(-f(3.0) * ext * _slab_len(p0, toward))
    # the direction is normalized in FP32 (an FMA or not in |d|^2 moves it by 1 ulp)
    np.testing.assert_allclose(got, want, rtol=1e-12 if sum(c != 0 for c in direction) <= 1 else 1e-6, atol=1e-15)
    o.close()


# Light grids (18, 20, 16), (24, 20, 16) and (18, 20, 18): the smallest light
# cell, which sets the step, is along y, x and z in turn. The directions leave
# the volume box through each of its six faces.
@pytest.mark.parametrize("grid", [(24, 20, 18), (96, 20, 18), (24, 20, 72)])
@pytest.mark.parametrize("direction", [(0.3, -1.0, 0.2), (-0.5, -0.4, 0.7), (0.0, 0.0, 1.0), (0.2, 0.3, -1.0), (-0.3, 0.8, 0.4)])
def test_dir_transmittance_matches_restatement_in_a_random_medium(direction, grid):
    """The march from each light-cell centre toward the light, restated from
    its definition: step ds = half the smallest light cell, samples at
    ds (k + 1/2) weighted sat(1/2 + (len - s)/ds) over the ray's length len in
    the volume box, the extinction channel sampled trilinearly at the sample's
    volume coordinates, T = exp(-3 ds sum). A random medium makes every
    coordinate offset and scale count (a uniform one cannot see them)."""
    rng = np.random.default_rng(7)
    o = _zero_scene(grid, dir_light=1, dir_light_x=direction[0], dir_light_y=direction[1], dir_light_z=direction[2])
    lnz, lny, lnx, _ = o.shape("light")
    light = np.zeros((lnz, lny, lnx, 4))
    light[..., 3] = rng.uniform(0.0, 0.6, (lnz, lny, lnx))
    o.write("light", light)
    L.or_dir_transmittance(o.h)
    got = o.read("dir_transmittance")
    ext = np.ascontiguousarray(light[..., 3])
    toward = -_unit32(direction)
    ds = 0.5 * min(f(2.10) / lnx, f(2.03) / lny, f(1.64) / lnz)
    for z in range(0, lnz, 2):
        for y in range(0, lny, 2):
            for x in range(0, lnx, 3):
                p0 = np.array([(x + 0.5) / lnx * f(2.10) - f(1.05), (y + 0.5) / lny * f(2.03) + f(0.02), (z + 0.5) / lnz * f(1.64) - f(0.82)])
                length = _slab_len(p0, toward)
                tau, k = 0.0, 0
                while True:
                    sk = ds * (k + 0.5)
                    wk = min(1.0, 0.5 + (length - sk) / ds)
                    if wk <= 0.0:
                        break
                    q = p0 + toward * sk
                    uvw = ((q[0] + f(1.05)) / f(2.10), (q[1] - f(0.02)) / f(2.03), (q[2] + f(0.82)) / f(1.64))
                    tau += wk * oracle.sample_grid(ext, (lnx - 1, lny - 1, lnz - 1), *uvw)
                    k += 1
                # (the FP32 normalization of the direction moves it by up to an ulp)
                assert got[z, y, x] == pytest.approx(math.exp(-f(3.0) * ds * tau), rel=1e-6), (x, y, z)
    assert got.min() < 0.5
    o.close()


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_dir_transmittance_looks_toward_the_light(axis):
    """Smoke in the upper half of one axis, the light arriving from that side:
    cells in the lower half are shadowed by at least the half's thickness less
    a light cell, the last row by at most two light cells of it; with the light
    from the other side, the lower cells are clear."""
    grid = (40, 36, 32)
    extent = [f(2.10), f(2.03), f(1.64)][axis]
    for sign, lower_shadowed in ((-1.0, True), (1.0, False)):
        d = [0.0, 0.0, 0.0]
        d[axis] = sign
        o = _zero_scene(grid, dir_light=1, dir_light_x=d[0], dir_light_y=d[1], dir_light_z=d[2])
        lnz, lny, lnx, _ = o.shape("light")
        n = [lnx, lny, lnz][axis]
        light = np.zeros((lnz, lny, lnx, 4))
        E = 0.5
        idx = [slice(None)] * 3
        idx[2 - axis] = slice(n // 2, None)
        light[tuple(idx) + (3,)] = E
        o.write("light", light)
        L.or_dir_transmittance(o.h)
        T = np.moveaxis(o.read("dir_transmittance"), 2 - axis, 0)  # axis first
        cell = extent / n
        half = (n - n // 2) * cell
        low = T[: n // 2 - 1]
        if lower_shadowed:
            assert low.max() <= math.exp(-3.0 * E * (half - cell)) * (1 + 1e-12)
            assert T[-1].min() >= math.exp(-3.0 * E * 2 * cell)
        else:
            assert np.all(low == 1.0)
        o.close()


def test_dir_light_zero_colour_changes_no_pixel_and_is_linear():
    grid = (16, 18, 14)
    em = [(8.0, 3.0, 7.0, 3.0, 1.0, 1500.0)]
    prm = compare.emitter_params(turbulence=0.7, raymarch_steps=32)
    state, t, rt = compare.warm_state(*grid, prm, em, steps=12, frame=(24, 18))
    frames = {}
    for key, kw in {"off": dict(dir_light=0), "zero": dict(dir_light=1, dir_light_r=0.0, dir_light_g=0.0, dir_light_b=0.0),
                    "one": dict(dir_light=1, dir_light_r=1.0, dir_light_g=0.5, dir_light_b=0.25),
                    "two": dict(dir_light=1, dir_light_r=2.0, dir_light_g=1.0, dir_light_b=0.5)}.items():
        o = oracle.OracleScene(*grid, 24, 18)
        o.set_params(compare.emitter_params(turbulence=0.7, raymarch_steps=32, dir_light_x=0.4, dir_light_y=-1.0, dir_light_z=0.3, **kw))
        o.set_emitters(em)
        o.reset()
        o.load_state(state)
        o.set_time(t, rt)
        o.run_stage(firesim.STAGE_LIGHT)
        frames[key] = o.render_hdr()
        o.close()
    np.testing.assert_array_equal(frames["off"], frames["zero"])
    d1 = frames["one"] - frames["off"]
    d2 = frames["two"] - frames["off"]
    assert d1[..., :3].min() >= -1e-15 and d1[..., :3].max() > 1e-4
    np.testing.assert_allclose(d2, 2.0 * d1, rtol=1e-9, atol=1e-14)


# ------------------------------------------------------------------ gizmos ----
def _camera(prm, fw, fh):
    yaw, pitch, dist = f(prm.camera_yaw), f(prm.camera_pitch), f(prm.camera_distance)
    cp = math.cos(pitch)
    eye = np.array([math.sin(yaw) * dist * cp, f(0.72) + math.sin(pitch) * dist, math.cos(yaw) * dist * cp])

    def norm(a):
        return a / math.sqrt(max(f(0.000001), float(a @ a)))

    fwd = norm(np.array([0.0, f(0.72), 0.0]) - eye)
    right = norm(np.cross(fwd, [0.0, 1.0, 0.0]))
    up = norm(np.cross(right, fwd))
    return dict(eye=eye, fwd=fwd, right=right, up=up, tan=f(0.50), aspect=f(fw / fh))


def _bytes(px):
    return (int(px) >> 16) & 0xFF, (int(px) >> 8) & 0xFF, int(px) & 0xFF


@pytest.mark.parametrize("active", [1, 3])
def test_gizmo_colours_at_full_alpha(active):
    """drawGizmos (:1931-1951) over a black frame: the pixel nearest the axis
    anchor (0.90, 0.84) is inside all three lines' full-alpha core, so the z
    axis's colour wins; pixels on the source ring (|dist - 0.125/depth| <=
    0.0015, outside the dot) and at its centre carry tint * 0.85 (tint by
    activeGizmo); with showGizmos 0 the frame stays black."""
    fw = fh = 1000
    kw = dict(camera_yaw=0.35, camera_pitch=0.10, camera_distance=2.8)
    o = _zero_scene((6, 6, 6), frame=(fw, fh), show_gizmos=1, active_gizmo=active, **kw)
    prm = o.get_params()
    hdr = np.zeros((fh, fw, 4))
    img = o.tonemap(hdr)
    assert _bytes(img[839, 899]) == (int(f(0.20) * 255), int(f(0.48) * 255), 255)
    cam = _camera(prm, fw, fh)
    src = np.array([f(prm.emitter_center_x), f(prm.emitter_height_norm) * f(2.03) + f(0.02), f(prm.emitter_center_z)])
    view = src - cam["eye"]
    depth = float(view @ cam["fwd"])
    cx = 0.5 + float(view @ cam["right"]) / (depth * cam["tan"] * cam["aspect"]) * 0.5
    cy = 0.5 - float(view @ cam["up"]) / (depth * cam["tan"]) * 0.5
    radius = f(0.125) / depth
    tint = (1.0, 0.34, 0.055) if active == 1 else (0.45, 0.18, 0.055)
    want = tuple(int(f(c) * f(0.85) * f(255.0)) for c in tint)
    ys, xs = np.mgrid[0:fh, 0:fw]
    dist = np.hypot((xs + 0.5) / fw - cx, (ys + 0.5) / fh - cy)
    ring = (np.abs(dist - radius) <= 0.0012) & (dist > 0.0125) & ((xs + 0.5) / fw < 0.8)
    assert ring.sum() > 100
    for y, x in zip(*np.nonzero(ring)):
        assert _bytes(img[y, x]) == want, (x, y)
    centre = np.unravel_index(np.argmin(dist), dist.shape)
    assert _bytes(img[centre]) == want
    o.close()
    o = _zero_scene((6, 6, 6), frame=(fw, fh), show_gizmos=0, **kw)
    assert np.all(o.tonemap(hdr) == 0xFF000000)
    o.close()


def test_gizmos_leave_debug_views_alone():
    o = _zero_scene((6, 6, 6), frame=(200, 100), show_gizmos=1, render_debug_mode=2)
    assert np.all(o.tonemap(np.zeros((100, 200, 4))) == 0xFF000000)
    o.close()


def _draw_gizmos_reference(rgb, fw, fh, prm):
    """drawGizmos (src/fire_cuda.cu:1909-1951) restated from the reference in
    numpy over a whole frame of unpacked colours (value / 255)."""
    cam = _camera(prm, fw, fh)
    ys, xs = np.mgrid[0:fh, 0:fw]
    uvx, uvy = (xs + 0.5) / fw, (ys + 0.5) / fh

    def ss(e0, e1, x):
        lo, hi = min(e0, e1), max(e0, e1)
        t = np.clip((x - lo) / max(f(0.000001), hi - lo), 0.0, 1.0)
        v = t * t * (3.0 - 2.0 * t)
        return 1.0 - v if e1 < e0 else v

    def blend(c, tint, a):
        a = np.clip(a, 0.0, 1.0)[..., None]
        return c + (np.array([f(t) for t in tint]) - c) * a

    def line(ax, ay, bx, by, th):
        pax, pay, bax, bay = uvx - ax, uvy - ay, bx - ax, by - ay
        h = np.clip((pax * bax + pay * bay) / max(f(0.000001), bax * bax + bay * bay), 0.0, 1.0)
        dx, dy = pax - bax * h, pay - bay * h
        return ss(th, th * f(0.35), np.sqrt(dx * dx + dy * dy))

    burner = prm.scene_id == 2 and prm.burner_center_count > 0
    src = np.array([f(prm.burner_center_x[0]), f(prm.burner_center_y[0]), f(prm.burner_center_z[0])]) if burner else \
        np.array([f(prm.emitter_center_x), f(prm.emitter_height_norm) * f(2.03) + f(0.02), f(prm.emitter_center_z)])
    view = src - cam["eye"]
    depth = float(view @ cam["fwd"])
    c = rgb.copy()
    if depth > f(0.02):
        cx = 0.5 + float(view @ cam["right"]) / (depth * cam["tan"] * cam["aspect"]) * 0.5
        cy = 0.5 - float(view @ cam["up"]) / (depth * cam["tan"]) * 0.5
        dist = np.sqrt((uvx - cx) ** 2 + (uvy - cy) ** 2)
        a = ss(f(0.006), f(0.006) * f(0.25), np.abs(dist - f(0.125) / depth)) + ss(f(0.012), f(0.012) * f(0.45), dist)
        c = blend(c, (1.0, 0.34, 0.055) if prm.active_gizmo == 1 else (0.45, 0.18, 0.055), a * f(0.85))
    bx, by = f(0.90), f(0.84)
    for axis, tint in ((0, (1.0, 0.12, 0.08)), (1, (0.22, 1.0, 0.30)), (2, (0.20, 0.48, 1.0))):
        tx, ty = bx + cam["right"][axis] * f(0.070), by - cam["up"][axis] * f(0.070)
        c = blend(c, tint, line(bx, by, tx, ty, f(0.005)))
    return c


@pytest.mark.parametrize("scene_id,active,yaw", [(3, 1, 0.35), (2, 4, -0.6), (0, 1, 1.2)])
def test_gizmos_match_reference_restatement(scene_id, active, yaw):
    """The overlay over a random frame: unpack the plain tonemap's bytes,
    apply the restated drawGizmos, pack with truncation; equal to the oracle's
    gizmo frame except where a channel sits within roundoff of a byte edge."""
    fw, fh = 240, 180
    # (a raised emitter, so the source ring's height term counts)
    kw = dict(scene_id=scene_id, burner_center_count=1 if scene_id == 2 else 0, camera_yaw=yaw, camera_pitch=0.15, camera_distance=2.4,
              emitter_height_norm=0.3)
    rng = np.random.default_rng(scene_id)
    hdr = np.concatenate([rng.uniform(0, 2, (fh, fw, 3)) ** 2, np.ones((fh, fw, 1))], -1)
    o = _zero_scene((6, 6, 6), frame=(fw, fh), show_gizmos=0, active_gizmo=active, **kw)
    plain = o.tonemap(hdr)
    o.close()
    o = _zero_scene((6, 6, 6), frame=(fw, fh), show_gizmos=1, active_gizmo=active, **kw)
    prm = o.get_params()
    got = o.tonemap(hdr)
    o.close()
    rgb = np.stack([(plain >> 16) & 0xFF, (plain >> 8) & 0xFF, plain & 0xFF], -1).astype(np.float64) / f(255.0)
    want_rgb = (np.clip(_draw_gizmos_reference(rgb, fw, fh, prm), 0.0, 1.0) * f(255.0)).astype(np.int64)
    got_rgb = np.stack([(got >> 16) & 0xFF, (got >> 8) & 0xFF, got & 0xFF], -1).astype(np.int64)
    changed = (got != plain).sum()
    assert changed > 100
    assert np.abs(got_rgb - want_rgb).max() <= 1
    assert (got_rgb == want_rgb).all(-1).mean() >= 0.999
