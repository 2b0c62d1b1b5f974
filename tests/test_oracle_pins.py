"""Pins of the oracle's passes beyond the primitives (CPU only).

The reference holds no field-level golden vectors, so each oracle function here
is pinned two ways:

  * closed forms of special cases that a one-constant slip, a dropped term or
    a swapped index fails (MacCormack limits, a pure heat gradient, uniform light
    propagation, soot-only extinction and shadow, grey camera response, bed
    masks at known points, room colour along fixed rays, Beer-Lambert
    transmittance of the raymarch);
  * an independent numpy restatement, written from src/fire_cuda.cu (not from
    the oracle), evaluated on random states, for the long per-cell chains
    (advect-react :934-1126 term by term, force :1128-1183, light build and
    propagate :1445-1565, metrics :1319-1418).

Reference literals are float32 (`f(...)` below) used in float64 arithmetic,
as the oracle does, so agreement is to roundoff (1e-12 relative).
The lattice noise and its primitives (sampler, fbm, blackbody) are pinned in
test_oracle.py and used here as building blocks.
"""
import ctypes
import math

import numpy as np
import pytest

import compare
import firesim
import oracle

L = oracle.load()
D3 = ctypes.c_double * 3


def f(x: float) -> float:
    """A reference float literal as the float64 value of its float32."""
    return float(np.float32(x))


def sat(x):
    return np.clip(x, 0.0, 1.0)


def ss(e0, e1, x):
    """smoothstepf (src/fire_cuda.cu:227-234): inverted edges give 1 - smoothstep."""
    inv = e1 < e0
    lo, hi = (e1, e0) if inv else (e0, e1)
    t = sat((np.asarray(x, dtype=np.float64) - lo) / max(f(0.000001), hi - lo))
    v = t * t * (3.0 - 2.0 * t)
    return 1.0 - v if inv else v


def lerp(a, b, t):
    return a + (b - a) * t


def lum(r, g, b):
    return r * f(0.2126) + g * f(0.7152) + b * f(0.0722)


def cell_coords(nx, ny, nz):
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return (x + 0.5) / nx, (y + 0.5) / ny, (z + 0.5) / nz


SCALARS = firesim.SCALAR_FIELDS  # heat fuel oxygen soot char ash pyrolysis progress turbulence soot_optics


def scene(grid, frame=(16, 12), emitters=(), **params):
    o = oracle.OracleScene(*grid, *frame)
    p = compare.emitter_params(**params)
    o.set_params(p)
    o.set_emitters(list(emitters))
    o.reset()
    return o, o.get_params()


# ------------------------------------------------------------- MacCormack ----
def test_maccormack_limits_and_correction():
    """sampleScalarMacCormack (:713-732): blend 0 is the first-order sample,
    blend 1 the corrected value clamped to the 8 corners of the sample cell,
    the correction is 1/2 (current - round trip), and it vanishes when the
    current value equals the round-trip sample."""
    rng = np.random.default_rng(11)
    nz, ny, nx = 5, 6, 7
    fld = rng.normal(size=(nz, ny, nx))
    dt = f(1 / 60)
    scale = (nx - 1, ny - 1, nz - 1)
    for trial in range(40):
        pu, pv, pw = rng.uniform(0.02, 0.98, 3)
        pvel = rng.uniform(-3, 3, 3)
        cur = float(rng.normal() * 3)
        pv3 = (ctypes.c_double * 3)(*pvel)
        first = oracle.sample_grid(fld, scale, pu, pv, pw)
        k = dt * f(0.62)
        rt = oracle.sample_grid(fld, scale, pu + pvel[0] * k, pv + pvel[1] * k, pw + pvel[2] * k)
        fx, fy, fz = pu * scale[0], pv * scale[1], pw * scale[2]
        x0, y0, z0 = int(math.floor(fx)), int(math.floor(fy)), int(math.floor(fz))
        corners = fld[z0:z0 + 2, y0:y0 + 2, x0:x0 + 2]
        lo, hi = corners.min(), corners.max()
        got0 = L.or_maccormack(oracle._dptr(fld), nx, ny, nz, dt, cur, pv3, pu, pv, pw, 0.0)
        got1 = L.or_maccormack(oracle._dptr(fld), nx, ny, nz, dt, cur, pv3, pu, pv, pw, 1.0)
        gotb = L.or_maccormack(oracle._dptr(fld), nx, ny, nz, dt, cur, pv3, pu, pv, pw, 0.37)
        assert got0 == pytest.approx(first, abs=1e-13)
        assert lo - 1e-13 <= got1 <= hi + 1e-13
        corrected = min(hi, max(lo, first + 0.5 * (cur - rt)))
        assert got1 == pytest.approx(corrected, abs=1e-13)
        assert gotb == pytest.approx(first + (corrected - first) * 0.37, abs=1e-13)
        same = L.or_maccormack(oracle._dptr(fld), nx, ny, nz, dt, rt, pv3, pu, pv, pw, 1.0)
        assert same == pytest.approx(first, abs=1e-13)
    # blend saturates: > 1 behaves as 1, < 0 as 0
    args = (oracle._dptr(fld), nx, ny, nz, dt, 5.0, (ctypes.c_double * 3)(1, 0, 0), 0.5, 0.5, 0.5)
    assert L.or_maccormack(*args, 7.0) == L.or_maccormack(*args, 1.0)
    assert L.or_maccormack(*args, -2.0) == L.or_maccormack(*args, 0.0)


# ----------------------------------------------------------- advect-react ----
def _react_reference(s, u, v, w, prm, emitters, cells, time, adv=None):
    """advectReactKernel (:934-1126) for scene 3 (no procedural bed: material
    0, so bed = 0), restated from the reference source. s: the scalar inputs
    (uniform values or per-cell arrays); adv: the MacCormack-advected values of
    the advected fields (default: s itself, exact for uniform fields, whose
    advection is the identity); u, v, w: cell-centre coordinates."""
    dt = f(prm.dt)
    h0, f0, o0, s0 = s["heat"], s["fuel"], s["oxygen"], s["soot"]
    ch0, a0, py0, pr0, k0, so0 = s["char"], s["ash"], s["pyrolysis"], s["progress"], s["turbulence"], s["soot_optics"]
    if adv is None:
        adv = s
    shape = u.shape
    x = np.broadcast_to
    # emitter membership (fs_emitter spheres, voxel units)
    xi, yi, zi = cells
    inside = [((xi + 0.5 - f(e[0])) ** 2 + (yi + 0.5 - f(e[1])) ** 2 + (zi + 0.5 - f(e[2])) ** 2) <= f(e[3]) ** 2 for e in emitters]
    any_em = np.zeros(shape, bool)
    for m in inside:
        any_em |= m
    # cold-air early-out (:970-1001)
    activity = h0 + f0 * f(0.18) + s0 * f(0.24) + py0 * f(0.20) + pr0 * f(0.14) + so0 * f(0.08) + k0 * f(0.06)
    sdx, sdy, sdz = u - f(prm.mouse_x), v - f(prm.mouse_y), w - f(0.5)
    outside = (prm.left_down == 0) | ((sdx * sdx + sdy * sdy + sdz * sdz) > f(0.090))
    early = (v > f(0.18)) & (activity < f(0.00035)) & (ch0 < f(0.00010)) & (a0 < f(0.00010)) & outside & ~any_em & (prm.right_down == 0)
    heat, fuel, oxygen, soot = (x(adv[n], shape).copy() for n in ("heat", "fuel", "oxygen", "soot"))
    char, ash = x(ch0, shape).copy(), x(a0, shape).copy()
    pyro = x(adv["pyrolysis"], shape) * f(0.55)
    progress, turb, sopt = (x(adv[n], shape).copy() for n in ("progress", "turbulence", "soot_optics"))
    # bed / pyrolysis / char (:1020-1045), scene 3: material 0, gasFeed = exposedChar * 0
    wx, wz = (u - f(0.50)) * f(2.10), (w - f(0.50)) * f(1.64)
    noise = np.vectorize(lambda a, b: L.or_fbm2(a, b, L.or_salt(prm.noise_seed, 1)))(
        wx * f(13.0) + wz * f(3.0) + time * f(0.38), v * f(31.0) + wz * f(11.0))
    source_noise = f(0.70) + f(0.48) * noise
    bed = 0.0
    tempK = f(293.0) + heat * f(335.0)
    radfb = ss(f(430.0), f(980.0), tempK)
    solid_o2 = sat(oxygen * f(1.28) - soot * f(0.030))
    flux = sat((tempK - f(405.0)) / f(900.0)) + bed * f(0.34) + sopt * f(0.024)
    char_rel = np.minimum(char, char * flux * solid_o2 * (f(0.14) + radfb * f(1.38)) * dt)
    pyro = pyro + (bed * (f(0.34) + radfb * f(1.05)) * dt + char_rel * (f(2.45) + source_noise * f(0.72))) * 1.0
    char = char + 0.0 * dt
    char = char - char_rel * f(0.92)
    ash = ash + char_rel * (f(0.16) + f(0.22) * sat(1.0 - oxygen))
    heat = heat + (bed * f(6.20) * dt * (1.0 + radfb * f(1.12)) + char_rel * f(3.35)) * 1.0
    fuel = fuel + pyro * f(3.70) * 1.0
    soot = soot + pyro * (f(0.090) + ash * f(0.032) + sat(1.0 - oxygen) * f(0.13)) * f(prm.smoke) * 1.0 * 1.0
    # interactive sources (:1047-1063)
    if prm.left_down:
        mx, my = f(prm.mouse_x), f(prm.mouse_y)
        src = np.exp(-((u - mx) ** 2) / f(0.0180) - ((w - f(0.5)) ** 2) / f(0.0550) - ((v - my) ** 2) / f(0.0020))
        heat = heat + src * dt * f(2.8)
        fuel = fuel + src * dt * f(1.4)
        pyro = pyro + src * dt * f(0.45)
        progress = progress + src * dt * f(0.75)
        turb = turb + src * dt * f(0.28)
        oxygen = np.maximum(oxygen, f(0.72) + src * f(0.20))
    if prm.right_down:
        mx, my = f(prm.mouse_x), f(prm.mouse_y)
        src = np.exp(-((u - mx) ** 2 + (w - f(0.5)) ** 2) / f(0.0100) - (v - my) ** 2 / f(0.0100))
        soot = soot + src * dt * f(8.0)
        heat = heat * (1.0 - src * dt * f(1.2))
        fuel = fuel * (1.0 - src * dt * f(0.7))
    for e, m in zip(emitters, inside):
        fuel = np.where(m, fuel + f(e[4]) * dt, fuel)
        heat = np.where(m, np.maximum(heat, (f(e[5]) - f(293.0)) / f(335.0)), heat)
    # combustion (:1065-1087)
    tempK = f(293.0) + heat * f(335.0)
    thermal = ss(f(575.0), f(1220.0), tempK)
    arr = np.exp(-f(3800.0) / np.maximum(f(640.0), tempK))
    oxy_lim = np.minimum(fuel, oxygen * f(1.42))
    burn = np.minimum(oxy_lim, oxy_lim * thermal * arr * (f(118.0) + f(prm.intensity) * f(34.0)) * dt)
    incomplete = sat(1.0 - oxygen * f(1.30) + fuel * f(0.065) + soot * f(0.015))
    release = burn * f(7.0)
    yield_ = burn * (f(0.035) + incomplete * f(0.34)) * f(prm.smoke) * 1.0
    oxidation = soot * oxygen * ss(f(760.0), f(1540.0), tempK) * dt * f(0.82)
    front = burn * (f(9.8) + turb * f(3.4)) + pyro * f(0.70) + char_rel * f(2.10)
    bA = sat(ss(f(0.70), f(0.96), v) + ss(f(0.11), f(0.03), u) + ss(f(0.89), f(0.97), u) + ss(f(0.11), f(0.03), w) + ss(f(0.89), f(0.97), w))
    upper = ss(f(0.52), f(0.98), v)
    heat = heat + release
    fuel = fuel - burn * f(1.20)
    oxygen = oxygen - burn * f(0.98)
    soot = soot + yield_ - oxidation
    progress = (progress + front) * np.exp(-dt * (f(0.72) + v * f(0.45) + oxygen * f(0.16) + bA * f(1.05) + upper * f(0.26)))
    oxygen = lerp(oxygen, 1.0, bA * dt * f(1.7))
    oxygen = oxygen + (1.0 - oxygen) * dt * f(0.08)
    # radiation and decays (:1089-1097)
    tempK = f(293.0) + heat * f(335.0)
    rb = np.maximum(0.0, (tempK - f(293.0)) / f(1560.0))
    heat = heat - (rb * rb) ** 2 * dt * (f(0.88) + soot * f(0.085) + sopt * f(0.018))
    heat = heat * np.exp(-dt * (f(0.31) + v * f(0.64) + soot * f(0.022) + bA * f(0.26) + upper * f(0.18)))
    fuel = fuel * np.exp(-dt * (f(0.58) + heat * f(0.20) + v * f(0.92) + bA * f(0.18)))
    soot = soot * np.exp(-dt * (f(0.060) + v * f(0.092) + oxygen * f(0.028) + bA * f(0.96) + upper * f(0.22)))
    pyro = pyro * np.exp(-dt * (f(1.55) + v * f(0.58) + bA * f(0.18)))
    # shear and turbulence energy (:1099-1111); s["shear"] is |du|+|dv|+|dw| of the cell
    activity_f = sat(progress * f(0.64) + burn * f(5.2) + heat * f(0.065))
    turb = turb + dt * (s.get("shear", 0.0) * (f(0.72) + progress * f(0.06)) + activity_f * f(prm.turbulence) * f(0.62) + front * f(0.085) + soot * f(0.018))
    turb = turb * np.exp(-dt * (f(0.54) + v * f(0.24) + bA * f(0.42) + upper * f(0.16)))
    growth = sat(incomplete * f(0.36) + ash * f(0.13) + soot * f(0.055) + ss(f(0.65), f(1.55), tempK / f(1000.0)) * f(0.18))
    sopt = soot * (f(1.04) + f(0.28) * np.sqrt(np.maximum(0.0, heat)) + f(0.24) * ash + f(0.46) * growth)
    sopt = lerp(sopt, sopt * f(0.48), sat(oxidation * f(0.92)))
    out = dict(heat=np.clip(heat, 0, f(8.8)), fuel=np.clip(fuel, 0, f(5.5)), oxygen=np.clip(oxygen, 0, 1.0), soot=np.clip(soot, 0, f(7.4)),
               char=np.clip(char, 0, f(2.8)), ash=np.clip(ash, 0, f(1.9)), pyrolysis=np.clip(pyro, 0, f(3.6)),
               progress=np.clip(progress, 0, f(3.4)), turbulence=np.clip(turb, 0, f(4.2)), soot_optics=np.clip(sopt, 0, f(10.0)))
    # early-out cells (:984-1001)
    eo = dict(heat=0.0, fuel=0.0, oxygen=np.minimum(1.0, o0 + (1.0 - o0) * dt * f(0.08)), soot=0.0, char=ch0, ash=a0, pyrolysis=0.0,
              progress=0.0, turbulence=k0 * max(0.0, 1.0 - dt * f(0.54)), soot_optics=0.0)
    for k in out:
        out[k] = np.where(early, eo[k], out[k])
    return out, float(np.where(early, 0.0, release).sum())


def react(state, grid=(6, 10, 6), emitters=(), shear=0.0, **params):
    """Runs the oracle's advect-react on uniform inputs; returns its fields,
    the heat release, and the restatement's fields and heat release."""
    o, prm = scene(grid, emitters=emitters, **params)
    nx, ny, nz = grid
    for name in SCALARS:
        o.write(name, np.full(o.shape(firesim.FIELD[name]), float(state.get(name, 1.0 if name == "oxygen" else 0.0))))
    if shear:
        # u faces linear in x: |uR - uL| = shear in every cell; v, w zero
        zu, yu, xu = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx + 1), indexing="ij")
        o.write("u", shear * xu.astype(np.float64))
    o.set_time(0.35, 0.35)
    o.run_stage(firesim.STAGE_ADVECT_REACT)
    got = {n: o.read(n) for n in SCALARS}
    rel = o.metrics()["heat_release_step"]
    u, v, w = cell_coords(nx, ny, nz)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    st = {n: float(state.get(n, 1.0 if n == "oxygen" else 0.0)) for n in SCALARS}
    st["shear"] = abs(shear)
    want, wrel = _react_reference(st, u, v, w, prm, list(emitters), (x, y, z), f(0.35))
    o.close()
    return got, rel, want, wrel


HOT = (1500.0 - 293.0) / 335.0
# Cases isolating one term each: all other inputs zero (oxygen 1 where the
# term needs none). Ids name the term of advectReactKernel they exercise.
REACT_CASES = {
    # burnMass, heat release 7B, fuel -1.2B, oxygen -0.98B, fuel-limited branch
    "burn_fuel_limited": dict(heat=HOT, fuel=0.2, oxygen=1.0),
    # oxygen-limited branch: min(fuel, 1.42 O) = 1.42 O
    "burn_oxygen_limited": dict(heat=HOT, fuel=4.0, oxygen=0.05),
    # soot yield B (0.035 + incomplete 0.34) smoke, incomplete from O, fuel, soot
    "soot_yield": dict(heat=2.2, fuel=0.9, oxygen=0.3),
    # soot oxidation soot O smoothstep(760, 1540, T) dt 0.82 and the optics lerp
    "soot_oxidation": dict(heat=3.0, soot=0.8, oxygen=0.9),
    # radiative loss ((T-293)/1560)^4 dt (0.88 + 0.085 soot + 0.018 optics)
    "radiative_loss": dict(heat=6.5, oxygen=1.0, soot_optics=0.4),
    # heat, fuel, soot, pyrolysis decay exponents (v, boundary air, upper)
    "decays": dict(heat=0.5, fuel=0.4, soot=0.3, pyrolysis=0.6, oxygen=1.0),
    # char release heat flux, solid O2, radiative feedback; ash and pyrolysis gains
    "char_ash_release": dict(heat=2.5, char=0.7, ash=0.2, oxygen=0.8),
    # turbulence energy: flame activity, front production, soot; decay
    "turbulence_energy": dict(heat=1.2, fuel=0.3, progress=0.9, turbulence=0.5, oxygen=0.7, soot=0.2),
    # soot optics growth: incomplete, ash, soot, smoothstep(0.65,1.55,T/1000), sqrt(heat)
    "soot_optics": dict(heat=2.8, soot=1.5, ash=0.4, oxygen=0.25, fuel=0.6),
    # progress front decay (oxygen, boundary air, upper entrainment)
    "progress_decay": dict(heat=0.8, progress=1.7, oxygen=0.6),
    # cold-air early-out (tiny activity) against the live path below v = 0.18
    "early_out": dict(oxygen=0.4, turbulence=0.004, char=0.00005, ash=0.00002),
    # clamps at the caps (:1116-1125)
    "clamps": dict(heat=8.79, fuel=5.49, soot=7.39, char=2.79, ash=1.89, pyrolysis=3.59, progress=3.39, turbulence=4.19,
                   soot_optics=9.99, oxygen=1.0),
}


@pytest.mark.parametrize("case", sorted(REACT_CASES))
def test_react_terms_match_restatement(case):
    got, rel, want, wrel = react(REACT_CASES[case], intensity=1.3, smoke=0.8, turbulence=0.9)
    for n in SCALARS:
        assert np.allclose(got[n], want[n], rtol=1e-12, atol=1e-15), (case, n, np.abs(got[n] - want[n]).max())
    assert rel == pytest.approx(wrel, rel=1e-12, abs=1e-300)


def test_react_shear_mouse_and_emitters_match_restatement():
    st = dict(heat=1.1, fuel=0.5, oxygen=0.9, soot=0.2, progress=0.3, turbulence=0.2)
    got, rel, want, wrel = react(st, shear=0.35, turbulence=1.1)
    for n in SCALARS:
        assert np.allclose(got[n], want[n], rtol=1e-12, atol=1e-15), ("shear", n)
    # mouse heat/fuel source and soot source, with an emitter sphere
    em = [(2.5, 3.0, 3.0, 2.2, 1.3, 1400.0)]
    for left, right in ((1, 0), (0, 1), (1, 1)):
        got, rel, want, wrel = react(dict(heat=0.3, fuel=0.2, oxygen=0.95), emitters=em, left_down=left, right_down=right,
                                     mouse_x=0.4, mouse_y=0.3)
        for n in SCALARS:
            assert np.allclose(got[n], want[n], rtol=1e-12, atol=1e-15), ((left, right), n, np.abs(got[n] - want[n]).max())
        assert rel == pytest.approx(wrel, rel=1e-12)


def test_react_conservation_bookkeeping():
    """Heat release is 7 burnMass (:1071): the summed release equals 7x the
    fuel burned, where fuel drops by 1.2 burnMass (:1078) before its decay."""
    grid = (4, 4, 4)
    st = dict(heat=HOT, fuel=0.3, oxygen=1.0)
    o, prm = scene(grid)
    for n in SCALARS:
        o.write(n, np.full(o.shape(firesim.FIELD[n]), st.get(n, 0.0)))
    o.run_stage(firesim.STAGE_ADVECT_REACT)
    release = o.metrics()["heat_release_step"]
    dt = f(prm.dt)
    u, v, w = cell_coords(*grid)
    bA = sat(ss(f(0.70), f(0.96), v) + ss(f(0.11), f(0.03), u) + ss(f(0.89), f(0.97), u) + ss(f(0.11), f(0.03), w) + ss(f(0.89), f(0.97), w))
    fuel = o.read("fuel")
    heat = o.read("heat")
    # undo the fuel decay exp(-dt(0.58 + 0.20 heat + 0.92 v + 0.18 bA)) to recover fuel - 1.2 B
    before_decay = fuel / np.exp(-dt * (f(0.58) + heat * f(0.20) + v * f(0.92) + bA * f(0.18)))
    burned = (0.3 - before_decay) / f(1.20)
    assert release == pytest.approx(7.0 * burned.sum(), rel=1e-10)
    o.close()


# ------------------------------------------------------------------ force ----
def _force_reference(H, S, K, C, U, V, W, prm, time):
    """forceVelocityKernel (:1128-1183) contributions, gathered onto faces in
    the build's fixed order, then velocityBoundaryKernel (:1261-1291) and
    divergenceKernel (:1185-1199)."""
    nz, ny, nx = H.shape
    u, v, w = cell_coords(nx, ny, nz)
    dt = f(prm.dt)
    T = f(prm.turbulence)

    def at(a, dx, dy, dz):  # clamped neighbour, the reference's scalarIndex clamps
        zi, yi, xi = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        return a[np.clip(zi + dz, 0, nz - 1), np.clip(yi + dy, 0, ny - 1), np.clip(xi + dx, 0, nx - 1)]

    gx, gy, gz = at(H, 1, 0, 0) - at(H, -1, 0, 0), at(H, 0, 1, 0) - at(H, 0, -1, 0), at(H, 0, 0, 1) - at(H, 0, 0, -1)
    uL, uR = U[:, :, :-1], U[:, :, 1:]
    vD, vU = V[:, :-1, :], V[:, 1:, :]
    wB, wF = W[:-1], W[1:]
    cx, cy, cz = (wF - wB) - (vU - vD), (uR - uL) - (wF - wB), (vU - vD) - (uR - uL)
    cmag = np.abs(cx) + np.abs(cy) + np.abs(cz)
    salt = L.or_salt(prm.noise_seed, 1)
    na = np.vectorize(lambda a, b: L.or_fbm2(a, b, salt))(u * f(9.0) + w * f(3.7) + time * f(0.22), v * f(13.0) - time * f(0.31)) - f(0.5)
    nb = np.vectorize(lambda a, b: L.or_fbm2(a, b, salt))(w * f(10.0) - time * f(0.19), u * f(8.0) + v * f(5.5) + time * f(0.27)) - f(0.5)
    thermal = ss(f(0.04), f(4.20), H + S * f(0.28))
    drag = ss(f(0.26), f(0.94), v)
    les = np.sqrt(np.maximum(0.0, K)) * (f(0.30) + C * f(0.54))
    curl_gain = T * 1.0 * (f(0.18) + thermal * f(0.82) + les * f(1.12))
    vortex = T * 1.0 * thermal * sat(cmag * f(0.46)) * f(0.68)
    fade = ss(f(0.96), f(0.08), v)
    damp = 0.5 * (vD + vU) * (f(0.10) + drag * f(0.72) + S * f(0.020))
    buoy = H * 1.0 * (f(1.62) + T * f(0.34) + C * f(0.16)) * fade - S * f(0.095) - gy * f(0.18) - damp
    prof = f(0.34) + v * f(0.96)
    fv = dt * (buoy + cy * vortex * f(0.22))
    fu = dt * (f(prm.wind_x) * prof + (-gz * f(1.36) + na * (f(0.20) + drag * f(0.16)) + cx * vortex) * curl_gain)
    fw = dt * (f(prm.wind_z) * prof + (gx * f(1.36) + nb * (f(0.20) + drag * f(0.16)) + cz * vortex) * curl_gain)
    Uo, Vo, Wo = U.copy(), V.copy(), W.copy()
    Uo[:, :, :-1] += fu
    Uo[:, :, 1:] += fu
    Vo[:, 1:, :] += fv
    Wo[:-1] += fw
    Wo[1:] += fw
    Uo = np.clip(Uo, f(-4.0), f(4.0))
    Uo[:, :, [0, -1]] = 0.0
    Vo = np.clip(Vo, f(-0.9), f(6.4))
    Vo[:, 0, :] = 0.0
    Wo = np.clip(Wo, f(-4.0), f(4.0))
    Wo[[0, -1]] = 0.0
    div = np.diff(Uo, axis=2) + np.diff(Vo, axis=1) + np.diff(Wo, axis=0)
    return Uo, Vo, Wo, -div


def test_force_matches_restatement_on_random_fields():
    grid = (9, 11, 7)
    rng = np.random.default_rng(21)
    o, prm = scene(grid, turbulence=0.8, wind_x=0.3, wind_z=-0.45, noise_seed=5)
    nx, ny, nz = grid
    H = rng.uniform(0, 5, (nz, ny, nx))
    S, K, C = rng.uniform(0, 2, (nz, ny, nx)), rng.uniform(0, 1.5, (nz, ny, nx)), rng.uniform(0, 1.2, (nz, ny, nx))
    # faces reach past the clamps u, w in [-4, 4] and v in [-0.9, 6.4] (:1270-1290)
    U = rng.uniform(-4.3, 4.3, (nz, ny, nx + 1))
    V = rng.uniform(-1.1, 6.7, (nz, ny + 1, nx))
    W = rng.uniform(-4.3, 4.3, (nz + 1, ny, nx))
    for n, a in (("heat", H), ("soot", S), ("turbulence", K), ("progress", C), ("u", U), ("v", V), ("w", W)):
        o.write(n, a)
    o.set_time(1.25, 1.25)
    o.run_stage(firesim.STAGE_FORCE)
    Uo, Vo, Wo, d = _force_reference(H, S, K, C, U, V, W, prm, float(np.float32(1.25)))
    assert np.allclose(o.read("u"), Uo, rtol=1e-12, atol=1e-14)
    assert np.allclose(o.read("v"), Vo, rtol=1e-12, atol=1e-14)
    assert np.allclose(o.read("w"), Wo, rtol=1e-12, atol=1e-14)
    assert np.allclose(o.read("divergence"), d, rtol=1e-12, atol=1e-13)
    assert np.all(o.read("pressure") == 0)
    o.close()


def test_force_pure_heat_gradient_closed_form():
    """With heat saturating the thermal gate (h >= 4.2), no soot, turbulence
    energy or velocity, the noise and vortex terms are the same for heat
    ramps of opposite sign along z (x), so the u (w) face difference is the
    gradient term alone: -2 dt T 1.36 (gradZ+ - gradZ-) on interior faces,
    since each u face receives two cell contributions (:1178-1180)."""
    grid = (6, 5, 8)
    nx, ny, nz = grid
    T, a = 0.9, 0.05
    out = {}
    for sign in (1, -1):
        for axis in ("z", "x"):
            o, prm = scene(grid, turbulence=T)
            zi, yi, xi = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
            ramp = (zi if axis == "z" else xi).astype(np.float64)
            o.write("heat", 5.0 + sign * a * ramp)
            o.run_stage(firesim.STAGE_FORCE)
            out[(sign, axis)] = (o.read("u"), o.read("w"))
            o.close()
    dt, Tf = f(1 / 60), f(T)
    # gradZ = 2a (interior z), a at z = 0 and nz-1 (clamped neighbour)
    gz = np.full(nz, 2 * a)
    gz[[0, -1]] = a
    du = out[(1, "z")][0] - out[(-1, "z")][0]
    want = 2 * dt * Tf * (-f(1.36)) * (2 * gz)  # (gradZ+ - gradZ-) = 2 gz
    assert np.allclose(du[:, :, 1:-1], want[:, None, None], rtol=1e-12)
    assert np.all(du[:, :, [0, -1]] == 0)
    gx = np.full(nx, 2 * a)
    gx[[0, -1]] = a
    dw = out[(1, "x")][1] - out[(-1, "x")][1]
    want_w = 2 * dt * Tf * f(1.36) * (2 * gx)
    assert np.allclose(dw[1:-1], want_w[None, None, :], rtol=1e-12)


# ------------------------------------------------------------------ light ----
def _sample(fld, u, v, w):
    nz, ny, nx = fld.shape
    return oracle.sample_grid(fld, (nx - 1, ny - 1, nz - 1), u, v, w)


def _light_build_reference(F, lnx, lny, lnz, prm, time):
    """buildSceneLightKernel (:1445-1525) at every light cell."""
    nz, ny, nx = F["heat"].shape
    salt3 = L.or_salt(prm.noise_seed, 2)
    light = np.zeros((lnz, lny, lnx, 4))
    shadow = np.zeros((lnz, lny, lnx))
    bb = D3()
    cX, cY, cZ = f(1.0 / nx), f(1.0 / ny), f(1.0 / nz)
    for lz in range(lnz):
        for ly in range(lny):
            for lx in range(lnx):
                u, v, w = (lx + 0.5) / lnx, (ly + 0.5) / lny, (lz + 0.5) / lnz
                fade = L.or_domain_fade(u, v, w)
                s = {n: _sample(F[n], u, v, w) for n in F}
                gx = _sample(F["heat"], u + cX, v, w) - _sample(F["heat"], u - cX, v, w)
                gy = _sample(F["heat"], u, v + cY, w) - _sample(F["heat"], u, v - cY, w)
                gz = _sample(F["heat"], u, v, w + cZ) - _sample(F["heat"], u, v, w - cZ)
                front = math.sqrt(gx * gx + gy * gy + gz * gz) + s["fuel"] * s["oxygen"] * f(0.035) + s["progress"] * f(0.16)
                fine = L.or_fbm3(u * f(34.0) + time * f(0.45), v * f(42.0) - time * f(0.82), w * f(34.0) + time * f(0.22), salt3)
                shear = L.or_fbm3(u * f(12.0) - time * f(0.38), v * f(11.0) - time * f(0.62), w * f(13.0) + time * f(0.20), salt3)
                rf = ss(f(0.035), f(0.54), front + s["pyrolysis"] * f(0.10))
                hfade = ss(f(0.70), f(0.070), v)
                lower = ss(f(0.62), f(0.06), v)
                plume = ss(f(0.12), f(0.88), v)
                gate = ss(f(0.56), f(0.94), fine * f(0.48) + shear * f(0.42) + rf * f(0.26) + s["heat"] * f(0.018)) * ss(f(0.04), f(0.86), rf) * hfade
                core = sat(gate * (f(0.42) + s["progress"] * f(0.22) + s["pyrolysis"] * f(0.34)))
                tK = f(293.0) + s["heat"] * f(360.0) + s["fuel"] * f(44.0) + s["pyrolysis"] * f(120.0) + s["progress"] * f(70.0) + core * f(1020.0)
                flame_e = (fade * ss(f(820.0), f(1880.0), tK) * ss(f(0.010), f(0.68), s["fuel"] + s["pyrolysis"] * f(0.30) + s["heat"] * f(0.040)) *
                           ss(f(0.030), f(0.82), s["oxygen"]) * rf * gate * (f(0.24) + core * f(1.42)))
                ember = ss(f(0.002), f(0.22), s["char"] + s["ash"] * f(0.18)) * ss(f(0.09), f(0.00), v) * (f(0.10) + s["heat"] * f(0.012) + s["pyrolysis"] * f(0.050))
                att = math.exp(-(s["soot_optics"] * f(0.25) + s["soot"] * f(0.090) + s["ash"] * f(0.035)))
                L.or_blackbody(tK, bb)
                k = flame_e * att * f(3.40)
                coal = ember * f(prm.reflection_gain) * f(0.20)
                ext = sat(s["soot_optics"] * (f(0.26) + plume * f(0.42)) + s["soot"] * (f(0.065) + plume * f(0.20)) + s["ash"] * f(0.050) +
                          s["progress"] * lower * f(0.020))
                light[lz, ly, lx] = (bb[0] * k + 1.0 * coal, bb[1] * k + f(0.22) * coal, bb[2] * k + f(0.040) * coal, ext)
                vdepth = v * (1.0 + plume * f(0.70))
                radial = math.exp(-((u - f(0.50)) ** 2 * f(5.5) + (w - f(0.50)) ** 2 * f(7.0)))
                optical = s["soot_optics"] * (f(0.62) + plume * f(1.15)) + s["soot"] * (f(0.22) + plume * f(0.46)) + s["ash"] * f(0.072)
                shadow[lz, ly, lx] = math.exp(-(optical * (f(0.72) + vdepth * f(0.42)) + radial * optical * f(0.34) + s["progress"] * f(0.035)))
    return light, shadow


def _propagate_reference(light, shadow, diffusion, decay):
    """propagateSceneLightKernel (:1527-1565)."""
    lnz, lny, lnx, _ = light.shape
    out = np.zeros_like(light)
    for lz in range(lnz):
        for ly in range(lny):
            for lx in range(lnx):
                nb, wsum = np.zeros(3), 0.0
                for dx, dy, dz in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
                    q = light[min(max(lz + dz, 0), lnz - 1), min(max(ly + dy, 0), lny - 1), min(max(lx + dx, 0), lnx - 1)]
                    occ = math.exp(-q[3] * f(0.42))
                    nb += q[:3] * occ
                    wsum += occ
                nb /= max(f(0.0001), wsum)
                direct = light[lz, ly, lx, :3]
                bounced = direct + (nb - direct) * diffusion
                lift = ss(f(0.10), f(0.00), (ly + 0.5) / lny)
                contact = np.array([1.0, f(0.40), f(0.12)]) * lift * lum(*direct) * f(0.18)
                out[lz, ly, lx, :3] = bounced * decay * (f(0.66) + shadow[lz, ly, lx] * f(0.34)) + contact
                out[lz, ly, lx, 3] = light[lz, ly, lx, 3]
    return out


def _smooth_fields(grid, rng):
    """A random-amplitude fire blob low in the domain: non-trivial flame gates,
    coal glow at the floor and soot above, oxygen depleted inside."""
    nx, ny, nz = grid
    u, v, w = cell_coords(nx, ny, nz)
    blob = np.exp(-((u - 0.5) ** 2 * 8 + (v - 0.25) ** 2 * 10 + (w - 0.5) ** 2 * 8))
    scale = dict(heat=6.0, fuel=2.0, soot=1.5, char=0.5, ash=0.3, pyrolysis=1.0, progress=1.5, turbulence=0.8, soot_optics=1.2)
    F = {n: blob * k * rng.uniform(0.6, 1.0, blob.shape) for n, k in scale.items()}
    F["oxygen"] = np.clip(1.0 - 0.6 * blob + rng.uniform(0, 0.1, blob.shape), 0, 1)
    return F


def test_light_build_and_propagate_match_restatement():
    grid = (10, 12, 8)
    rng = np.random.default_rng(31)
    o, prm = scene(grid, noise_seed=9, reflection_gain=1.3)
    F = _smooth_fields(grid, rng)
    for n, a in F.items():
        o.write(n, a)
    o.set_time(0.8, 0.8)
    L.or_light_build(o.h)
    lnx, lny, lnz = firesim.light_dims(*grid)
    want_l, want_s = _light_build_reference(F, lnx, lny, lnz, prm, float(np.float32(0.8)))
    got_l = o.read("light").reshape(lnz, lny, lnx, 4)
    got_s = o.read("shadow").reshape(lnz, lny, lnx)
    assert want_l[..., :3].max() > 0.01, "the fields must light some cells"
    assert np.allclose(got_l, want_l, rtol=1e-11, atol=1e-14)
    assert np.allclose(got_s, want_s, rtol=1e-11, atol=1e-14)
    L.or_light_propagate(o.h, f(0.48), f(0.92))
    p1 = _propagate_reference(want_l, want_s, f(0.48), f(0.92))
    assert np.allclose(o.read("light").reshape(p1.shape), p1, rtol=1e-11, atol=1e-14)
    # the full pass is build, then (0.48, 0.92), then (0.34, 0.96) (:3313-3314)
    o.run_stage(firesim.STAGE_LIGHT)
    p2 = _propagate_reference(p1, want_s, f(0.34), f(0.96))
    assert np.allclose(o.read("light").reshape(p2.shape), p2, rtol=1e-11, atol=1e-14)
    o.close()


def test_light_build_on_a_warm_even_field_matches_restatement():
    """A hot field with almost no heat gradient (heat 2 with a 2% ripple): the
    reaction front is then carried by fuel x oxygen and progress alone and sits
    on the ramps of its smoothsteps, and of the flame sheet's gate on it, in
    cells hot enough to emit, which a fire blob (steep edges, saturated front
    wherever it is hot) never shows."""
    grid = (10, 12, 8)
    rng = np.random.default_rng(33)
    o, prm = scene(grid, noise_seed=9, reflection_gain=1.3)
    shape = (grid[2], grid[1], grid[0])
    level = dict(heat=2.0, fuel=1.0, oxygen=0.5, soot=0.2, char=0.1, ash=0.05, pyrolysis=0.3, progress=0.9, turbulence=0.3, soot_optics=0.4)
    F = {n: v * (1.0 + 0.02 * rng.uniform(-1, 1, shape)) for n, v in level.items()}
    for n, a in F.items():
        o.write(n, a)
    o.set_time(0.8, 0.8)
    L.or_light_build(o.h)
    lnx, lny, lnz = firesim.light_dims(*grid)
    want_l, want_s = _light_build_reference(F, lnx, lny, lnz, prm, float(np.float32(0.8)))
    got_l = o.read("light").reshape(lnz, lny, lnx, 4)
    got_s = o.read("shadow").reshape(lnz, lny, lnx)
    assert want_l[..., :3].max() > 0.01, "the fields must light some cells"
    assert np.allclose(got_l, want_l, rtol=1e-11, atol=1e-14)
    assert np.allclose(got_s, want_s, rtol=1e-11, atol=1e-14)
    o.close()


def test_light_propagate_uniform_closed_form():
    """A uniform light field (rgb c, extinction e) with uniform shadow s: the
    occlusion-weighted neighbour mean is c, so next = c decay (0.66 + 0.34 s)
    + (1, 0.40, 0.12) smoothstep(0.10, 0, v) lum(c) 0.18; extinction is kept."""
    grid = (8, 8, 8)
    o, prm = scene(grid)
    lnx, lny, lnz = firesim.light_dims(*grid)
    c = np.array([0.7, 0.3, 0.1])
    e, s = 0.4, 0.6
    light = np.zeros((lnz, lny, lnx, 4))
    light[..., :3] = c
    light[..., 3] = e
    o.write("light", light)
    o.write("shadow", np.full((lnz, lny, lnx), s))
    L.or_light_propagate(o.h, 0.48, 0.92)
    got = o.read("light").reshape(lnz, lny, lnx, 4)
    lift = ss(f(0.10), f(0.00), (np.arange(lny) + 0.5) / lny)
    lc = c[0] * f(0.2126) + c[1] * f(0.7152) + c[2] * f(0.0722)
    for ch, tint in enumerate((1.0, f(0.40), f(0.12))):
        want = c[ch] * 0.92 * (f(0.66) + s * f(0.34)) + tint * lift * lc * f(0.18)
        assert np.allclose(got[..., ch], want[None, :, None], rtol=1e-13)
    assert np.all(got[..., 3] == e)
    assert lift[0] > 0 and lift[-1] == 0  # only the floor rows get the contact bounce
    o.close()


@pytest.mark.parametrize("field,value", [("soot", 0.9), ("soot_optics", 0.7), ("ash", 0.01), ("progress", 1.3)])
def test_light_build_absorber_only_closed_form(field, value):
    """One uniform absorber, no oxygen: no flame (oxygen gate) and no coal
    (char + 0.18 ash < 0.002), so radiance is 0; extinction and shadow are the
    reference's closed forms in v, u, w (:1506-1524)."""
    grid = (12, 14, 10)
    o, prm = scene(grid)
    for n in SCALARS:
        o.write(n, np.zeros(o.shape(firesim.FIELD[n])))
    o.write(field, np.full(o.shape(firesim.FIELD[field]), value))
    L.or_light_build(o.h)
    lnx, lny, lnz = firesim.light_dims(*grid)
    got = o.read("light").reshape(lnz, lny, lnx, 4)
    sh = o.read("shadow").reshape(lnz, lny, lnx)
    assert np.all(got[..., :3] == 0)
    u, v, w = cell_coords(lnx, lny, lnz)
    plume, lower = ss(f(0.12), f(0.88), v), ss(f(0.62), f(0.06), v)
    so = value if field == "soot_optics" else 0.0
    sd = value if field == "soot" else 0.0
    ash = value if field == "ash" else 0.0
    pr = value if field == "progress" else 0.0
    ext = sat(so * (f(0.26) + plume * f(0.42)) + sd * (f(0.065) + plume * f(0.20)) + ash * f(0.050) + pr * lower * f(0.020))
    optical = so * (f(0.62) + plume * f(1.15)) + sd * (f(0.22) + plume * f(0.46)) + ash * f(0.072)
    vdepth = v * (1.0 + plume * f(0.70))
    radial = np.exp(-((u - f(0.50)) ** 2 * f(5.5) + (w - f(0.50)) ** 2 * f(7.0)))
    shadow = np.exp(-(optical * (f(0.72) + vdepth * f(0.42)) + radial * optical * f(0.34) + pr * f(0.035)))
    assert np.allclose(got[..., 3], ext, rtol=1e-12, atol=1e-15)
    assert np.allclose(sh, shadow, rtol=1e-12)
    o.close()


def test_light_build_coal_glow_closed_form():
    """Char only (no oxygen): the ember emitter smoothstep(0.002, 0.22, char)
    smoothstep(0.09, 0, v) 0.10 lights the floor rows with (1, 0.22, 0.04)
    reflectionGain 0.20 (:1497-1505); rows above v = 0.09 stay dark."""
    grid = (12, 40, 10)
    o, prm = scene(grid, reflection_gain=1.7)
    for n in SCALARS:
        o.write(n, np.zeros(o.shape(firesim.FIELD[n])))
    o.write("char", np.full(o.shape(firesim.FIELD["char"]), 0.1))
    L.or_light_build(o.h)
    lnx, lny, lnz = firesim.light_dims(*grid)
    got = o.read("light").reshape(lnz, lny, lnx, 4)
    v = (np.arange(lny) + 0.5) / lny
    ember = ss(f(0.002), f(0.22), 0.1) * ss(f(0.09), f(0.00), v) * f(0.10)
    for ch, tint in enumerate((1.0, f(0.22), f(0.040))):
        assert np.allclose(got[..., ch], (tint * ember * f(1.7) * f(0.20))[None, :, None], rtol=1e-12, atol=1e-16)
    assert got[:, 0, :, 0].min() > 0 and np.all(got[:, v >= 0.09, :, :3] == 0)
    o.close()


# --------------------------------------------------------- camera response ----
GAMMA = f(1.0 / f(2.2))  # powf(c, 1.0f / 2.2f): the float quotient (:2271-2273)

def _aces(x):
    return min(1.0, max(0.0, x * (f(2.51) * x + f(0.03)) / (x * (f(2.43) * x + f(0.59)) + f(0.14))))


@pytest.mark.parametrize("grey,exposure", [(0.02, 1.0), (0.005, 0.84), (0.3, 1.0), (0.12, 1.7), (1.9, 1.0), (6.0, 0.84), (0.9, 2.0)])
def test_camera_response_grey_closed_form(grey, exposure):
    """cameraResponse (:2261-2274) on a grey input: luminance weights sum to 1,
    so the hue-preserving ACES gives L = aces(g e); then the toe lerp
    (0.82 below luma 0.055), the shoulder (1/max(1, L/0.94) weighted 0.10 by
    smoothstep(0.78, 1.08, L)), the highlight lift (1 + 0.28 smoothstep(0.56,
    0.90, L)) and gamma 1/2.2. Inputs cover toe, mid range and shoulder."""
    e = f(exposure)
    wsum = f(0.2126) + f(0.7152) + f(0.0722)
    l0 = grey * e * wsum
    c = grey * e * (_aces(l0) / l0)  # every channel
    luma = c * wsum  # the toe, shoulder and lift gates all use this luma (:2264-2270)
    toe = ss(0.0, f(0.055), luma)
    c = c * f(0.82) + (c - c * f(0.82)) * toe
    sh = ss(f(0.78), f(1.08), luma)
    c = c + (c / max(1.0, luma / f(0.94)) - c) * (sh * f(0.10))
    c = c * (1.0 + ss(f(0.56), f(0.90), luma) * f(0.28))
    want = min(1.0, max(0.0, c)) ** GAMMA
    out = D3()
    L.or_camera_response(D3(grey, grey, grey), e, out)
    assert out[0] == pytest.approx(want, rel=1e-12) and out[1] == out[0] == out[2]


def test_camera_response_preserves_hue_below_saturation():
    """The tone curve is a per-pixel scalar on rgb (hue preserving, :2255-2259):
    out^2.2 is proportional to the input while no channel saturates."""
    out = D3()
    for rgb in ((0.3, 0.1, 0.05), (0.02, 0.05, 0.01), (0.6, 0.5, 0.2)):
        L.or_camera_response(D3(*rgb), 1.0, out)
        lin = np.array(out) ** (1.0 / GAMMA)
        assert max(lin) < 1.0
        assert lin / lin[0] == pytest.approx(np.array(rgb) / rgb[0], rel=1e-9)


# ---------------------------------------------------------------- metrics ----
def test_metrics_match_numpy_restatement():
    """accumulateMetricsKernel (:1319-1418) and the host reductions
    (:3495-3546): sums, maxima of positive values, L2 norms, flame height, the
    light/shadow means and the proxies, from numpy over the same fields."""
    grid = (9, 10, 8)
    nx, ny, nz = grid
    rng = np.random.default_rng(41)
    o, prm = scene(grid)
    F = {n: rng.uniform(0, 3, (nz, ny, nx)) for n in SCALARS}
    F["oxygen"] = rng.uniform(0, 1, (nz, ny, nx))
    for n, a in F.items():
        o.write(n, a)
    U, V, W = rng.normal(size=(nz, ny, nx + 1)), rng.normal(size=(nz, ny + 1, nx)), rng.normal(size=(nz + 1, ny, nx))
    P, Dv = rng.normal(size=(nz, ny, nx)), rng.normal(size=(nz, ny, nx))
    for n, a in (("u", U), ("v", V), ("w", W), ("pressure", P), ("divergence", Dv)):
        o.write(n, a)
    lnx, lny, lnz = firesim.light_dims(*grid)
    light = rng.uniform(0, 1, (lnz, lny, lnx, 4))
    shadow = rng.uniform(0, 1, (lnz, lny, lnx))
    o.write("light", light)
    o.write("shadow", shadow)
    m = o.metrics()
    cells = nx * ny * nz
    u, v, w = cell_coords(nx, ny, nz)
    names = dict(heat="heat_sum", fuel="fuel_sum", oxygen="oxygen_sum", soot="soot_sum", char="char_sum", ash="ash_sum",
                 pyrolysis="pyrolysis_sum", progress="progress_sum", turbulence="turbulence_energy_sum", soot_optics="soot_optical_depth_sum")
    for n, key in names.items():
        assert m[key] == pytest.approx(math.fsum(F[n].ravel()), rel=1e-13)
    for n, key in (("heat", "max_heat"), ("fuel", "max_fuel"), ("soot", "max_soot"), ("pyrolysis", "max_pyrolysis"),
                   ("progress", "max_progress"), ("turbulence", "max_turbulence_energy")):
        assert m[key] == F[n].max()
    div_after = np.diff(U, axis=2) + np.diff(V, axis=1) + np.diff(W, axis=0)
    assert m["divergence_before_l2"] == pytest.approx(math.sqrt(math.fsum((Dv ** 2).ravel()) / cells), rel=1e-13)
    assert m["divergence_after_l2"] == pytest.approx(math.sqrt(math.fsum((div_after ** 2).ravel()) / cells), rel=1e-13)
    assert m["divergence_before_max"] == np.abs(Dv).max()
    assert m["divergence_after_max"] == pytest.approx(np.abs(div_after).max(), rel=1e-15)
    assert m["divergence_reduction"] == pytest.approx(1 - m["divergence_after_l2"] / m["divergence_before_l2"], rel=1e-13)
    H, Fu, O, S, A, Py, Pr, So = (F[n] for n in ("heat", "fuel", "oxygen", "soot", "ash", "pyrolysis", "progress", "soot_optics"))
    marker = H + Fu * f(0.24) + Pr * f(0.55)
    assert m["flame_height_meters"] == pytest.approx((v * f(2.03))[marker > f(0.74)].max(), rel=1e-15)
    tK = f(293.0) + H * f(335.0)
    hrr = (Fu + Py * f(0.42)) * O * ss(f(610.0), f(1280.0), tK) * np.exp(-f(5200.0) / np.maximum(f(650.0), tK))
    od = So + S * (f(0.24) + f(0.05) * np.sqrt(H))
    flame = ss(f(0.72), f(2.10), H + Pr * f(0.62) + Py * f(0.24)) * ss(f(0.035), f(0.66), O) * ss(f(0.012), f(0.52), Fu + Py * f(0.36))
    smoke = ss(f(0.035), f(1.16), So + S * f(0.38) + A * f(0.12)) * ss(f(0.08), f(0.88), v)
    assert m["heat_release_proxy"] == pytest.approx(hrr.sum(), rel=1e-12)
    assert m["mean_optical_depth"] == pytest.approx(od.mean(), rel=1e-12)
    assert m["flame_mass_proxy"] == pytest.approx(flame.sum(), rel=1e-12)
    assert m["smoke_mass_proxy"] == pytest.approx(smoke.sum(), rel=1e-12)
    assert m["flame_smoke_overlap_proxy"] == pytest.approx((flame * smoke).sum(), rel=1e-12)
    # light trilinear at cell centres, shadow nearest (:618-664)
    Ls = np.zeros((nz, ny, nx, 4))
    for c in range(4):
        for idx in np.ndindex(nz, ny, nx):
            Ls[idx + (c,)] = oracle.sample_grid(light[..., c], (lnx - 1, lny - 1, lnz - 1), u[idx], v[idx], w[idx])
    near = lambda t, n: np.clip((t * (n - 1) + 0.5).astype(int), 0, n - 1)  # noqa: E731
    Sh = shadow[near(w, lnz), near(v, lny), near(u, lnx)]
    lumL = lum(Ls[..., 0], Ls[..., 1], Ls[..., 2])
    vol = Sh * np.exp(-Ls[..., 3] * f(0.80))
    assert m["mean_scene_light"] == pytest.approx(lumL.mean(), rel=1e-12)
    assert m["mean_scene_shadow"] == pytest.approx(Sh.mean(), rel=1e-12)
    assert m["mean_volumetric_shadow"] == pytest.approx(vol.mean(), rel=1e-12)
    assert m["mean_room_irradiance"] == pytest.approx((lumL * vol).mean(), rel=1e-12)
    assert m["invalid_cells"] == 0
    # a non-finite cell is counted and left out of every sum (:1363-1368)
    F["soot"][1, 2, 3] = np.nan
    o.write("soot", F["soot"])
    m2 = o.metrics()
    assert m2["invalid_cells"] == 1
    assert m2["heat_sum"] == pytest.approx(math.fsum(F["heat"].ravel()) - F["heat"][1, 2, 3], rel=1e-13)
    o.close()


# ------------------------------------------------------------------- beds ----
def test_capsule_mask_closed_form():
    """orientedCapsuleMask (:734-744): 1 on the capsule's axis segment, 0 at
    distance >= radius from it, the inverted smoothstep in between."""
    cx, cz, ang, hl, r = 0.1, -0.2, 0.6, 0.3, 0.08
    ax, az = math.cos(ang), math.sin(ang)
    for t in (-0.3, 0.0, 0.17, 0.3):  # along the axis, within the half length
        assert L.or_capsule_mask(cx + t * ax, cz + t * az, cx, cz, ang, hl, r) == 1.0
    nxp, nzp = -az, ax  # unit normal
    for d in (0.0, 0.02, 0.05, 0.07, 0.08, 0.2):
        got = L.or_capsule_mask(cx + 0.1 * ax + d * nxp, cz + 0.1 * az + d * nzp, cx, cz, ang, hl, r)
        assert got == pytest.approx(float(ss(r, r * f(0.22), d)), abs=1e-12)
    # beyond the segment end the distance is to the end point
    got = L.or_capsule_mask(cx + (hl + 0.05) * ax, cz + (hl + 0.05) * az, cx, cz, ang, hl, r)
    assert got == pytest.approx(float(ss(r, r * f(0.22), 0.05)), abs=1e-12)


def test_gas_burner_ring_and_ports_closed_form():
    """fuelBedMaterial scene 2 (:765-785): on the ring radius R between ports
    (cos(32 angle) = -1, port mask 0) the material is the ring's 0.035; on a
    port (cos = 1) it saturates; at 1.3 R and at the centre (pilot 0.045 R-scaled
    smoothstep) it takes its closed forms."""
    o, prm = scene((8, 8, 8), scene_id=2, emitter_radius=0.5, burner_center_count=1, burner_center_x=[0.1, 0, 0, 0],
                   burner_center_y=[0.2, 0, 0, 0], burner_center_z=[-0.05, 0, 0, 0])
    R = f(0.5)
    cx, cz = f(0.1), f(-0.05)
    between = math.pi / 32  # cos(32 a) = -1
    assert L.or_bed_material(o.h, cx + R * math.cos(between), cz + R * math.sin(between)) == pytest.approx(f(0.035), rel=1e-9)
    assert L.or_bed_material(o.h, cx + R, cz) == 1.0
    assert L.or_bed_material(o.h, cx + 1.3 * R, cz) == 0.0
    # centre: only the pilot term, smoothstep(0.22 R, 0, 0) 0.045
    assert L.or_bed_material(o.h, cx, cz) == pytest.approx(f(0.045), rel=1e-9)
    # inner ring at 0.52 R between ports: 0.055
    assert L.or_bed_material(o.h, cx + f(0.52) * R * math.cos(between), cz + f(0.52) * R * math.sin(between)) == pytest.approx(f(0.055), rel=1e-6)
    # source (:795-803): the height gate smoothstep(max(.005, band .34),
    # max(.0015, band .055), |h - h0|) times the material's breathing
    # 0.84 + 0.16 fbm(9x + 0.035 t, 11z - 0.025 t)
    band, h0 = f(prm.emitter_height_band_norm), f(prm.emitter_height_norm)
    t = 0.7
    breathing = f(0.84) + f(0.16) * L.or_fbm2((cx + R) * f(9.0) + t * f(0.035), cz * f(11.0) - t * f(0.025), L.or_salt(prm.noise_seed, 1))
    assert L.or_bed_source(o.h, cx + R, cz, h0, t) == pytest.approx(min(1.0, breathing), rel=1e-12)
    half = 0.5 * (max(f(0.005), band * f(0.34)) + max(f(0.0015), band * f(0.055)))
    want = float(ss(max(f(0.005), band * f(0.34)), max(f(0.0015), band * f(0.055)), half)) * min(1.0, breathing)
    assert L.or_bed_source(o.h, cx + R, cz, h0 + half, t) == pytest.approx(want, rel=1e-9)
    off = max(f(0.005), band * f(0.34)) + 0.01
    assert L.or_bed_source(o.h, cx + R, cz, h0 + off, t) == 0.0
    o.close()


def test_room_tray_and_campfire_support():
    """Scene 0's tray (:786-793) is zero beyond |x| = 1.08 or |z| = 0.56; the
    campfire (:747-760) is zero far from its logs and coal bed, and its log A
    centre carries at least the log's 0.72 weight (less the crack term)."""
    o0, _ = scene((8, 8, 8), scene_id=0)
    for x, z in ((1.09, 0.0), (0.0, 0.57), (-1.2, 0.3), (0.5, -0.6)):
        assert L.or_bed_material(o0.h, x, z) == 0.0
    assert L.or_bed_material(o0.h, 0.0, 0.0) > 0.0
    o0.close()
    o1, _ = scene((8, 8, 8), scene_id=1)
    for x, z in ((0.9, 0.0), (-0.8, 0.6), (0.0, 0.7)):
        assert L.or_bed_material(o1.h, x, z) == 0.0
    mA = L.or_bed_material(o1.h, f(-0.06), f(0.01))
    assert mA >= f(0.72) * (1 - f(0.24)) - 1e-12
    o1.close()


# ------------------------------------------------------------ room colour ----
def _rd(x, y, z):
    n = math.sqrt(x * x + y * y + z * z)
    return (x / n, y / n, z / n)


def test_room_background_fixed_rays_closed_form():
    """roomBackgroundRay (:1683-1865) with an empty light volume (shadow 1):
    no indirect light, so a ray that hits the plain ceiling returns
    (0.011, 0.011, 0.010) + grain, and a ray that hits the side wall the wall
    colour with its soot darkening; both scaled by the corner term, fog and
    vignette. The camera sits inside the room (distance 1.5)."""
    o, prm = scene((12, 12, 12), camera_distance=1.5, camera_yaw=0.0, camera_pitch=0.0, smoke_darkness=1.4)
    o.write("shadow", np.ones(o.shape(firesim.FIELD["shadow"])))
    o.set_time(0.5, 0.5)
    eye = D3()
    L.or_camera_eye(o.h, eye)
    eye = np.array(eye)
    assert eye == pytest.approx([0.0, 0.72, 1.5], abs=1e-6)
    salt = L.or_salt(prm.noise_seed, 5)
    rt = float(np.float32(0.5))
    uvx, uvy = 0.1, 0.1  # outside the source-model overlay box
    vign = ss(f(0.78), f(0.20), math.sqrt((uvx - f(0.50)) ** 2 + (uvy - f(0.46)) ** 2))

    def expect(base, hit, t, corner):
        gx = math.floor(hit[0] * f(19.0) + hit[2] * f(13.0))
        gy = math.floor(hit[1] * f(15.0) + rt * f(0.08))
        grain = (L.or_hash2(int(gx), int(gy), salt) - f(0.5)) * f(0.006)
        fog = ss(f(3.8), f(0.2), t)
        return [(c * (f(0.84) + ss(0.0, f(0.54), corner) * f(0.16)) + grain) * (f(0.12) + fog * f(0.56)) * (f(0.58) + vign * f(0.30))
                for c in base]

    out = D3()
    # ceiling: hit (0.73, 2.18, 1.5) away from the panel lines and fixtures
    rd = _rd(0.5, 1.0, 0.0)
    t = (f(2.18) - eye[1]) / rd[1]
    hit = eye + t * np.array(rd)
    L.or_room_background(o.h, uvx, uvy, D3(*rd), 0.0, out)
    want = expect((f(0.011), f(0.011), f(0.010)), hit, t, min(f(2.55) - abs(hit[0]), f(2.55) - abs(hit[2])))
    assert list(out) == pytest.approx(want, rel=1e-9)
    # side wall x = +2.55 at height 0.72: base 0.014, soot darkening toward (0.010, 0.011, 0.012)
    rd = (1.0, 0.0, 0.0)
    t = (f(2.55) - eye[0]) / rd[0]
    hit = eye + t * np.array(rd)
    soot = math.exp(-(hit[0] ** 2 * f(0.82) + hit[2] ** 2 * f(1.12))) * float(ss(f(0.36), f(1.92), hit[1]))
    k = soot * f(0.82) * f(1.4)
    base = [f(0.014) + (tgt - f(0.014)) * k for tgt in (f(0.010), f(0.011), f(0.012))]
    corner = min(f(2.55) - abs(hit[2]), min(hit[1], f(2.18) - hit[1]))
    L.or_room_background(o.h, uvx, uvy, D3(*rd), 0.0, out)
    want = expect(base, hit, t, corner)
    assert list(out) == pytest.approx(want, rel=1e-9)
    o.close()


# ------------------------------------------------------------- raymarch ----
def test_render_transmittance_is_beer_lambert():
    """renderKernel (:1953-2244) with soot and no flame (no oxygen): debug mode
    2 accumulates the soot optical depth sum(ext dt) and mode 3 the
    transmittance prod(exp(-ext dt)) along the same jittered samples, so where
    the depth stays below 1 (no saturation) and the march does not stop early
    (trans > 0.018), transmittance = exp(-optical depth) (Beer-Lambert)."""
    grid = (16, 18, 14)
    rng = np.random.default_rng(51)
    o, prm = scene(grid, frame=(24, 20), raymarch_steps=48, render_debug_mode=2)
    F = _smooth_fields(grid, rng)
    F["oxygen"] = np.zeros_like(F["oxygen"])
    for n, a in F.items():
        o.write(n, a * (0.25 if n in ("soot", "soot_optics") else 1.0))
    o.run_stage(firesim.STAGE_LIGHT)
    od = o.render_hdr()[..., 0]
    p3 = o.get_params()
    p3.render_debug_mode = 3
    o.set_params(p3)
    trans = o.render_hdr()[..., 0]
    ok = (od > 1e-4) & (od < 0.99) & (trans > 0.05)
    assert ok.sum() > 20, "the volume must attenuate a fair share of pixels"
    assert np.allclose(trans[ok], np.exp(-od[ok]), rtol=1e-12)
    assert np.all(trans[od == 0] == 1.0)
    o.close()


# ------------------------------------------ beds, scenes 0-2 (restatement) ----
def _capsule(x, z, cx, cz, ang, hl, r):
    """orientedCapsuleMask (:734-744)."""
    ca, sa = math.cos(ang), math.sin(ang)
    px, pz = x - cx, z - cz
    qx, qz = px * ca + pz * sa, -px * sa + pz * ca
    over = max(abs(qx) - hl, 0.0)
    return float(ss(r, r * f(0.22), math.sqrt(over * over + qz * qz)))


def _bed_material(x, z, prm):
    """fuelBedMaterial (:746-793); scene 3 has no bed."""
    sid = prm.scene_id
    salt = L.or_salt(prm.noise_seed, 1)

    def fbm(a, b):
        return L.or_fbm2(a, b, salt)

    if sid == 3:
        return 0.0
    if sid == 1:
        A = _capsule(x, z, f(-0.06), f(0.01), f(0.38), f(0.34), f(0.060))
        B = _capsule(x, z, f(0.07), f(-0.02), f(-0.58), f(0.33), f(0.058))
        C = _capsule(x, z, f(0.01), f(0.08), f(1.34), f(0.24), f(0.050))
        contact = min(1.0, max(0.0, min(A, B) * f(1.25) + min(B, C) * f(1.10) + min(C, A) * f(0.92)))
        coal = float(ss(f(0.38), f(0.08), math.sqrt(x * x * f(1.35) + z * z * f(2.40)))) * float(
            ss(f(0.34), f(0.88), fbm(x * f(9.4) + z * f(1.7), z * f(11.2) - x * f(3.1))))
        cracks = float(ss(f(0.65), f(0.93), fbm(x * f(21.0) - z * f(6.0), z * f(24.0) + x * f(4.0))))
        return min(1.0, max(0.0, (A * f(0.72) + B * f(0.70) + C * f(0.54) + coal * f(0.84) + contact * f(1.18)) * (1.0 - cracks * f(0.24))))
    if sid == 2:
        R = max(f(0.04), f(prm.emitter_radius))
        cx = f(prm.burner_center_x[0]) if prm.burner_center_count > 0 else f(prm.emitter_center_x)
        cz = f(prm.burner_center_z[0]) if prm.burner_center_count > 0 else f(prm.emitter_center_z)
        px, pz = x - cx, z - cz
        radius = math.sqrt(px * px + pz * pz)
        angle = math.atan2(pz, px)
        ring = float(ss(R * f(0.072), R * f(0.014), abs(radius - R)))
        inner = float(ss(R * f(0.052), R * f(0.012), abs(radius - R * f(0.52))))
        port = float(ss(f(0.80), f(0.992), f(0.5) + f(0.5) * math.cos(angle * f(32.0))))
        pilot = float(ss(R * f(0.22), f(0.000), radius)) * f(0.045)
        src = max(0.0, (ring * (f(0.035) + port * f(1.72)) + inner * f(0.055) + pilot) * float(ss(R * f(1.28), R * f(1.05), radius)))
        return min(1.0, src)
    tray = float(ss(f(1.08), f(0.030), abs(x)) * ss(f(0.56), f(0.026), abs(z)))
    ember = float(ss(f(1.00), f(0.018), abs(x)) * ss(f(0.50), f(0.020), abs(z)))
    mound = math.exp(-(x * x * f(0.70) + z * z * f(2.15)))
    log_mass = fbm(x * f(3.7) + z * f(0.9), z * f(5.4) - x * f(0.6))
    chips = fbm(x * f(15.0) + z * f(4.0), z * f(18.0) - x * f(5.0))
    strand = f(0.5) + f(0.5) * math.sin((x * f(19.0) + z * f(7.0)) + fbm(z * f(4.0), x * f(3.0)) * f(4.2))
    cracks = float(ss(f(0.44), f(0.74), chips) * ss(f(0.20), f(0.84), log_mass))
    voids = float(ss(f(0.70), f(0.93), chips * f(0.72) + strand * f(0.34)))
    clumps = float(ss(f(0.22), f(0.86), log_mass * f(0.82) + chips * f(0.36) + strand * f(0.18)))
    porous = (f(0.24) + clumps * f(0.94) + ember * f(0.20)) * (f(0.64) + mound * f(0.46))
    return min(1.0, max(0.0, tray * porous * (1.0 - cracks * f(0.46)) * (1.0 - voids * f(0.34))))


def _bed_source(x, z, h, prm, time):
    """fuelBedSource (:795-803)."""
    sid = prm.scene_id
    mat = _bed_material(x, z, prm)
    band, h0 = f(prm.emitter_height_band_norm), f(prm.emitter_height_norm)
    height = float(ss(max(f(0.005), band * f(0.34)), max(f(0.0015), band * f(0.055)), abs(h - h0)) if sid == 2
                   else ss(f(0.245) if sid == 1 else f(0.145), f(0.00), h))
    breathing = f(0.84) + f(0.16) * L.or_fbm2(x * f(9.0) + time * f(0.035), z * f(11.0) - time * f(0.025), L.or_salt(prm.noise_seed, 1))
    return min(1.0, max(0.0, height * min(1.0, max(0.0, mat * breathing))))


BED_PARAMS = {
    0: dict(scene_id=0),
    1: dict(scene_id=1),
    # burner height 0.146875 puts the source band on the cell centres v = 0.0625 of an 8-cell column
    2: dict(scene_id=2, emitter_radius=0.42, burner_center_count=1, burner_center_x=[0.05, 0, 0, 0], burner_center_y=[0.146875, 0, 0, 0],
            burner_center_z=[-0.03, 0, 0, 0], emitter_height_band_norm=0.09),
}


@pytest.mark.parametrize("sid", [0, 1, 2])
def test_bed_material_and_source_match_restatement(sid):
    o, prm = scene((8, 8, 8), noise_seed=13, **BED_PARAMS[sid])
    rng = np.random.default_rng(60 + sid)
    pts = rng.uniform([-1.15, -0.9], [1.15, 0.9], (400, 2))
    if sid == 2:  # concentrate on the burner's rings and ports
        r = rng.uniform(0, 0.6, 300)
        a = rng.uniform(0, 2 * np.pi, 300)
        pts = np.concatenate([pts, np.stack([0.05 + r * np.cos(a), -0.03 + r * np.sin(a)], 1)])
    if sid == 1:  # the logs and coal bed lie within |x| < 0.45, |z| < 0.3
        pts = np.concatenate([pts, rng.uniform([-0.45, -0.3], [0.45, 0.3], (300, 2))])
    nonzero = 0
    for x, z in pts:
        want = _bed_material(x, z, prm)
        assert L.or_bed_material(o.h, x, z) == pytest.approx(want, rel=1e-12, abs=1e-15), (x, z)
        nonzero += want > 0
        h = rng.uniform(0, 0.4)
        assert L.or_bed_source(o.h, x, z, h, 0.9) == pytest.approx(_bed_source(x, z, h, prm, 0.9), rel=1e-12, abs=1e-15)
    assert nonzero > 50
    o.close()


# ------------------------------------ advect-react on the beds (scenes 0-2) ----
def _react_bed_reference(s, u, v, w, prm, time):
    """advectReactKernel (:934-1126) on uniform inputs with zero velocity for
    scenes 0-2: the procedural bed's material and source, the char/ash and
    pyrolysis feed and the per-scene gains (:1020-1045, :1067), no mouse or
    emitter sources; the later chain is that of _react_reference."""
    sid = prm.scene_id
    dt = f(prm.dt)
    shape = u.shape
    out = {n: np.empty(shape) for n in SCALARS}
    salt = L.or_salt(prm.noise_seed, 1)
    for idx in np.ndindex(shape):
        uu, vv, ww = u[idx], v[idx], w[idx]
        heat, fuel, oxygen, soot = s["heat"], s["fuel"], s["oxygen"], s["soot"]
        char, ash, pyro = s["char"], s["ash"], s["pyrolysis"] * f(0.55)
        progress, turb, sopt = s["progress"], s["turbulence"], s["soot_optics"]
        activity = heat + fuel * f(0.18) + soot * f(0.24) + s["pyrolysis"] * f(0.20) + progress * f(0.14) + sopt * f(0.08) + turb * f(0.06)
        if vv > f(0.18) and activity < f(0.00035) and char < f(0.00010) and ash < f(0.00010):
            vals = dict(heat=0.0, fuel=0.0, oxygen=min(1.0, oxygen + (1.0 - oxygen) * dt * f(0.08)), soot=0.0, char=char, ash=ash,
                        pyrolysis=0.0, progress=0.0, turbulence=turb * max(0.0, 1.0 - dt * f(0.54)), soot_optics=0.0)
            for n in SCALARS:
                out[n][idx] = vals[n]
            continue
        wx, wz = (uu - f(0.50)) * f(2.10), (ww - f(0.50)) * f(1.64)
        mat = _bed_material(wx, wz, prm)
        snoise = f(0.70) + f(0.48) * L.or_fbm2(wx * f(13.0) + wz * f(3.0) + time * f(0.38), vv * f(31.0) + wz * f(11.0), salt)
        exposed = min(1.0, max(0.0, char / max(f(0.055), char + ash * f(1.75))))
        edges = float(ss(f(0.06), f(0.78), mat) * ss(f(0.92), f(0.02), ash))
        feed = f(1.18) if sid == 2 else exposed * edges
        g_py, g_fu, g_he, g_so = {2: (f(0.42), f(0.46), f(1.18), f(0.18)), 1: (f(1.34), f(1.18), f(1.08), f(1.22))}.get(sid, (1.0, 1.0, 1.0, 1.0))
        bed = _bed_source(wx, wz, vv, prm, time) * snoise * feed * f(prm.intensity)
        tK = f(293.0) + heat * f(335.0)
        radfb = float(ss(f(430.0), f(980.0), tK))
        solid = min(1.0, max(0.0, oxygen * f(1.28) - soot * f(0.030)))
        flux = min(1.0, max(0.0, (tK - f(405.0)) / f(900.0))) + bed * f(0.34) + sopt * f(0.024)
        crel = min(char, char * flux * solid * (f(0.14) + radfb * f(1.38)) * dt)
        pyro += (bed * (f(0.34) + radfb * f(1.05)) * dt + crel * (f(2.45) + snoise * f(0.72))) * g_py
        char += mat * dt * (f(0.00005) if sid == 2 else (f(0.050) if sid == 1 else f(0.026)))
        char -= crel * f(0.92)
        ash += crel * (f(0.025) if sid == 2 else f(0.16) + f(0.22) * min(1.0, max(0.0, 1.0 - oxygen)))
        heat += (bed * (f(10.40) if sid == 2 else f(6.20)) * dt * (1.0 + radfb * f(1.12)) + crel * f(3.35)) * g_he
        fuel += pyro * (f(1.34) if sid == 2 else f(3.70)) * g_fu
        soot += pyro * (f(0.090) + ash * f(0.032) + min(1.0, max(0.0, 1.0 - oxygen)) * f(0.13)) * f(prm.smoke) * (
            f(0.010) if sid == 2 else (f(0.92) if sid == 1 else 1.0)) * g_so
        tK = f(293.0) + heat * f(335.0)
        burn_lim = min(fuel, oxygen * f(1.42))
        burn = min(burn_lim, burn_lim * float(ss(f(575.0), f(1220.0), tK)) * math.exp(-f(3800.0) / max(f(640.0), tK)) *
                   (f(118.0) + f(prm.intensity) * f(34.0)) * dt)
        inc = min(1.0, max(0.0, 1.0 - oxygen * f(1.30) + fuel * f(0.065) + soot * f(0.015)))
        yld = burn * (f(0.035) + inc * f(0.34)) * f(prm.smoke) * (f(0.006) if sid == 2 else (f(0.96) if sid == 1 else 1.0))
        oxid = soot * oxygen * float(ss(f(760.0), f(1540.0), tK)) * dt * f(0.82)
        front = burn * (f(9.8) + turb * f(3.4)) + pyro * f(0.70) + crel * f(2.10)
        bA = float(sat(ss(f(0.70), f(0.96), vv) + ss(f(0.11), f(0.03), uu) + ss(f(0.89), f(0.97), uu) + ss(f(0.11), f(0.03), ww) +
                       ss(f(0.89), f(0.97), ww)))
        up = float(ss(f(0.52), f(0.98), vv))
        heat += burn * f(7.0)
        fuel -= burn * f(1.20)
        oxygen -= burn * f(0.98)
        soot = soot + yld - oxid
        progress = (progress + front) * math.exp(-dt * (f(0.72) + vv * f(0.45) + oxygen * f(0.16) + bA * f(1.05) + up * f(0.26)))
        oxygen = oxygen + (1.0 - oxygen) * (bA * dt * f(1.7))
        oxygen += (1.0 - oxygen) * dt * f(0.08)
        tK = f(293.0) + heat * f(335.0)
        rb = max(0.0, (tK - f(293.0)) / f(1560.0))
        heat -= (rb * rb) ** 2 * dt * (f(0.88) + soot * f(0.085) + sopt * f(0.018))
        heat *= math.exp(-dt * (f(0.31) + vv * f(0.64) + soot * f(0.022) + bA * f(0.26) + up * f(0.18)))
        fuel *= math.exp(-dt * (f(0.58) + heat * f(0.20) + vv * f(0.92) + bA * f(0.18)))
        soot *= math.exp(-dt * (f(0.060) + vv * f(0.092) + oxygen * f(0.028) + bA * f(0.96) + up * f(0.22)))
        pyro *= math.exp(-dt * (f(1.55) + vv * f(0.58) + bA * f(0.18)))
        act = min(1.0, max(0.0, progress * f(0.64) + burn * f(5.2) + heat * f(0.065)))
        turb += dt * (act * f(prm.turbulence) * f(0.62) + front * f(0.085) + soot * f(0.018))
        turb *= math.exp(-dt * (f(0.54) + vv * f(0.24) + bA * f(0.42) + up * f(0.16)))
        growth = min(1.0, max(0.0, inc * f(0.36) + ash * f(0.13) + soot * f(0.055) + float(ss(f(0.65), f(1.55), tK / f(1000.0))) * f(0.18)))
        so = soot * (f(1.04) + f(0.28) * math.sqrt(max(0.0, heat)) + f(0.24) * ash + f(0.46) * growth)
        so = so + (so * f(0.48) - so) * min(1.0, max(0.0, oxid * f(0.92)))
        vals = dict(heat=(heat, 8.8), fuel=(fuel, 5.5), oxygen=(oxygen, 1.0), soot=(soot, 7.4), char=(char, 2.8), ash=(ash, 1.9),
                    pyrolysis=(pyro, 3.6), progress=(progress, 3.4), turbulence=(turb, 4.2), soot_optics=(so, 10.0))
        for n, (val, cap) in vals.items():
            out[n][idx] = min(f(cap), max(0.0, val))
    return out


@pytest.mark.parametrize("sid", [0, 1, 2])
def test_react_on_procedural_beds_matches_restatement(sid):
    grid = (24, 8, 20)
    # oxygen 0.5 keeps the solid-oxygen term 1.28 O - 0.03 soot below its saturation
    st = dict(heat=0.9, fuel=0.4, oxygen=0.5, soot=0.1, char=0.6, ash=0.15, pyrolysis=0.3, progress=0.2, turbulence=0.1, soot_optics=0.2)
    o, prm = scene(grid, noise_seed=17, intensity=1.2, smoke=0.9, **BED_PARAMS[sid])
    for n in SCALARS:
        o.write(n, np.full(o.shape(firesim.FIELD[n]), st[n]))
    o.set_time(0.6, 0.6)
    o.run_stage(firesim.STAGE_ADVECT_REACT)
    u, v, w = cell_coords(*grid)
    lit = sum(_bed_source((uu - f(0.5)) * f(2.10), (ww - f(0.5)) * f(1.64), vv, prm, f(0.6)) > 0 for uu, vv, ww in zip(u.ravel(), v.ravel(), w.ravel()))
    assert lit >= 8, "the bed must feed some cells"
    want = _react_bed_reference(st, u, v, w, prm, f(0.6))
    for n in SCALARS:
        got = o.read(n)
        assert np.allclose(got, want[n], rtol=1e-12, atol=1e-15), (sid, n, np.abs(got - want[n]).max())
    o.close()


@pytest.mark.parametrize("case,grid,state", [
    ("activity_above_threshold", (6, 10, 6), dict(turbulence=0.0060)),   # activity 0.00036 >= 0.00035: live path
    ("activity_below_threshold", (6, 10, 6), dict(turbulence=0.0057)),   # 0.000342: early-out above v = 0.18
    ("char_above_threshold", (6, 10, 6), dict(char=0.000105)),
    ("ash_above_threshold", (6, 10, 6), dict(ash=0.000105)),
    ("height_threshold", (4, 50, 4), dict(turbulence=0.001)),            # a cell row at v = 0.19
])
def test_react_early_out_thresholds(case, grid, state):
    got, rel, want, wrel = react(state, grid=grid)
    for n in SCALARS:
        assert np.allclose(got[n], want[n], rtol=1e-12, atol=1e-15), (case, n)


def test_react_mouse_raises_oxygen():
    """The heat/fuel mouse source lifts oxygen to at least 0.72 + 0.20 source (:1056)."""
    got, rel, want, wrel = react(dict(heat=0.3, fuel=0.2, oxygen=0.3), left_down=1, mouse_x=0.5, mouse_y=0.35)
    for n in SCALARS:
        assert np.allclose(got[n], want[n], rtol=1e-12, atol=1e-15), n
    assert got["oxygen"].max() > 0.72


@pytest.mark.parametrize("sid", [0, 1, 2])
def test_force_scene_scales_match_restatement(sid):
    """The per-scene curl and buoyancy scales and the burner's low buoyancy
    fade (:1168-1172)."""
    grid = (7, 9, 6)
    rng = np.random.default_rng(70 + sid)
    o, prm = scene(grid, turbulence=1.1, wind_x=-0.2, noise_seed=3, **BED_PARAMS[sid])
    nx, ny, nz = grid
    H = rng.uniform(0, 5, (nz, ny, nx))
    S, K, C = rng.uniform(0, 2, (nz, ny, nx)), rng.uniform(0, 1.5, (nz, ny, nx)), rng.uniform(0, 1.2, (nz, ny, nx))
    U, V, W = rng.uniform(-1, 1, (nz, ny, nx + 1)), rng.uniform(-0.5, 2, (nz, ny + 1, nx)), rng.uniform(-1, 1, (nz + 1, ny, nx))
    for n, a in (("heat", H), ("soot", S), ("turbulence", K), ("progress", C), ("u", U), ("v", V), ("w", W)):
        o.write(n, a)
    o.set_time(0.4, 0.4)
    o.run_stage(firesim.STAGE_FORCE)
    curl = {2: f(0.38), 1: f(1.18)}.get(sid, 1.0)
    buoy = {2: f(0.42), 1: f(1.06)}.get(sid, 1.0)
    Uo, Vo, Wo, d = _force_reference_scaled(H, S, K, C, U, V, W, prm, f(0.4), curl, buoy, sid == 2)
    assert np.allclose(o.read("u"), Uo, rtol=1e-12, atol=1e-14)
    assert np.allclose(o.read("v"), Vo, rtol=1e-12, atol=1e-14)
    assert np.allclose(o.read("w"), Wo, rtol=1e-12, atol=1e-14)
    assert np.allclose(o.read("divergence"), d, rtol=1e-12, atol=1e-13)
    o.close()


def _force_reference_scaled(H, S, K, C, U, V, W, prm, time, curl_scale, buoy_scale, burner):
    """_force_reference with the scene's curl and buoyancy scales and the
    burner's buoyancy fade smoothstep(0.24, 0.035, v)."""
    nz, ny, nx = H.shape
    u, v, w = cell_coords(nx, ny, nz)
    dt, T = f(prm.dt), f(prm.turbulence)
    zi, yi, xi = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")

    def at(a, dx, dy, dz):
        return a[np.clip(zi + dz, 0, nz - 1), np.clip(yi + dy, 0, ny - 1), np.clip(xi + dx, 0, nx - 1)]

    gx, gy, gz = at(H, 1, 0, 0) - at(H, -1, 0, 0), at(H, 0, 1, 0) - at(H, 0, -1, 0), at(H, 0, 0, 1) - at(H, 0, 0, -1)
    uL, uR, vD, vU, wB, wF = U[:, :, :-1], U[:, :, 1:], V[:, :-1, :], V[:, 1:, :], W[:-1], W[1:]
    cx, cy, cz = (wF - wB) - (vU - vD), (uR - uL) - (wF - wB), (vU - vD) - (uR - uL)
    salt = L.or_salt(prm.noise_seed, 1)
    na = np.vectorize(lambda a, b: L.or_fbm2(a, b, salt))(u * f(9.0) + w * f(3.7) + time * f(0.22), v * f(13.0) - time * f(0.31)) - f(0.5)
    nb = np.vectorize(lambda a, b: L.or_fbm2(a, b, salt))(w * f(10.0) - time * f(0.19), u * f(8.0) + v * f(5.5) + time * f(0.27)) - f(0.5)
    thermal = ss(f(0.04), f(4.20), H + S * f(0.28))
    drag = ss(f(0.26), f(0.94), v)
    les = np.sqrt(np.maximum(0.0, K)) * (f(0.30) + C * f(0.54))
    gain = T * curl_scale * (f(0.18) + thermal * f(0.82) + les * f(1.12))
    vortex = T * curl_scale * thermal * sat((np.abs(cx) + np.abs(cy) + np.abs(cz)) * f(0.46)) * f(0.68)
    fade = ss(f(0.24), f(0.035), v) if burner else ss(f(0.96), f(0.08), v)
    damp = 0.5 * (vD + vU) * (f(0.10) + drag * f(0.72) + S * f(0.020))
    b = H * buoy_scale * (f(1.62) + T * f(0.34) + C * f(0.16)) * fade - S * f(0.095) - gy * f(0.18) - damp
    prof = f(0.34) + v * f(0.96)
    fv = dt * (b + cy * vortex * f(0.22))
    fu = dt * (f(prm.wind_x) * prof + (-gz * f(1.36) + na * (f(0.20) + drag * f(0.16)) + cx * vortex) * gain)
    fw = dt * (f(prm.wind_z) * prof + (gx * f(1.36) + nb * (f(0.20) + drag * f(0.16)) + cz * vortex) * gain)
    Uo, Vo, Wo = U.copy(), V.copy(), W.copy()
    Uo[:, :, :-1] += fu
    Uo[:, :, 1:] += fu
    Vo[:, 1:, :] += fv
    Wo[:-1] += fw
    Wo[1:] += fw
    Uo = np.clip(Uo, f(-4.0), f(4.0))
    Uo[:, :, [0, -1]] = 0.0
    Vo = np.clip(Vo, f(-0.9), f(6.4))
    Vo[:, 0, :] = 0.0
    Wo = np.clip(Wo, f(-4.0), f(4.0))
    Wo[[0, -1]] = 0.0
    return Uo, Vo, Wo, -(np.diff(Uo, axis=2) + np.diff(Vo, axis=1) + np.diff(Wo, axis=0))


def test_camera_response_shoulder_on_coloured_input():
    """Above luma 0.94 the shoulder scales every channel by
    lerp(1, 1/(luma/0.94), 0.10 smoothstep(0.78, 1.08, luma)): a saturated
    green keeps red and blue below 1 where the shoulder is observable."""
    for rgb, e in (((0.5, 12.0, 0.2), 1.0), ((0.05, 9.0, 0.8), 1.4), ((0.3, 0.25, 0.02), 0.9), ((0.2, 2.3, 0.05), 1.0)):
        x = np.array(rgb) * f(e)
        l0 = max(f(0.000001), lum(*x))
        c = x * (_aces(l0) / l0)
        luma = lum(*c)
        c = c * f(0.82) + (c - c * f(0.82)) * ss(0.0, f(0.055), luma)
        c = c + (c / max(1.0, luma / f(0.94)) - c) * (ss(f(0.78), f(1.08), luma) * f(0.10))
        c = c * (1.0 + ss(f(0.56), f(0.90), luma) * f(0.28))
        want = np.clip(c, 0, 1) ** GAMMA
        if luma > f(0.94):
            assert min(c) < 1.0, "a channel must stay below saturation to observe the shoulder"
        out = D3()
        L.or_camera_response(D3(*rgb), f(e), out)
        assert list(out) == pytest.approx(list(want), rel=1e-12)


def test_metrics_flame_height_marker_weights():
    """The flame marker heat + 0.24 fuel + 0.55 progress > 0.74 (:1386). Rows
    0-5 burn (heat 1); one higher row holds only fuel (progress) just below or
    just above the threshold with the reference's weight, so the flame height
    is that row's only when its marker crosses 0.74: 2.95 * 0.24 = 0.708 does
    not, 3.2 * 0.24 = 0.768 does; 1.3 * 0.55 = 0.715 does not, 1.4 * 0.55 =
    0.77 does."""
    grid = (4, 20, 4)
    for field, value, row, top in (("fuel", 2.95, 17, 5), ("fuel", 3.2, 17, 17), ("progress", 1.30, 15, 5), ("progress", 1.4, 15, 15)):
        o, prm = scene(grid)
        for n in SCALARS:
            o.write(n, np.zeros(o.shape(firesim.FIELD[n])))
        heat = np.zeros(o.shape(0))
        heat[:, :6, :] = 1.0
        o.write("heat", heat)
        arr = np.zeros(o.shape(0))
        arr[:, row, :] = value
        o.write(field, arr)
        m = o.metrics()
        assert m["flame_height_meters"] == pytest.approx((top + 0.5) / 20 * f(2.03), rel=1e-12), (field, value)
        o.close()


# --------------------------------------------------- room colour (restatement) ----
def _camera(prm, fw, fh):
    """computeCameraFromSettings (:315-329)."""
    yaw, pitch, dist = f(prm.camera_yaw), f(prm.camera_pitch), f(prm.camera_distance)
    cp = math.cos(pitch)
    eye = np.array([math.sin(yaw) * dist * cp, f(0.72) + math.sin(pitch) * dist, math.cos(yaw) * dist * cp])

    def norm(a):
        return a / math.sqrt(max(f(0.000001), float(a @ a)))

    fwd = norm(np.array([0.0, f(0.72), 0.0]) - eye)
    right = norm(np.cross(fwd, [0.0, 1.0, 0.0]))
    up = norm(np.cross(right, fwd))
    return dict(eye=eye, fwd=fwd, right=right, up=up, tan=f(0.50), aspect=f(fw / fh))


def _light_at(light, uvw):
    lnz, lny, lnx, _ = light.shape
    return np.array([oracle.sample_grid(np.ascontiguousarray(light[..., c]), (lnx - 1, lny - 1, lnz - 1), *uvw) for c in range(4)])


def _shadow_at(shadow, uvw):
    lnz, lny, lnx = shadow.shape
    u, v, w = (min(1.0, max(0.0, t)) for t in uvw)
    x = min(max(int(u * (lnx - 1) + f(0.5)), 0), lnx - 1)
    y = min(max(int(v * (lny - 1) + f(0.5)), 0), lny - 1)
    z = min(max(int(w * (lnz - 1) + f(0.5)), 0), lnz - 1)
    return shadow[z, y, x]


def _rect(px, py, cx, cy, hx, hy, fe):
    return float(ss(fe, -fe, max(abs(px - cx) - hx, abs(py - cy) - hy)))


def _line(px, py, ax, ay, bx, by, th):
    pa, ba = np.array([px - ax, py - ay]), np.array([bx - ax, by - ay])
    h = min(1.0, max(0.0, float(pa @ ba) / max(f(0.000001), float(ba @ ba))))
    d = pa - ba * h
    return float(ss(th, th * f(0.25), math.sqrt(float(d @ d))))


def _dir_of(prm):
    """The directional light's travel direction, normalized in FP32 (a zero
    vector means straight down)."""
    d = np.array([prm.dir_light_x, prm.dir_light_y, prm.dir_light_z], dtype=np.float32)
    n = np.float32(np.sqrt(np.float32(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])))
    if not (n > np.float32(1e-6) and np.isfinite(n)):
        return np.array([0.0, -1.0, 0.0])
    return (d / n).astype(np.float64)


def _dir_colour(prm):
    return np.array([f(prm.dir_light_r), f(prm.dir_light_g), f(prm.dir_light_b)])


def _room_reference(prm, light, shadow, cam, uv, rd, glow, time, dirT=None):
    """roomBackgroundRay (:1683-1865) with gatherSceneIrradiance (:1589-1626),
    volumeShadowRay (:1628-1645) and sourceModelOverlay (:1661-1681);
    floorGrid (:1567-1573) is identically 0."""
    sid, cin = prm.scene_id, prm.cinematic_mode != 0
    half, ceil_y = f(2.55), f(2.18)
    eye, rd = cam["eye"], np.asarray(rd, dtype=np.float64)
    best, hit, nrm, surf = f(1.0e6), np.zeros(3), np.array([0.0, 1.0, 0.0]), 0
    if rd[1] < -f(0.0001):
        t = -eye[1] / rd[1]
        q = eye + rd * t
        if t > 0 and abs(q[0]) <= half and abs(q[2]) <= half and t < best:
            best, hit, nrm, surf = t, q, np.array([0.0, 1.0, 0.0]), 1
    if rd[1] > f(0.0001):
        t = (ceil_y - eye[1]) / rd[1]
        q = eye + rd * t
        if t > 0 and abs(q[0]) <= half and abs(q[2]) <= half and t < best:
            best, hit, nrm, surf = t, q, np.array([0.0, -1.0, 0.0]), 2
    if abs(rd[0]) > f(0.0001):
        px = half if rd[0] > 0 else -half
        t = (px - eye[0]) / rd[0]
        q = eye + rd * t
        if t > 0 and 0 <= q[1] <= ceil_y and abs(q[2]) <= half and t < best:
            best, hit, nrm, surf = t, q, np.array([-1.0 if px > 0 else 1.0, 0.0, 0.0]), 3
    if abs(rd[2]) > f(0.0001):
        pz = half if rd[2] > 0 else -half
        t = (pz - eye[2]) / rd[2]
        q = eye + rd * t
        if t > 0 and 0 <= q[1] <= ceil_y and abs(q[0]) <= half and t < best:
            best, hit, nrm, surf = t, q, np.array([0.0, 0.0, -1.0 if pz > 0 else 1.0]), 4
    F3 = lambda *a: np.array([f(x) for x in a])  # noqa: E731
    col, alb, gis = F3(0.014, 0.014, 0.014), F3(0.26, 0.26, 0.25), f(0.24)
    hx, hy, hz = hit
    rg = f(prm.reflection_gain)
    if surf == 1:
        slab = math.sin(hx * f(2.20) + hz * f(1.40)) * f(0.50) + math.sin(hz * f(3.10) - hx * f(0.80)) * f(0.50)
        col = F3(0.016, 0.016, 0.015) if cin else F3(0.014, 0.014, 0.013)
        col = col + F3(0.014, 0.014, 0.013) * (slab * f(0.018))
        alb, gis = F3(0.26, 0.26, 0.25), f(0.22)
        if sid == 1:
            mat = _bed_material(hx, hz, prm)
            coal = math.exp(-(hx * hx * f(1.25) + hz * hz * f(2.45))) * float(ss(f(0.28), f(0.90), L.or_fbm2(hx * f(13.0), hz * f(15.0), L.or_salt(prm.noise_seed, 1))))
            scorch = math.exp(-(hx * hx * f(0.66) + hz * hz * f(1.15)))
            col = lerp(col, F3(0.018, 0.015, 0.012), min(1.0, max(0.0, mat * f(0.78) + scorch * f(0.20))))
            col = col + F3(0.090, 0.020, 0.004) * (coal * glow * f(0.026))
            col = col + F3(0.17, 0.11, 0.055) * (mat * f(0.050))
            alb = lerp(alb, F3(0.040, 0.030, 0.022), mat * f(0.86))
            gis = lerp(gis, f(0.44), mat * f(0.80))
        elif sid == 2:
            radius = math.sqrt(hx * hx + hz * hz)
            plate = float(ss(f(0.78), f(0.70), radius))
            burner = _bed_material(hx, hz, prm)
            col = lerp(col, F3(0.018, 0.019, 0.020), plate * f(0.70))
            col = lerp(col, F3(0.005, 0.005, 0.005), burner * f(0.72))
            col = col + F3(0.090, 0.032, 0.006) * (burner * glow * f(0.012))
            alb = lerp(alb, F3(0.032, 0.033, 0.034), plate * f(0.82))
            gis = lerp(gis, f(0.42), plate * f(0.72))
        else:
            body = _rect(hx, hz, 0.0, 0.0, f(1.08), f(0.58), f(0.022))
            inner = _rect(hx, hz, 0.0, 0.0, f(0.96), f(0.47), f(0.020))
            rim = min(1.0, max(0.0, body - inner * f(0.74)))
            ember = inner * math.exp(-(hx * hx * f(0.74) + hz * hz * f(2.20)))
            contact = inner * math.exp(-(hx * hx * f(1.35) + hz * hz * f(3.35)))
            scorch = math.exp(-(hx * hx * f(0.82) + hz * hz * f(1.75)))
            col = lerp(col, F3(0.017, 0.014, 0.012), body * f(0.82))
            col = lerp(col, F3(0.020, 0.019, 0.018), scorch * (1.0 - body) * f(0.16))
            col = lerp(col, F3(0.010, 0.009, 0.008), contact * f(0.22))
            col = col + F3(0.056, 0.014, 0.003) * (ember * glow * f(0.26))
            col = col + F3(0.30, 0.28, 0.24) * (rim * f(0.10))
            alb = lerp(alb, F3(0.030, 0.024, 0.020), body * f(0.75))
            gis = lerp(gis, f(0.48), body * f(0.85))
        core = math.exp(-(hx * hx * f(2.45) + (hz + f(0.18)) ** 2 * f(1.95))) * glow
        col = col + F3(0.55, 0.11, 0.023) * (core * rg * f(0.48))
    elif surf == 2:
        col = F3(0.012, 0.012, 0.012) if cin else F3(0.011, 0.011, 0.010)
        alb, gis = F3(0.24, 0.24, 0.23), f(0.20)
        pa = _line(hx, hz, f(-1.2), f(-2.0), f(-1.2), f(2.0), f(0.010))
        pb = _line(hx, hz, 0.0, f(-2.0), 0.0, f(2.0), f(0.008))
        fl = math.exp(-((hx + f(1.50)) ** 2 + (hz - f(0.92)) ** 2) * f(120.0))
        fr = math.exp(-((hx - f(1.72)) ** 2 + (hz - f(0.80)) ** 2) * f(120.0))
        col = col + F3(0.10, 0.096, 0.088) * ((pa + pb) * f(0.08))
        col = col + F3(0.26, 0.20, 0.13) * ((fl + fr) * f(0.10))
    else:
        col = F3(0.015, 0.015, 0.014) if cin else col
        alb, gis = F3(0.25, 0.25, 0.24), f(0.24)
        base = float(ss(f(0.34), f(0.02), hy))
        crown = float(ss(ceil_y - f(0.32), ceil_y - f(0.02), hy))
        col = lerp(col, F3(0.012, 0.012, 0.011), min(1.0, max(0.0, base * f(0.30) + crown * f(0.20))))
        if surf == 4 and hz > 0:
            recess = _rect(hx, hy, 0.0, f(1.00), f(0.92), f(0.72), f(0.055))
            edge = _line(hx, hy, f(-0.92), f(0.28), f(-0.92), f(1.72), f(0.010)) + _line(hx, hy, f(0.92), f(0.28), f(0.92), f(1.72), f(0.010))
            col = lerp(col, F3(0.015, 0.015, 0.014), recess * f(0.42))
            col = col + F3(0.050, 0.046, 0.038) * (edge * f(0.030))
        soot = math.exp(-(hx * hx * f(0.82) + hz * hz * f(1.12))) * float(ss(f(0.36), f(1.92), hy))
        col = lerp(col, F3(0.010, 0.011, 0.012), soot * f(0.82) * f(prm.smoke_darkness))
        alb = lerp(alb, F3(0.030, 0.032, 0.034), soot * f(0.75))
        wglow = math.exp(-(hx * hx * f(1.1) + hz * hz * f(1.3) + (hy - f(0.70)) ** 2 * f(1.0))) * glow
        col = col + F3(0.15, 0.050, 0.018) * (wglow * f(0.090))
        if surf == 3 and hx < 0:
            win = _rect(hz, hy, f(-1.28), f(0.92), f(0.06), f(0.62), f(0.030))
            glass = f(0.55) + f(0.25) * math.sin(hy * f(23.0) + hz * f(15.0))
            col = lerp(col, F3(0.009, 0.011, 0.012), win * f(0.92))
            col = col + F3(0.10, 0.12, 0.13) * (win * glass * f(0.16))
    vuv = [min(1.0, max(0.0, (hx + f(1.05)) / f(2.10))), min(1.0, max(0.0, (hy - f(0.02)) / f(2.03))), min(1.0, max(0.0, (hz + f(0.82)) / f(1.64)))]
    gi = _light_at(light, vuv)
    sh = _shadow_at(shadow, vuv)
    # volumeShadowRay
    od, src = 0.0, np.array([f(0.50), f(0.12), f(0.50)])
    for i in range(1, 6):
        t = f(i / 6.0)
        uvw = src + (np.array(vuv) - src) * t
        lt = _light_at(light, uvw)
        od += lt[3] * (f(0.18) + t * f(0.22)) + (1.0 - _shadow_at(shadow, uvw)) * (f(0.08) + t * f(0.16))
    vis = math.exp(-od)
    # gatherSceneIrradiance
    irr = np.zeros(3)
    for suv in ((f(0.50), f(0.16), f(0.50)), (f(0.34), f(0.26), f(0.43)), (f(0.66), f(0.26), f(0.57))):
        lt = _light_at(light, suv)
        if lum(*lt[:3]) <= f(0.00001):
            continue
        source = np.array([suv[0] * f(2.10) - f(1.05), suv[1] * f(2.03) + f(0.02), suv[2] * f(1.64) - f(0.82)])
        to = source - hit
        d2 = max(f(0.045), float(to @ to))
        wi = to * (1.0 / math.sqrt(d2))
        ndl = max(0.0, float(nrm @ wi))
        hocc = math.exp(-max(0.0, hy - source[1]) * lt[3] * f(1.15))
        v_ = _shadow_at(shadow, suv) * math.exp(-lt[3] * f(0.92)) * hocc
        geo = (f(0.12) + ndl * f(0.88)) * math.exp(-d2 * f(0.22)) / (f(0.12) + d2)
        irr = irr + lt[:3] * (v_ * geo)
    graze = 1.0 + float(ss(0.0, f(0.45), 1.0 - max(0.0, nrm[1]))) * f(0.36)
    irr = irr * (f(4.20) * graze)
    indirect = (gi[:3] * f(0.42) + irr) * vis
    fall = f(0.62) if surf == 2 else 1.0
    cocc = float(ss(f(0.005), f(0.42), sh)) * math.exp(-gi[3] * f(0.72)) * (f(0.30) + vis * f(0.70))
    lift = f(0.66) + float(ss(f(0.006), f(0.13), lum(*indirect))) * f(0.72)
    warm = f(1.72) if surf == 1 else (f(0.88) if surf == 2 else f(1.16))
    col = col + (indirect * alb) * (gis * fall * cocc * lift * warm)
    if prm.dir_light and dirT is not None:
        # the directional key light's diffuse term on the room's surfaces:
        # albedo x colour x max(0, n . -d) x the transmittance toward the
        # light at the hit x 0.10 (the extension's definition)
        ndl = max(0.0, -float(nrm @ _dir_of(prm)))
        col = col + alb * _dir_colour(prm) * (ndl * _sample(dirT, *vuv) * f(0.10))
    ssoot = (1.0 - vis) * float(ss(f(0.18), f(0.82), 1.0 - sh))
    col = lerp(col, F3(0.006, 0.007, 0.008), ssoot * (f(0.38) if surf == 2 else f(0.26)))
    edge = half - abs(hz) if surf == 3 else half - abs(hx)
    corner = min(half - abs(hx), half - abs(hz)) if surf in (1, 2) else min(edge, min(hy, ceil_y - hy))
    col = col * (f(0.84) + float(ss(0.0, f(0.54), corner)) * f(0.16))
    g = (L.or_hash2(int(math.floor(hx * f(19.0) + hz * f(13.0))), int(math.floor(hy * f(15.0) + time * f(0.08))), L.or_salt(prm.noise_seed, 5)) - f(0.5)) * f(0.006)
    col = col + g
    # sourceModelOverlay
    ux, uy = uv
    if not (ux < f(0.24) or ux > f(0.76) or uy < f(0.42) or uy > f(0.88)) and sid != 1:
        burner = sid == 2 and prm.burner_center_count > 0
        pt = np.array([f(prm.burner_center_x[0]) if burner else f(prm.emitter_center_x),
                       f(prm.burner_center_y[0]) if burner else f(prm.emitter_height_norm) * f(2.03) + f(0.02),
                       f(prm.burner_center_z[0]) if burner else f(prm.emitter_center_z)])
        view = pt - eye
        depth = float(view @ cam["fwd"])
        if depth <= f(0.02):
            cx, cy = -10.0, -10.0
        else:
            cx = 0.5 + float(view @ cam["right"]) / (depth * cam["tan"] * cam["aspect"]) * 0.5
            cy = 0.5 - float(view @ cam["up"]) / (depth * cam["tan"]) * 0.5
        tray = float(ss(f(0.135), f(0.105), abs(ux - cx)) * ss(f(0.050), f(0.036), abs(uy - cy)))
        inner = float(ss(f(0.112), f(0.086), abs(ux - cx)) * ss(f(0.039), f(0.027), abs(uy - cy)))
        alpha = min(1.0, max(0.0, (tray - inner * f(0.55)) * f(0.65)))
        model = lerp(F3(0.015, 0.013, 0.011), F3(0.13, 0.12, 0.10), alpha)
        if alpha > f(0.001):
            col = lerp(col, model, alpha)
    fog = float(ss(f(3.8), f(0.2), best))
    vig = float(ss(f(0.78), f(0.20), math.sqrt((ux - f(0.50)) ** 2 + (uy - f(0.46)) ** 2)))
    return col * ((f(0.12) + fog * f(0.56)) * (f(0.58) + vig * f(0.30)))


ROOM_FEATURES = [
    (-2.55, 1.54, -1.28), (-2.55, 0.30, -1.28), (-2.55, 0.92, -1.22), (-2.55, 0.92, -1.34), (-2.55, 0.92, -1.28),  # window (:1853-1858)
    (-0.92, 1.0, 2.55), (0.92, 0.5, 2.55), (0.0, 1.72, 2.55), (0.92, 1.72, 2.55), (0.5, 0.28, 2.55),  # recess (:1838-1843)
    (0.92, 0.29, 2.55), (-0.92, 0.3, 2.55), (-0.92, 1.71, 2.55),
    (-1.2, 2.18, 0.5), (0.0, 2.18, -0.7), (-1.5, 2.18, 0.92), (1.72, 2.18, 0.8),  # ceiling panels, fixtures (:1820-1827)
    (-1.2, 2.18, -2.0), (-1.2, 2.18, 2.0), (0.0, 2.18, -2.0), (0.0, 2.18, 2.0),  # ends of the ceiling panel lines
    (1.08, 0.0, 0.2), (0.3, 0.0, 0.58), (0.96, 0.0, -0.1), (0.2, 0.0, 0.47),  # tray body and inner edges (:1770-1772)
    (0.47, 0.0, -0.03), (0.05, 0.0, 0.1884), (0.05, 0.0, -0.03), (0.0, 0.0, 0.75),  # burner ring, inner ring, pilot, plate (:1754-1760)
]


@pytest.mark.parametrize("sid,cinematic", [(0, 0), (1, 1), (2, 0), (3, 1)])
def test_room_background_matches_restatement(sid, cinematic):
    """Random rays from cameras inside and outside the room, against a random
    light volume (so the indirect light, contact occlusion and shadow-ray terms
    are live), for every scene's floor and both cinematic settings."""
    grid = (16, 20, 14)
    rng = np.random.default_rng(80 + sid)
    base = dict(BED_PARAMS.get(sid, dict(scene_id=3)))
    base.update(cinematic_mode=cinematic, smoke_darkness=1.3, reflection_gain=1.6, noise_seed=19, emitter_center_x=0.1,
                emitter_center_z=-0.1, emitter_height_norm=0.05)
    lnx, lny, lnz = firesim.light_dims(*grid)
    light = rng.uniform(0, 1, (lnz, lny, lnx, 4))
    light[..., 3] *= 0.6
    shadow = rng.uniform(0.2, 1, (lnz, lny, lnx))
    n_surf = set()
    for cam_d, yaw, pitch in ((1.6, 0.3, 0.1), (2.2, -2.0, 0.4), (3.35, 0.0, 0.18), (1.9, 2.6, -0.3)):
        o, prm = scene(grid, frame=(40, 30), camera_distance=cam_d, camera_yaw=yaw, camera_pitch=pitch, **base)
        o.write("light", light)
        o.write("shadow", shadow)
        o.set_time(0.9, 1.7)
        cam = _camera(prm, 40, 30)
        eye = D3()
        L.or_camera_eye(o.h, eye)
        assert list(eye) == pytest.approx(list(cam["eye"]), rel=1e-12, abs=1e-15)
        out = D3()
        for _ in range(60):
            rd = rng.normal(size=3)
            rd /= np.linalg.norm(rd)
            uv = rng.uniform(0, 1, 2)
            glow = float(rng.uniform(0, 1))
            L.or_room_background(o.h, uv[0], uv[1], D3(*rd), glow, out)
            want = _room_reference(prm, light, shadow, cam, uv, rd, glow, f(1.7))
            assert list(out) == pytest.approx(list(want), rel=1e-12, abs=1e-15), (cam_d, yaw, pitch, rd, uv)
        # rays aimed at the thin features (window and recess edges, ceiling
        # panel lines and fixtures, tray edges, burner rings), with offsets
        # across their smoothstep bands
        for P in ROOM_FEATURES:
            for off in (0.0, 0.004, -0.009, 0.02):
                # (a small fixed offset keeps hits off exact half-cell ties of
                # the nearest-neighbour shadow lookup, where either side is right)
                rd = np.array(P) + off + np.array([0.00137, 0.0, 0.00071]) - cam["eye"]
                rd /= np.linalg.norm(rd)
                uv = rng.uniform(0, 1, 2)
                glow = float(rng.uniform(0, 1))
                L.or_room_background(o.h, uv[0], uv[1], D3(*rd), glow, out)
                want = _room_reference(prm, light, shadow, cam, uv, rd, glow, f(1.7))
                assert list(out) == pytest.approx(list(want), rel=1e-12, abs=1e-15), (P, off)
        o.close()


def _advected(o, prm, F, U, V, W):
    """The MacCormack-advected fields of advectReactKernel (:1002-1018): the
    backtrace from the velocity at the cell centre, the velocity there, and
    sampleScalarMacCormack with the reference's per-field blends, built from
    the pinned sampler and MacCormack primitives."""
    nz, ny, nx = F["heat"].shape
    dt = f(prm.dt)
    anti = min(1.0, max(0.0, f(0.30) + f(prm.detail) * f(0.34)))
    blends = dict(heat=anti, fuel=anti * f(0.92), oxygen=anti * f(0.70), soot=anti, pyrolysis=anti * f(0.72), progress=anti * f(0.82),
                  turbulence=anti * f(0.52), soot_optics=anti * f(0.72))

    def vel(a, b, c):
        return np.array([oracle.sample_grid(U, (nx, ny - 1, nz - 1), a, b, c), oracle.sample_grid(V, (nx - 1, ny, nz - 1), a, b, c),
                         oracle.sample_grid(W, (nx - 1, ny - 1, nz), a, b, c)])

    out = {n: np.empty((nz, ny, nx)) for n in blends}
    k = dt * f(0.62)
    for z, y, x in np.ndindex(nz, ny, nx):
        c = np.array([(x + 0.5) / nx, (y + 0.5) / ny, (z + 0.5) / nz])
        prev = c - vel(*c) * k
        pv = (ctypes.c_double * 3)(*vel(*prev))
        for n, b in blends.items():
            out[n][z, y, x] = L.or_maccormack(oracle._dptr(np.ascontiguousarray(F[n])), nx, ny, nz, dt, F[n][z, y, x], pv, *prev, b)
    return out


def test_react_advection_of_nonuniform_fields_matches_primitives():
    """Advect-react on random fields and velocities: the advected values come
    from the pinned sampler and MacCormack primitives with the reference's
    backtrace scale 0.62 and blends, the reaction from the restatement."""
    grid = (7, 9, 6)
    nx, ny, nz = grid
    rng = np.random.default_rng(91)
    o, prm = scene(grid, detail=1.1, intensity=1.4, turbulence=0.7)
    F = {n: rng.uniform(0.0, 1.5, (nz, ny, nx)) for n in SCALARS}
    F["oxygen"] = rng.uniform(0.2, 1.0, (nz, ny, nx))
    F["heat"] = rng.uniform(0.0, 5.0, (nz, ny, nx))
    U, V, W = rng.uniform(-2, 2, (nz, ny, nx + 1)), rng.uniform(-2, 4, (nz, ny + 1, nx)), rng.uniform(-2, 2, (nz + 1, ny, nx))
    for n, a in list(F.items()) + [("u", U), ("v", V), ("w", W)]:
        o.write(n, a)
    o.set_time(0.5, 0.5)
    o.run_stage(firesim.STAGE_ADVECT_REACT)
    adv = _advected(o, prm, F, U, V, W)
    adv["char"], adv["ash"] = F["char"], F["ash"]
    u, v, w = cell_coords(*grid)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    st = dict(F)
    st["shear"] = np.abs(np.diff(U, axis=2)) + np.abs(np.diff(V, axis=1)) + np.abs(np.diff(W, axis=0))
    want, _ = _react_reference(st, u, v, w, prm, [], (x, y, z), f(0.5), adv=adv)
    for n in SCALARS:
        assert np.allclose(o.read(n), want[n], rtol=1e-11, atol=1e-14), (n, np.abs(o.read(n) - want[n]).max())
    o.close()


# --------------------------------------------------- raymarch (restatement) ----
def _render_pixel_reference(prm, F, light, shadow, cam, x, y, fw, fh, time, dirT=None):
    """renderKernel (:1953-2244) for one pixel, with the build's departures
    (integer-hash jitter of the pixel, step cap 256) and debug modes 0-6.
    Fields, light and noise are sampled at the warped position; the height
    gates use the unwarped fv, as the reference does."""
    sid = prm.scene_id
    nz, ny, nx = F["heat"].shape
    s3 = L.or_salt(prm.noise_seed, 2)

    def S(n, a, b, c):
        return oracle.sample_grid(F[n], (nx - 1, ny - 1, nz - 1), a, b, c)

    def fbm3(a, b, c):
        return L.or_fbm3(a, b, c, s3)

    uv = ((x + 0.5) / fw, (y + 0.5) / fh)
    sx = (uv[0] * 2.0 - 1.0) * cam["aspect"] * cam["tan"]
    sy = (1.0 - uv[1] * 2.0) * cam["tan"]
    rd = cam["fwd"] + cam["right"] * sx + cam["up"] * sy
    rd = rd / math.sqrt(max(f(0.000001), float(rd @ rd)))
    glow = min(1.0, max(0.0, S("heat", f(0.50), f(0.06), f(0.50)) * f(0.42)))
    accum, trans, fglow, socc = np.zeros(3), 1.0, 0.0, 0.0
    dflame = dod = dtemp = dfuel = dvel = 0.0
    eye = cam["eye"]
    # intersectBox (:1867-1907)
    bmin, bmax = np.array([f(-1.05), f(0.02), f(-0.82)]), np.array([f(1.05), f(2.05), f(0.82)])
    tn, tf = 0.0, f(1.0e6)
    for a in range(3):
        d = rd[a]
        inv = 1.0 / (math.copysign(f(0.00001), d) if abs(d) < f(0.00001) else d)
        t0, t1 = (bmin[a] - eye[a]) * inv, (bmax[a] - eye[a]) * inv
        if t0 > t1:
            t0, t1 = t1, t0
        tn, tf = max(tn, t0), min(tf, t1)
    em_h, em_b = f(prm.emitter_height_norm), f(prm.emitter_height_band_norm)
    if tf > tn:
        steps = min(max(prm.raymarch_steps, 24), 256)
        t0 = max(0.0, tn)
        t1 = min(tf, t0 + f(3.05))
        dt_s = (t1 - t0) / steps
        jitter = L.or_hash2(x, y, L.or_salt(prm.noise_seed, 3))
        i = 0
        while i < steps and trans > f(0.018):
            tt = t0 + (i + jitter) * dt_s
            wp = eye + rd * tt
            fu = min(1.0, max(0.0, (wp[0] + f(1.05)) / f(2.10)))
            fv = min(1.0, max(0.0, (wp[1] - f(0.02)) / f(2.03)))
            fwz = min(1.0, max(0.0, (wp[2] + f(0.82)) / f(1.64)))
            fade = L.or_domain_fade(fu, fv, fwz)
            if fade <= f(0.001):
                i += 1
                continue
            cF = S("fuel", fu, fv, fwz)
            coarse = (S("heat", fu, fv, fwz) + cF * f(0.42) + S("soot", fu, fv, fwz) * f(0.34) + S("progress", fu, fv, fwz) * f(0.32)) * fade
            if coarse <= f(0.00150):
                i += 2
                continue
            wa = fbm3(fu * f(7.0) + time * f(0.18), fv * f(5.0) - time * f(0.11), fwz * f(9.0) + time * f(0.07)) - f(0.5)
            wb = fbm3(fwz * f(8.0) - time * f(0.14), fu * f(8.0) + time * f(0.10), fv * f(6.0) - time * f(0.08)) - f(0.5)
            wg = f(0.028) + fv * f(0.060)
            u_, v_, w_ = min(1.0, max(0.0, fu + wa * wg)), min(1.0, max(0.0, fv + (wa - wb) * f(0.020))), min(1.0, max(0.0, fwz + wb * wg))
            q = {n: S(n, u_, v_, w_) for n in SCALARS}
            lt = _light_at(light, (u_, v_, w_))
            sh = _shadow_at(shadow, (u_, v_, w_))
            heat, fuel, oxy, soot, char, ash, pyro, prog, turb, so = (q[n] for n in SCALARS)
            active = (heat + fuel * f(0.42) + soot * f(0.34) + prog * f(0.32) + pyro * f(0.18)) * fade
            if active <= f(0.0025):
                i += 1
                continue
            dfuel = max(dfuel, fade * min(1.0, max(0.0, fuel * f(0.18) + char * f(0.22) + pyro * f(0.30))))
            if prm.render_debug_mode == 6:  # debugVelocity (:2054): the staggered velocity at the warped sample
                vx = oracle.sample_grid(F["u"], (nx, ny - 1, nz - 1), u_, v_, w_)
                vy = oracle.sample_grid(F["v"], (nx - 1, ny, nz - 1), u_, v_, w_)
                vz = oracle.sample_grid(F["w"], (nx - 1, ny - 1, nz), u_, v_, w_)
                dvel = max(dvel, min(1.0, max(0.0, math.sqrt(vx * vx + vy * vy + vz * vz) * f(0.22))))
            cx, cy, cz = f(1.0 / nx), f(1.0 / ny), f(1.0 / nz)
            gx = S("heat", u_ + cx, v_, w_) - S("heat", u_ - cx, v_, w_)
            gy = S("heat", u_, v_ + cy, w_) - S("heat", u_, v_ - cy, w_)
            gz = S("heat", u_, v_, w_ + cz) - S("heat", u_, v_, w_ - cz)
            front = math.sqrt(gx * gx + gy * gy + gz * gz) + abs(fuel - cF) * f(0.62)
            rf = float(ss(f(0.020), f(0.46), front * f(1.35) + fuel * oxy * f(0.040) + prog * f(0.18) + pyro * f(0.080)))
            fine = fbm3(u_ * f(42.0) + time * f(0.72), v_ * f(54.0) - time * f(1.28), w_ * f(43.0) + time * f(0.38))
            shear = fbm3(u_ * f(15.0) - time * f(0.72), v_ * f(13.0) - time * f(1.08), w_ * f(17.0) + time * f(0.32))
            sat1 = lambda a: min(1.0, max(0.0, a))  # noqa: E731
            sst = lambda a, b, c: float(ss(f(a), f(b), c))  # noqa: E731
            plume = sat1(fine * f(0.58) + shear * f(0.42))
            hole = sat1((1.0 - fine) * f(0.36) + shear * f(0.64))
            les = sat1(turb * f(0.48) + prog * f(0.28))
            ragged = sst(0.18, 0.92, fine * (f(0.46) + les * f(0.30)) + shear * f(0.38) + plume * f(0.32) + sat1(active * f(0.08)))
            edge = sst(0.05, 2.30, heat + fuel * f(0.32)) * (f(0.10) + rf * f(0.90))
            fil = sst(0.60, 1.0, fine + les * f(0.10)) * sst(0.38, 0.92, shear + (f(0.72) - fv) * f(0.12) + rf * f(0.10)) * sst(1.0, 0.02, fv) * edge
            srcv = max(0.0, fv - em_h) if sid == 2 else fv
            low = (float(ss(max(f(0.012), em_b * f(0.68)), max(f(0.003), em_b * f(0.10)), abs(fv - em_h))) * sst(0.05, 0.58, fuel + heat * f(0.025)) *
                   sst(0.50, 1.0, oxy)) if sid == 2 else 0.0
            port = low * sst(0.68, 0.98, fine + shear * f(0.18)) if sid == 2 else 0.0
            tongue = sst(0.035, 0.72, fuel + pyro * f(0.50) + char * f(0.12)) * sst(0.76, 0.035, fv) if sid == 1 else 0.0
            traysh = sst(0.018, 0.42, fuel + pyro * f(0.72) + char * f(0.08)) * sst(0.30, 0.015, fv) if sid == 0 else 0.0
            cgate = float(ss(em_h + em_b * f(1.20), em_h - em_b * f(0.08), fv)) if sid == 2 else sst(0.82 if sid == 1 else 0.88, 0.030, fv)
            conv = sst(0.42, 0.88, shear + heat * f(0.040) + plume * f(0.14) + les * f(0.10) + port * f(0.56) + tongue * f(0.20) + traysh * f(0.24)) * cgate * edge
            brk = sst(0.34, 0.88, fine * f(0.48) + shear * f(0.42) + hole * f(0.24) + heat * f(0.026))
            vfade = sst(0.96, 0.025, fv)
            hfade = float(ss(em_h + em_b * f(0.72), em_h - em_b * f(0.05), fv)) if sid == 2 else sst(0.72 if sid == 1 else 0.76, 0.10, fv)
            lwhite = sst(0.060, 0.00, srcv) * sst(0.05, 1.30, pyro + fuel * f(0.20))
            tK = f(293.0) + heat * f(360.0) + fuel * f(44.0) + pyro * f(90.0) + fil * f(560.0) + conv * f(420.0) + lwhite * f(110.0) + prog * f(64.0)
            comb = sst(740.0, 1630.0, tK) * sst(0.016, 0.62, fuel + heat * f(0.055)) * sst(0.045, 0.62, oxy) * vfade * hfade
            thin = sst(0.04, 0.56, front + prog * f(0.10)) * sst(0.02, 0.92, rf)
            tear = sst(0.52, 0.88, hole + shear * f(0.18) + les * f(0.30) + sat1(1.0 - oxy) * f(0.12))
            shole = sst(0.54, 0.90, hole + les * f(0.26) + sat1(1.0 - oxy) * f(0.12))
            sheet = max(sst(0.55, 0.94, fine * f(0.46) + shear * f(0.44) + plume * f(0.18) + rf * f(0.15) + heat * f(0.014) - fv * f(0.050)) *
                        (1.0 - shole * (f(0.74) + les * f(0.12))) * (1.0 - tear * f(0.18)),
                        sst(0.18, 0.030, srcv) * sst(0.75, 2.0, heat) * thin * f(0.030))
            coh = sat1(fil * f(0.86) + conv * f(0.55) + thin * f(0.74))
            conf = sat1(coh + rf * f(0.08))
            boost = (f(0.72) + traysh * f(1.24)) if sid == 0 else 1.0
            fden = (fade * comb * rf * sheet * (f(0.010) + fil * f(0.92) + conv * f(0.24) + thin * f(1.22)) *
                    (f(0.052) + brk * f(0.78) + ragged * f(0.56) + les * f(0.34)) * (f(0.18) + thin * f(2.55)) *
                    (f(0.060) + conf * f(0.94) + port * f(0.72) + tongue * f(0.34)) * hfade * boost)
            void = sst(0.54, 0.90, hole + fine * f(0.26) + shear * f(0.22) + fv * f(0.20))
            top = float(ss(f(0.82) if sid == 1 else f(0.88), f(0.34) if sid == 1 else f(0.42), fv))
            rplume = 1.0 - void * sst(0.18, 0.74, fv) * f(0.96)
            upper = (sst(0.14, 0.48, fv) * sst(0.030, 1.20, so + soot * f(0.42)) *
                     (f(0.025) + fine * f(0.035) + shear * f(0.034) + plume * f(0.34) + ragged * f(0.12) + turb * f(0.035)) * top * rplume)
            sscale = f(0.035) if sid == 2 else (f(0.56) if sid == 1 else 1.0)
            raw = fade * sscale * (sst(0.020, 1.05, so + soot * f(0.32)) * sst(0.16, 0.42, fv) *
                                   (f(0.028) + fine * f(0.040) + shear * f(0.052) + plume * f(0.38) + ragged * f(0.16) + ash * f(0.016)) +
                                   upper * (f(0.22) + plume * f(0.30)) + sst(0.004, 0.34, so) * (f(0.020) + plume * f(0.034)))
            ffm = sat1(fden * f(4.8) + comb * rf * sheet * (f(0.08) + thin * f(0.56) + fil * f(0.18)))
            emis = sat1(ffm)
            rsheet = sst(0.030, 0.32, emis) * sst(0.035, 0.72, oxy)
            smask = 1.0 - sst(0.025, 0.20, rsheet)
            sep = sat1(smask * (1.0 - sst(0.02, 0.16, emis) * f(0.58)))
            clean = 1.0 - sst(0.025, 0.34, rsheet) * f(0.82)
            sheight = sst(0.20, 0.70, fv) * (1.0 - sst(0.02, 0.20, emis))
            neutral = sst(0.18, 0.78, fv) * sep
            absd = raw * clean * (f(1.10) + sheight * f(2.20) + neutral * f(2.40))
            scd = raw * sep * sst(0.10, 0.55, fv + so * f(0.10)) * (f(0.24) + (1.0 - sheight) * f(0.34))
            prad = sat1(f(0.08) + so * f(0.028) + ash * f(0.085) + sat1(1.0 - oxy) * f(0.12))
            sabs = absd * (f(3.10) + prad * f(4.60) + so * f(0.70) + neutral * f(2.85)) * f(prm.smoke_darkness)
            sscat = scd * (f(0.018) + (1.0 - prad) * f(0.050)) * (f(0.20) + f(prm.smoke) * f(0.10))
            sext = sabs + sscat
            fext = fden * (f(0.12) + (1.0 - rsheet) * f(0.12))
            ext = (sext + fext) * dt_s
            salpha = 1.0 - math.exp(-sabs * dt_s)
            scalpha = 1.0 - math.exp(-sscat * dt_s)
            lvis = sh * (f(0.35) + sh * f(0.65))
            irr = lt[:3]
            sload = sat1(so * f(0.74) + soot * f(0.20) + upper * f(0.46))
            white = (sst(2450.0, 3450.0, tK) * comb * sst(0.32, 0.90, oxy) * sst(0.72, 0.02, fv) * (f(0.02) + thin * f(0.48) + fil * f(0.18)) *
                     (1.0 - sst(0.16, 0.74, sload) * f(0.68)))
            rb = sat1((tK - f(760.0)) / f(2000.0))
            rcurve = rb * rb * (f(0.82) + rb * f(0.18))
            rpow = rcurve * (f(1.42) + fil * f(3.15) + conv * f(1.60) + prog * f(0.26) + white * f(0.92))
            core = sat1(white * (f(0.080) + fil * f(0.28) + thin * f(0.32)))
            wfil = core * (f(0.55) + fine * f(0.30) + shear * f(0.18))
            orange = (1.0 - wfil * f(0.72)) * thin * rf * sheet * (f(0.52) + brk * f(0.95)) * sst(0.08, 0.80, oxy)
            bb = D3()
            L.or_blackbody(tK, bb)
            fcol = lerp(np.array([f(1.32), f(0.27), f(0.038)]), np.array(bb), sat1(white * f(0.92) + wfil * f(0.82)))
            if sid == 2:
                blue = port * sst(0.010, 0.11, fden + comb * f(0.18))
                fcol = lerp(np.array([f(0.020), f(0.18), f(2.40)]), np.array([f(0.36), f(0.58), f(1.95)]), sat1(white * f(0.34) + blue * f(0.20)))
            elif sid == 1:
                fcol = lerp(fcol, np.array([f(1.52), f(0.24), f(0.028)]), sat1(tongue * (1.0 - white) * f(0.58)))
            femit = fcol * (fden * rpow * dt_s * (f(1.10) + white * f(0.38)))
            if sid == 2:
                femit = femit * f(0.42)
                femit = femit + np.array([f(0.012), f(0.28), f(5.20)]) * (port * comb * dt_s * f(4.40))
            femit = femit + np.array([f(1.10), f(1.02), f(0.86)]) * (wfil * wfil * fden * rpow * dt_s * f(0.82))
            femit = femit + np.array([f(1.70), f(0.30), f(0.040)]) * (orange * fden * rpow * dt_s * (f(0.06) if sid == 2 else (f(3.20) if sid == 1 else f(3.10))))
            femit = femit + np.array([f(1.36), f(0.070), f(0.008)]) * (fil * (1.0 + tear * f(0.55)) * fden * rpow * dt_s *
                                                                       (f(0.035) if sid == 2 else (f(1.55) if sid == 1 else f(1.10))) * (1.0 - white * f(0.62)))
            vash = sat1(ash * f(0.014) + sst(0.30, 0.88, fv) * (1.0 - sload) * f(0.003))
            scol = lerp(np.array([f(0.008), f(0.0092), f(0.0115)]), np.array([f(0.0018), f(0.0020), f(0.0024)]), sload)
            scol = lerp(scol, np.array([f(0.014), f(0.0165), f(0.0210)]), vash)
            scol = scol * smask
            il = lum(*irr)
            damp = 1.0 - neutral * f(0.985)
            leak = (1.0 - neutral) * (1.0 - sheight) * f(0.18)
            warm = f(0.010) if sid == 2 else (f(0.18) if sid == 1 else f(0.30))
            back = scalpha * lvis * (f(0.00085) + il * f(0.0011) * damp + glow * f(0.00010) * warm) * (f(0.018) + (1.0 - prad) * f(0.018)) * sep
            scol = scol + np.array([f(0.010), f(0.012), f(0.016)]) * back
            sgain = scalpha * (f(0.00046) + lvis * f(0.00072) * damp + fden * f(0.00008) * warm + turb * f(0.00018)) * (f(0.038) + (1.0 - prad) * f(0.045)) * sep
            scene_sc = scalpha * sep * (f(0.0065) + (1.0 - prad) * f(0.014)) * lvis
            neutral_l = np.array([il * f(0.052), il * f(0.064), il * f(0.086)])
            sc_light = lerp(neutral_l, irr, leak)
            coal = sst(0.004, 0.18, char) * sst(0.54, 0.02, fv) * (f(0.070) + pyro * f(0.12) + heat * f(0.009))
            dflame = max(dflame, sat1(fden * rpow * f(0.26)))
            dod = sat1(dod + sext * dt_s)
            dtemp = max(dtemp, sat1((tK - f(520.0)) / f(1900.0)))
            accum = accum + femit * trans
            accum = accum + np.array([1.0, f(0.24), f(0.035)]) * (trans * coal * dt_s)
            accum = accum + scol * (trans * sgain * smask)
            accum = accum + sc_light * (trans * scene_sc * (f(0.18) + damp * f(0.42)))
            if prm.dir_light and dirT is not None:
                # single scattering of the directional light: colour x the view
                # transmittance x the sample's scattering opacity x the
                # transmittance toward the light x 2.0 (the extension's definition)
                accum = accum + _dir_colour(prm) * (trans * scalpha * _sample(dirT, u_, v_, w_) * f(2.0))
            fglow = max(fglow, fden * rpow * (f(0.12) + fil * f(0.18) + wfil * f(0.20)))
            socc = sat1(socc + salpha * trans)
            trans *= math.exp(-ext)
            i += 1
    mode = prm.render_debug_mode
    if mode:
        def heatcol(v):  # debugHeatColor (:1909-1929)
            return _debug_heat_color(v)
        if mode == 1:
            return np.append(heatcol(dflame), 1.0)
        if mode == 2:
            od = min(1.0, max(0.0, dod))
            return np.array([od, od * f(0.90), od * f(0.78), 1.0])
        if mode == 3:
            return np.array([trans, trans, trans, 1.0])
        if mode == 4:
            return np.append(heatcol(dtemp), 1.0)
        if mode == 5:
            return np.array([dfuel * f(0.82), dfuel * f(0.48), dfuel * f(0.18), 1.0])
        return np.array([dvel * f(0.20), dvel * f(0.72), dvel, 1.0])
    col = accum
    if trans > f(0.006):
        col = _room_reference(prm, light, shadow, cam, uv, rd, glow, time, dirT) * trans + accum
    bloom = min(1.0, max(0.0, fglow * f(0.10))) * (1.0 - min(1.0, max(0.0, socc * f(0.95))))
    col = col + np.array([1.0, f(0.25), f(0.045)]) * (bloom * bloom * f(0.017) * f(prm.reflection_gain))
    return np.append(np.maximum(0.0, col), 1.0)


def _debug_heat_color(value):
    """debugHeatColor (:370-379)."""
    value = min(1.0, max(0.0, value))
    red, orange, white = float(ss(f(0.05), f(0.45), value)), float(ss(f(0.24), f(0.78), value)), float(ss(f(0.72), f(1.0), value))
    return np.array([min(1.0, red + white * f(0.70)), min(1.0, orange * f(0.58) + white * f(0.78)), min(1.0, white * f(0.62))])


# (2, 0, narrow band): a burner band so narrow that the port jet's width floor
# (0.012 of the height) takes over from 0.68 of the band.
# (0, 0, _heat): heat weak enough that the glow sampled at (0.5, 0.06, 0.5) for
# the room stays below saturation, so its sample point and gain show.
# (3, 0, _scale): every scalar so thin that the coarse density along the rays
# crosses the empty-space skip threshold (0.0015) and the activity threshold.
# (3, 6): the velocity view over random staggered velocities.
@pytest.mark.parametrize("sid,mode,extra", [(0, 0, {}), (1, 0, {}), (2, 0, {}), (2, 0, dict(emitter_height_band_norm=0.008)), (3, 0, {}), (3, 1, {}),
                                            (3, 2, {}), (3, 4, {}), (0, 5, {}), (3, 6, {}), (0, 0, dict(_heat=0.12)),
                                            (3, 0, dict(_scale=0.0012))])
def test_render_pixels_match_restatement(sid, mode, extra):
    """Selected pixels of a small frame over a fire blob (flame, soot, coal glow,
    room behind), every scene's flame model and the debug views."""
    grid = (18, 22, 16)
    fw, fh = 24, 18
    extra = dict(extra)
    heat_k, all_k = extra.pop("_heat", None), extra.pop("_scale", None)
    rng = np.random.default_rng(100 + sid)
    base = dict(BED_PARAMS.get(sid, dict(scene_id=3)))
    base.update(render_debug_mode=mode, raymarch_steps=40, ember_count=0, noise_seed=23, camera_distance=2.6, camera_pitch=0.12,
                reflection_gain=1.4, smoke=1.1, smoke_darkness=1.2, **extra)
    o, prm = scene(grid, frame=(fw, fh), **base)
    F = _smooth_fields(grid, rng)
    nx, ny, nz = grid
    if heat_k is not None:
        F["heat"] = F["heat"] * heat_k
        glow = min(1.0, max(0.0, oracle.sample_grid(F["heat"], (nx - 1, ny - 1, nz - 1), f(0.50), f(0.06), f(0.50)) * f(0.42)))
        assert 0.05 < glow < 0.95, glow
    if all_k is not None:
        F = {n: a * all_k if n != "oxygen" else a for n, a in F.items()}
        assert 0.002 < F["heat"].max() < 0.02
    if mode == 6:
        F["u"] = rng.uniform(-2, 2, (nz, ny, nx + 1))
        F["v"] = rng.uniform(-1, 4, (nz, ny + 1, nx))
        F["w"] = rng.uniform(-2, 2, (nz + 1, ny, nx))
    for n, a in F.items():
        o.write(n, a)
    o.set_time(0.45, 0.45)
    o.run_stage(firesim.STAGE_LIGHT)
    lnx, lny, lnz = firesim.light_dims(*grid)
    light = o.read("light").reshape(lnz, lny, lnx, 4)
    shadow = o.read("shadow").reshape(lnz, lny, lnx)
    hdr = o.render_hdr()
    cam = _camera(prm, fw, fh)
    pix = [(x, y) for y in range(2, fh, 3) for x in range(1, fw, 3)]
    emissive = 0
    for x, y in pix:
        want = _render_pixel_reference(prm, F, light, shadow, cam, x, y, fw, fh, f(0.45))
        assert list(hdr[y, x]) == pytest.approx(list(want), rel=1e-12, abs=1e-15), (sid, mode, x, y)
        emissive += want[0] > 0.05
    if mode == 0 and heat_k is None and all_k is None:
        assert emissive >= 3, "the blob must emit in some of the pixels checked"
    if mode == 6:
        assert emissive >= 3, "the velocity view must show in some of the pixels checked"
    o.close()


@pytest.mark.parametrize("sid", [0, 1, 2])
def test_reset_beds_match_restatement(sid):
    """resetScalarsKernel (:805-847): every scalar from the bed material, its
    source at time 0 and the chunk noise, with the per-scene scales; oxygen 1,
    velocity, pressure and divergence 0."""
    grid = (24, 8, 20)
    o, prm = scene(grid, noise_seed=29, **BED_PARAMS[sid])
    u, v, w = cell_coords(*grid)
    salt = L.or_salt(prm.noise_seed, 1)
    want = {n: np.zeros(u.shape) for n in SCALARS}
    for idx in np.ndindex(u.shape):
        wx, wz = (u[idx] - f(0.50)) * f(2.10), (w[idx] - f(0.50)) * f(1.64)
        mat = _bed_material(wx, wz, prm)
        chunk = L.or_fbm2(wx * f(18.0) + wz * f(2.0), wz * f(22.0) - wx * f(5.0), salt)
        bed = _bed_source(wx, wz, v[idx], prm, 0.0)
        cs = f(0.004) if sid == 2 else (f(1.18) if sid == 1 else 1.0)
        ssc = f(0.004) if sid == 2 else (f(0.92) if sid == 1 else 1.0)
        vals = dict(
            heat=bed * (f(1.34) if sid == 2 else (f(0.78) + chunk * f(0.26) if sid == 1 else f(0.86) + chunk * f(0.24))),
            fuel=bed * (f(1.42) if sid == 2 else (f(1.05) + chunk * f(0.30) if sid == 1 else f(1.08) + chunk * f(0.26))),
            oxygen=1.0,
            soot=bed * (f(0.00065) if sid == 2 else f(0.018)) * ssc,
            char=mat * (f(0.92) + chunk * f(0.72)) * cs,
            ash=mat * (f(0.018) + float(ss(f(0.56), f(0.86), chunk)) * f(0.055)) * (f(0.10) if sid == 2 else 1.0),
            pyrolysis=bed * (f(0.018) if sid == 2 else (f(0.060) + chunk * f(0.034) if sid == 1 else f(0.058) + chunk * f(0.028))),
            progress=bed * (f(0.11) if sid == 2 else (f(0.24) if sid == 1 else f(0.21))),
            turbulence=bed * (f(0.012) if sid == 2 else (f(0.082) if sid == 1 else f(0.065))),
            soot_optics=bed * (f(0.00050) if sid == 2 else f(0.018)) * ssc)
        for n in SCALARS:
            want[n][idx] = vals[n] if mat != 0.0 or n == "oxygen" else (1.0 if n == "oxygen" else 0.0)
    assert want["char"].max() > 0 and want["heat"].max() > 0
    for n in SCALARS:
        assert np.allclose(o.read(n), want[n], rtol=1e-12, atol=1e-15), (sid, n)
    for n in ("u", "v", "w", "pressure", "divergence", "light", "shadow"):
        assert np.all(o.read(n) == 0)
    o.close()


def test_reset_zeroes_the_clocks():
    """resetSimulation (:3053-3055): simulation time, render time and frame
    index return to 0, whatever they were."""
    o, _ = scene((8, 8, 8), **BED_PARAMS[1])
    o.step(2)
    o.set_time(1.3, 2.7)
    assert o.get_time()[:2] == pytest.approx((1.3, 2.7))
    o.reset()
    assert o.get_time() == (0.0, 0.0, 0)
    o.close()


def _embers_reference(prm, F, cam, fw, fh, time):
    """emberHdrKernel (:2323-2452) into a zero frame, spark after spark, with
    the build's integer hash of (spark, k) in place of hash21(spark, const)."""
    sid = prm.scene_id
    nz, ny, nx = F["heat"].shape
    out = np.zeros((fh, fw, 3))
    salt = L.or_salt(prm.noise_seed, 6)
    sat1 = lambda a: min(1.0, max(0.0, a))  # noqa: E731

    def S(n, a, b, c):
        return oracle.sample_grid(F[n], (nx - 1, ny - 1, nz - 1), a, b, c)

    def vel(a, b, c):
        return (oracle.sample_grid(F["u"], (nx, ny - 1, nz - 1), a, b, c), oracle.sample_grid(F["v"], (nx - 1, ny, nz - 1), a, b, c),
                oracle.sample_grid(F["w"], (nx - 1, ny - 1, nz), a, b, c))

    def project(pt):
        view = np.array(pt) - cam["eye"]
        depth = float(view @ cam["fwd"])
        if depth <= f(0.02):
            return (-10.0, -10.0, -1.0)
        return (0.5 + float(view @ cam["right"]) / (depth * cam["tan"] * cam["aspect"]) * 0.5,
                0.5 - float(view @ cam["up"]) / (depth * cam["tan"]) * 0.5, depth)

    glow = sat1(S("heat", f(0.50), f(0.06), f(0.50)) * f(0.42))
    gate = float(ss(f(0.02), f(0.30), glow))
    if gate <= f(0.0001) or prm.render_debug_mode != 0:
        return out
    for i in range(min(max(prm.ember_count, 0), 192)):
        h0, h1, h2, h3 = (L.or_hash2(i, k, salt) for k in range(4))
        if (sid == 1 and h2 > f(0.34)) or (sid == 2 and h0 > f(0.14)):
            continue
        lt = time * (f(0.20) + h2 * f(0.18) + (f(0.36) if sid == 2 else 0.0)) + h0
        life = lt - math.floor(lt)
        ang = h0 * f(6.2831853)
        if sid == 2:
            R = max(f(0.04), f(prm.emitter_radius))
            cx = f(prm.burner_center_x[0]) if prm.burner_center_count > 0 else f(prm.emitter_center_x)
            cz = f(prm.burner_center_z[0]) if prm.burner_center_count > 0 else f(prm.emitter_center_z)
            sx0, sz0 = cx + math.cos(ang) * R * (f(0.82) + h1 * f(0.18)), cz + math.sin(ang) * R * (f(0.82) + h1 * f(0.18))
        elif sid == 1:
            la = _capsule(math.cos(ang) * f(0.34), math.sin(ang) * f(0.25), f(-0.06), f(0.01), f(0.38), f(0.34), f(0.095))
            lb = _capsule(math.cos(ang + f(1.7)) * f(0.35), math.sin(ang + f(1.7)) * f(0.25), f(0.07), f(-0.02), f(-0.58), f(0.33), f(0.092))
            cb = sat1(la + lb + h3 * f(0.35))
            r = f(0.16) + (f(0.46) - f(0.16)) * math.sqrt(h1)
            sx0, sz0 = math.cos(ang) * r * (f(0.64) + cb * f(0.36)), math.sin(ang) * r * (f(0.46) + cb * f(0.26))
        else:
            sx0, sz0 = math.cos(ang) * f(0.48) * math.sqrt(h1), math.sin(ang) * f(0.48) * math.sqrt(h1)
        su, sw = sat1(sx0 / f(2.10) + f(0.50)), sat1(sz0 / f(1.64) + f(0.50))
        sv = f(prm.emitter_height_norm) if sid == 2 else (f(0.035) + h2 * f(0.055) if sid == 1 else f(0.045))
        heat, char, pyro, turb = S("heat", su, sv, sw), S("char", su, sv, sw), S("pyrolysis", su, sv, sw), S("turbulence", su, sv, sw)
        vx, _, vz = vel(su, sv, sw)
        mgate = float(ss(f(0.018), f(0.40), char + pyro * f(0.34) + heat * f(0.075)))
        if mgate <= f(0.0001):
            continue
        tK = f(620.0) + heat * f(330.0) + pyro * f(210.0) + char * f(42.0)
        lift = f(0.86) if sid == 1 else (f(0.16) if sid == 2 else f(0.80))
        t = life * (f(1.24) if sid == 1 else (f(0.34) if sid == 2 else f(1.10)))
        dk = f(4.8) if sid == 2 else f(1.35)
        rise, fall = f(0.26) + h3 * f(0.62) + turb * f(0.11), f(0.22) + h2 * f(0.16)
        dx_, dz_ = vx * f(0.34) + (h1 - f(0.5)) * f(0.20) + f(prm.wind_x) * f(0.26), vz * f(0.28) + (h2 - f(0.5)) * f(0.18)
        world = (sx0 + dx_ * t * math.exp(-t * dk), sv * f(2.03) + f(0.02) + rise * t * lift - fall * t * t, sz0 + dz_ * t * math.exp(-t * dk))
        pt = max(0.0, t - f(0.045))
        prev = (sx0 + dx_ * pt * math.exp(-pt * dk), sv * f(2.03) + f(0.02) + rise * pt * lift - fall * pt * pt, sz0 + dz_ * pt * math.exp(-pt * dk))
        sc, ps = project(world), project(prev)
        if sc[2] <= 0.0 or sc[0] < f(-0.05) or sc[0] > f(1.05) or sc[1] < f(-0.05) or sc[1] > f(1.05):
            continue
        size = float(ss(f(720.0), f(1580.0), tK)) * mgate
        rad = ((f(0.00042) if sid == 2 else f(0.00086)) + h3 * (f(0.00042) if sid == 2 else f(0.00145))) * (f(0.72) + size * f(0.90)) / max(f(0.55), sc[2])
        cxp, cyp = sc[0] * fw - f(0.5), sc[1] * fh - f(0.5)
        tr = np.array([sc[0] - ps[0], sc[1] - ps[1]])
        tl = math.sqrt(float(tr @ tr))
        capped = min(tl, f(0.016) if sid == 1 else f(0.005))
        td = tr / tl if tl > f(0.000001) else np.array([0.0, -1.0])
        pad = max(2, int(math.ceil((rad + capped) * max(fw, fh) * f(3.0))))
        x0, x1 = max(0, int(math.floor(cxp)) - pad), min(fw - 1, int(math.ceil(cxp)) + pad)
        y0, y1 = max(0, int(math.floor(cyp)) - pad), min(fh - 1, int(math.ceil(cyp)) + pad)
        hot = np.array([f(0.36), f(0.58), f(1.04)]) if sid == 2 else np.array([1.0, f(0.42), f(0.090)])
        cool = np.array([f(0.055), f(0.10), f(0.18)]) if sid == 2 else np.array([f(0.18), f(0.026), f(0.008)])
        colr = lerp(hot, cool, float(ss(f(0.16), f(0.92), life)) * (1.0 - sat1(heat * f(0.045))))
        fade = (float(ss(1.0, f(0.08), life) * ss(0.0, f(0.16), life)) * gate * mgate * size * (f(0.055) if sid == 2 else (f(0.58) if sid == 1 else f(0.82))))
        streak = rad * (1.0 + capped * f(130.0))
        for y in range(y0, y1 + 1):
            for x in range(x0, x1 + 1):
                d = np.array([(x + 0.5) / fw - sc[0], (y + 0.5) / fh - sc[1]])
                along = float(d @ td)
                ac = d - td * along
                core = math.exp(-(float(ac @ ac) / max(f(0.0000002), rad * rad) + along * along / max(f(0.0000002), streak * streak)))
                halo = math.exp(-float(d @ d) / max(f(0.0000002), rad * rad * f(8.0))) * f(0.16)
                spark = (core + halo) * fade
                if spark > f(0.00001):
                    out[y, x] += colr * spark
    return out


# amp 0.25: fields weak enough that many sparks sit on the ramp of their
# material gate instead of at its top. _heat 0.05: heat alone so weak that the
# glow sampled at (0.5, 0.06, 0.5) lies on the ramp of the spark gate (0.02 to
# 0.30) while char and pyrolysis keep the material gates open. The burner with
# its smallest radius (0.035) starts sparks inside the 0.04 floor.
@pytest.mark.parametrize("sid,amp,extra", [(0, 1.0, {}), (1, 1.0, {}), (2, 1.0, {}), (3, 1.0, {}), (0, 0.25, {}), (1, 0.25, {}),
                                           (0, 1.0, dict(_heat=0.05)), (3, 1.0, dict(_heat=0.05)), (2, 1.0, dict(emitter_radius=0.035))])
def test_embers_match_restatement(sid, amp, extra):
    """The ember increments of the HDR frame (frame with sparks minus frame
    without) equal the restated splat of every spark."""
    grid = (18, 22, 16)
    # the burner's sparks are a few thousandths of the frame wide: a finer frame catches them
    fw, fh = (320, 240) if sid == 2 else (48, 36)
    rng = np.random.default_rng(120 + sid)
    base = dict(BED_PARAMS.get(sid, dict(scene_id=3)))
    extra = dict(extra)
    heat_k = extra.pop("_heat", None)
    base.update(raymarch_steps=24, ember_count=0, noise_seed=31, camera_distance=2.4, wind_x=0.3, **extra)
    o, prm = scene(grid, frame=(fw, fh), **base)
    F = _smooth_fields(grid, rng)
    F["heat"] = F["heat"] * 1.5
    for n in ("heat", "char", "pyrolysis"):
        F[n] = F[n] * amp
    nx, ny, nz = grid
    if heat_k is not None:
        F["heat"] = F["heat"] * heat_k
        glow = min(1.0, max(0.0, oracle.sample_grid(F["heat"], (nx - 1, ny - 1, nz - 1), f(0.50), f(0.06), f(0.50)) * f(0.42)))
        assert 0.02 < glow < 0.30, glow
    F["u"] = rng.uniform(-1, 1, (nz, ny, nx + 1))
    F["v"] = rng.uniform(-1, 2, (nz, ny + 1, nx))
    F["w"] = rng.uniform(-1, 1, (nz + 1, ny, nx))
    for n, a in F.items():
        o.write(n, a)
    o.set_time(1.3, 2.2)
    plain = o.render_hdr()
    p2 = o.get_params()
    p2.ember_count = 160
    o.set_params(p2)
    prm = o.get_params()
    sparks = o.render_hdr() - plain
    want = _embers_reference(prm, F, _camera(prm, fw, fh), fw, fh, f(2.2))
    assert want.max() > 1e-4, "some sparks must draw"
    assert np.allclose(sparks[..., :3], want, rtol=1e-9, atol=1e-13)
    assert np.all(sparks[..., 3] == 0)
    o.close()


def test_tonemap_matches_restatement():
    """tonemapKernel (:2276-2297): camera response, a dither of
    (hash - 0.5) / 255 per pixel (integer hash of the pixel), truncating 8-bit
    channels and the 0xff alpha, packed as B, G, R, A bytes."""
    o, prm = scene((6, 6, 6), frame=(40, 30), exposure=1.3)
    rng = np.random.default_rng(130)
    hdr = np.concatenate([rng.uniform(0, 3, (30, 40, 3)) ** 2, np.ones((30, 40, 1))], -1)
    hdr[0, :5, :3] = 0.0
    got = o.tonemap(hdr)
    salt = L.or_salt(prm.noise_seed, 4)
    out = D3()
    for y in range(30):
        for x in range(40):
            L.or_camera_response(D3(*hdr[y, x, :3]), f(prm.exposure), out)
            d = (L.or_hash2(x, y, salt) - f(0.5)) / f(255.0)
            r, g, b = (int(min(1.0, max(0.0, c + d)) * f(255.0)) for c in out)
            assert int(got[y, x]) == (0xFF << 24) | (r << 16) | (g << 8) | b, (x, y)
    o.close()


# ------------------------------------------- directional light (restatement) ----
DIR_LIGHT = dict(dir_light=1, dir_light_x=0.35, dir_light_y=-1.0, dir_light_z=0.25, dir_light_r=1.6, dir_light_g=1.5, dir_light_b=1.35)


@pytest.mark.parametrize("sid", [0, 3])
def test_room_background_with_directional_light_matches_restatement(sid):
    """The room colour with the directional key light on, against a random
    transmittance volume: the restatement adds albedo x colour x max(0, n.-d)
    x T x 0.10 to the pinned room colour, so a slip in the gain, the sign of
    n.d or the sampling point fails here."""
    grid = (16, 20, 14)
    rng = np.random.default_rng(180 + sid)
    base = dict(BED_PARAMS.get(sid, dict(scene_id=3)))
    base.update(cinematic_mode=1, noise_seed=19, **DIR_LIGHT)
    lnx, lny, lnz = firesim.light_dims(*grid)
    light = rng.uniform(0, 1, (lnz, lny, lnx, 4))
    light[..., 3] *= 0.6
    shadow = rng.uniform(0.2, 1, (lnz, lny, lnx))
    dirT = rng.uniform(0.05, 1, (lnz, lny, lnx))
    o, prm = scene(grid, frame=(40, 30), camera_distance=1.9, camera_yaw=0.7, camera_pitch=0.2, **base)
    o.write("light", light)
    o.write("shadow", shadow)
    o.write("dir_transmittance", dirT)
    o.set_time(0.9, 1.7)
    cam = _camera(prm, 40, 30)
    out = D3()
    lit = 0
    for _ in range(120):
        rd = rng.normal(size=3)
        rd /= np.linalg.norm(rd)
        uv = rng.uniform(0, 1, 2)
        glow = float(rng.uniform(0, 1))
        L.or_room_background(o.h, uv[0], uv[1], D3(*rd), glow, out)
        want = _room_reference(prm, light, shadow, cam, uv, rd, glow, f(1.7), dirT)
        assert list(out) == pytest.approx(list(want), rel=1e-12, abs=1e-15), rd
        lit += np.any(np.abs(want - _room_reference(prm, light, shadow, cam, uv, rd, glow, f(1.7))) > 1e-6)
    assert lit > 30, "the key light must reach many of the surfaces hit"
    o.close()


@pytest.mark.parametrize("sid", [1, 3])
def test_render_pixels_with_directional_light_match_restatement(sid):
    """The raymarch with the directional light: the restatement adds colour x
    trans x the sample's scattering opacity x T x 2.0 at every active sample
    (and the room's diffuse term behind), on the light pass's own
    transmittance volume."""
    grid = (18, 22, 16)
    fw, fh = 24, 18
    rng = np.random.default_rng(200 + sid)
    base = dict(BED_PARAMS.get(sid, dict(scene_id=3)))
    base.update(raymarch_steps=40, ember_count=0, noise_seed=23, camera_distance=2.6, camera_pitch=0.12, reflection_gain=1.4, smoke=1.1,
                smoke_darkness=1.2, **DIR_LIGHT)
    o, prm = scene(grid, frame=(fw, fh), **base)
    F = _smooth_fields(grid, rng)
    for n, a in F.items():
        o.write(n, a)
    o.set_time(0.45, 0.45)
    o.run_stage(firesim.STAGE_LIGHT)
    lnx, lny, lnz = firesim.light_dims(*grid)
    light = o.read("light").reshape(lnz, lny, lnx, 4)
    shadow = o.read("shadow").reshape(lnz, lny, lnx)
    dirT = o.read("dir_transmittance").reshape(lnz, lny, lnx)
    assert dirT.min() < 0.9 and dirT.max() > 0.1
    hdr = o.render_hdr()
    cam = _camera(prm, fw, fh)
    changed = 0
    for y in range(2, fh, 3):
        for x in range(1, fw, 3):
            want = _render_pixel_reference(prm, F, light, shadow, cam, x, y, fw, fh, f(0.45), dirT)
            assert list(hdr[y, x]) == pytest.approx(list(want), rel=1e-12, abs=1e-15), (sid, x, y)
            without = _render_pixel_reference(prm, F, light, shadow, cam, x, y, fw, fh, f(0.45))
            changed += np.any(np.abs(want - without) > 1e-6)
    assert changed > 10, "the key light must change many of the pixels checked"
    o.close()
