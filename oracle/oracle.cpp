// oracle.cpp - CPU FP64 restatement of the reference fire/smoke step.
//
// TEST INFRASTRUCTURE ONLY. This library is the correctness oracle and the
// CPU baseline. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it; the product library never links or calls it.
//
// Every pass follows one reference kernel in src/fire_cuda.cu (cited per
// function). Arithmetic is FP64 throughout; the reference's float literals are
// kept as float literals, so each constant has exactly the reference's value
// and is then used in double arithmetic. Host bookkeeping that the reference
// does in FP32 (the time accumulators, src/fire_cuda.cu:3136-3139) is done in
// FP32 here as well.
//
// Documented departures from the reference (the same ones the HIP path makes):
//  * noise: the reference's fract(sin(x)*43758.5453) lattice hash
//    (src/fire_cuda.cu:381-388) is replaced by the PCG integer hash of the
//    integer lattice point (Jarzynski & Olano, "Hash Functions for GPU
//    Rendering", JCGT 2020), mapped to [0,1) by its top 24 bits. The value-noise
//    interpolation and fbm octave constants are the reference's. Render jitter
//    and tonemap dither hash the integer pixel coordinates.
//  * force: the racy atomicAdd scatter of forceVelocityKernel (:1178-1182) is
//    computed as a gather from the pre-force fields: u face x receives the
//    contributions of cells x-1 and x, v face y that of cell y-1, w face z those
//    of cells z-1 and z, summed in that order.
//  * scene 3 has no procedural bed (material 0); fs_emitter spheres inject
//    fuel and heat at the reference's interactive-source point (:1047-1063) and
//    keep their cells out of the cold-air early-out (:984-989).
//  * wind_z drives w with the same height profile the reference uses for
//    wind on u (:1179); SOR omega and iteration count are parameters.
//
// Every function is pinned by tests/test_oracle.py (primitives) and
// tests/test_oracle_pins.py (passes: closed forms and independent numpy
// restatements from src/fire_cuda.cu); oracle/mutation_sweep.py checks that
// one-constant slips fail those tests.

#include "../include/firesim.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef double R;

struct V3 {
    R x, y, z;
};
struct V4 {
    R x, y, z, w;
};
static inline V3 v3(R x, R y, R z) { return V3{x, y, z}; }
static inline V3 add3(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline V3 sub3(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline V3 mul3(V3 a, R s) { return v3(a.x * s, a.y * s, a.z * s); }
static inline V3 mul3v(V3 a, V3 b) { return v3(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline R dot3(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline V3 cross3(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}

// src/fire_cuda.cu:219-225
static inline R sat(R v) { return std::min<R>(1.0, std::max<R>(0.0, v)); }
static inline R lerp(R a, R b, R t) { return a + (b - a) * t; }
static inline V3 lerp3(V3 a, V3 b, R t) { return add3(a, mul3(sub3(b, a), t)); }
static inline R luminance3(V3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }  // :274-276

// src/fire_cuda.cu:285-288
static inline V3 normalize3(V3 v) {
    const R len = std::sqrt(std::max<R>(0.000001f, dot3(v, v)));
    return v3(v.x / len, v.y / len, v.z / len);
}

// Non-recursive smoothstep with inverted-edge handling, src/fire_cuda.cu:227-234.
static inline R smoothstep(R e0, R e1, R x) {
    const bool inverted = e1 < e0;
    const R lo = inverted ? e1 : e0;
    const R hi = inverted ? e0 : e1;
    const R t = sat((x - lo) / std::max<R>(0.000001f, hi - lo));
    const R value = t * t * (3.0 - 2.0 * t);
    return inverted ? 1.0 - value : value;
}

// ---------------------------------------------------------------- noise ----
// PCG hash (Jarzynski & Olano 2020).
static inline uint32_t pcg(uint32_t v) {
    const uint32_t state = v * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
static inline uint32_t salt_of(uint32_t seed, uint32_t k) { return pcg(seed + k * 0x9E3779B9u); }
enum { kSaltNoise2 = 1, kSaltNoise3 = 2, kSaltJitter = 3, kSaltDither = 4, kSaltGrain = 5, kSaltEmber = 6 };

static inline R unit24(uint32_t h) { return static_cast<R>(h >> 8) * (1.0 / 16777216.0); }
static inline R hash2(int32_t x, int32_t y, uint32_t salt) {
    return unit24(pcg(pcg(salt + static_cast<uint32_t>(y)) + static_cast<uint32_t>(x)));
}
static inline R hash3(int32_t x, int32_t y, int32_t z, uint32_t salt) {
    return unit24(pcg(pcg(pcg(salt + static_cast<uint32_t>(z)) + static_cast<uint32_t>(y)) + static_cast<uint32_t>(x)));
}

// Value noise with the reference's smoothstep weights, src/fire_cuda.cu:390-399.
static inline R value_noise2(R px, R py, uint32_t salt) {
    const R fx0 = std::floor(px), fy0 = std::floor(py);
    const int32_t ix = static_cast<int32_t>(fx0), iy = static_cast<int32_t>(fy0);
    const R fx = px - fx0, fy = py - fy0;
    const R a = hash2(ix, iy, salt);
    const R b = hash2(ix + 1, iy, salt);
    const R c = hash2(ix, iy + 1, salt);
    const R d = hash2(ix + 1, iy + 1, salt);
    const R ux = fx * fx * (3.0 - 2.0 * fx);
    const R uy = fy * fy * (3.0 - 2.0 * fy);
    return lerp(lerp(a, b, ux), lerp(c, d, ux), uy);
}

// src/fire_cuda.cu:401-421
static inline R value_noise3(R px, R py, R pz, uint32_t salt) {
    const R fx0 = std::floor(px), fy0 = std::floor(py), fz0 = std::floor(pz);
    const int32_t ix = static_cast<int32_t>(fx0), iy = static_cast<int32_t>(fy0), iz = static_cast<int32_t>(fz0);
    const R fx = px - fx0, fy = py - fy0, fz = pz - fz0;
    const R ux = fx * fx * (3.0 - 2.0 * fx);
    const R uy = fy * fy * (3.0 - 2.0 * fy);
    const R uz = fz * fz * (3.0 - 2.0 * fz);
    const R c000 = hash3(ix, iy, iz, salt);
    const R c100 = hash3(ix + 1, iy, iz, salt);
    const R c010 = hash3(ix, iy + 1, iz, salt);
    const R c110 = hash3(ix + 1, iy + 1, iz, salt);
    const R c001 = hash3(ix, iy, iz + 1, salt);
    const R c101 = hash3(ix + 1, iy, iz + 1, salt);
    const R c011 = hash3(ix, iy + 1, iz + 1, salt);
    const R c111 = hash3(ix + 1, iy + 1, iz + 1, salt);
    const R c00 = lerp(c000, c100, ux);
    const R c10 = lerp(c010, c110, ux);
    const R c01 = lerp(c001, c101, ux);
    const R c11 = lerp(c011, c111, ux);
    return lerp(lerp(c00, c10, uy), lerp(c01, c11, uy), uz);
}

// 5 octaves, src/fire_cuda.cu:423-432
static inline R fbm2(R px, R py, uint32_t salt) {
    R sum = 0.0, amp = 0.5;
    for (int i = 0; i < 5; ++i) {
        sum += amp * value_noise2(px, py, salt);
        const R nx = px * 2.03f + 17.7f;
        const R ny = py * 2.01f - 9.2f;
        px = nx;
        py = ny;
        amp *= 0.5;
    }
    return sum;
}

// 3 octaves then x1.142857 and saturate, src/fire_cuda.cu:434-443
static inline R fbm3(R px, R py, R pz, uint32_t salt) {
    R sum = 0.0, amp = 0.5;
    for (int i = 0; i < 3; ++i) {
        sum += amp * value_noise3(px, py, pz, salt);
        const R nx = px * 2.03f + 17.7f;
        const R ny = py * 2.01f - 9.2f;
        const R nz = pz * 2.07f + 5.1f;
        px = nx;
        py = ny;
        pz = nz;
        amp *= 0.5;
    }
    return sat(sum * 1.142857f);
}

// ------------------------------------------------------------- params ----
// Reference clamps src/fire_cuda.cu:3135-3179 with the extensions listed in
// include/firesim.h.
static inline float clampf(float v, float lo, float hi) { return std::max(lo, std::min(hi, v)); }
static inline int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return std::max(lo, std::min(hi, v)); }

static void clamp_params(const fs_params& in, fs_params& o) {
    o = in;
    o.dt = clampf(in.dt, 0.001f, 1.0f / 30.0f);
    o.mouse_x = clampf(in.mouse_x, 0.0f, 1.0f);
    o.mouse_y = clampf(in.mouse_y, 0.0f, 1.0f);
    o.scene_id = clampi(in.scene_id, 0, 3);
    o.wind_x = clampf(in.wind_x, -1.0f, 1.0f);
    o.wind_z = clampf(in.wind_z, -1.0f, 1.0f);
    o.turbulence = clampf(in.turbulence, 0.0f, 1.8f);
    o.detail = clampf(in.detail, 0.1f, 1.5f);
    o.smoke = clampf(in.smoke, 0.0f, 1.6f);
    o.intensity = clampf(in.intensity, 0.2f, 2.4f);
    o.camera_pitch = clampf(in.camera_pitch, -0.75f, 0.75f);
    o.camera_distance = clampf(in.camera_distance, 1.5f, 6.0f);
    o.cinematic_mode = in.cinematic_mode != 0 ? 1 : 0;
    o.raymarch_steps = clampi(in.raymarch_steps, 24, 256);
    o.ember_count = clampi(in.ember_count, 0, 192);
    o.render_debug_mode = clampi(in.render_debug_mode, 0, 6);
    o.exposure = clampf(in.exposure, 0.45f, 2.40f);
    o.reflection_gain = clampf(in.reflection_gain, 0.0f, 2.2f);
    o.smoke_darkness = clampf(in.smoke_darkness, 0.35f, 2.4f);
    o.emitter_center_x = clampf(in.emitter_center_x, -1.05f, 1.05f);
    o.emitter_center_z = clampf(in.emitter_center_z, -0.82f, 0.82f);
    o.emitter_height_norm = clampf(in.emitter_height_norm, 0.0f, 0.96f);
    o.emitter_height_band_norm = clampf(in.emitter_height_band_norm, 0.006f, 0.20f);
    o.emitter_radius = clampf(in.emitter_radius, 0.035f, 0.80f);
    o.burner_center_count = clampi(in.burner_center_count, 0, 4);
    for (int i = 0; i < 4; ++i) {
        o.burner_center_x[i] = clampf(in.burner_center_x[i], -1.05f, 1.05f);
        o.burner_center_y[i] = clampf(in.burner_center_y[i], 0.0f, 2.03f);
        o.burner_center_z[i] = clampf(in.burner_center_z[i], -0.82f, 0.82f);
    }
    if (o.scene_id == 2 && o.burner_center_count > 0) {
        o.emitter_height_norm = clampf((o.burner_center_y[0] - 0.02f) / 2.03f, 0.0f, 0.96f);
    }
    o.sor_omega = clampf(in.sor_omega, 0.05f, 1.95f);
    o.sor_iterations = clampi(in.sor_iterations, 1, 1000);
    // extensions (no reference counterpart)
    o.closed_top = in.closed_top != 0 ? 1 : 0;
    o.noise_forcing = in.noise_forcing != 0 ? 1 : 0;
    o.curl_noise = clampf(in.curl_noise, 0.0f, 4.0f);
    o.noise_mode = clampi(in.noise_mode, 0, 1);
    o.show_gizmos = in.show_gizmos != 0 ? 1 : 0;
    o.active_gizmo = in.active_gizmo;
    o.dir_light = in.dir_light != 0 ? 1 : 0;
    o.dir_light_r = clampf(in.dir_light_r, 0.0f, 20.0f);
    o.dir_light_g = clampf(in.dir_light_g, 0.0f, 20.0f);
    o.dir_light_b = clampf(in.dir_light_b, 0.0f, 20.0f);
}

struct Cam {
    V3 eye, forward, right, up;
    R tanHalfFov, aspect;
};

// Per-pass parameters in FP64 (SimParams src/fire_cuda.cu:16-68).
struct P {
    int nx, ny, nz, lnx, lny, lnz, fw, fh;
    R dt, time;
    R mouseX, mouseY;
    int leftDown, rightDown, sceneId;
    R windX, windZ, turbulence, detail, smokeGain, intensity;
    int cinematic, raymarchSteps, debugMode;
    R exposure, reflectionGain, smokeDarkness;
    R emitterCenterX, emitterCenterZ, emitterHeightNorm, emitterHeightBandNorm, emitterRadius;
    int burnerCount;
    R burnerX[4], burnerY[4], burnerZ[4];
    R omega;
    int iterations;
    uint32_t saltNoise2, saltNoise3, saltJitter, saltDither, saltGrain, saltEmber;
    uint32_t saltCurl[3];
    int emberCount;
    int closedTop;
    int curlOn;
    R curlNoise;
    int showGizmos, activeGizmo;
    int dirLight;
    R dirLightDir[3], dirLightColor[3];
    Cam cam;
    int emitterCount;
    R ex[FS_MAX_EMITTERS], ey[FS_MAX_EMITTERS], ez[FS_MAX_EMITTERS], er[FS_MAX_EMITTERS];
    R efuel[FS_MAX_EMITTERS], eheat[FS_MAX_EMITTERS];
};

// Orbit camera, src/fire_cuda.cu:315-329.
static Cam make_camera(const fs_params& s, int fw, int fh) {
    Cam c;
    const V3 target = v3(0.0, 0.72f, 0.0);
    const R pitch = s.camera_pitch, yaw = s.camera_yaw, dist = s.camera_distance;
    const R cp = std::cos(pitch);
    c.eye = v3(std::sin(yaw) * dist * cp, target.y + std::sin(pitch) * dist, std::cos(yaw) * dist * cp);
    c.forward = normalize3(sub3(target, c.eye));
    c.right = normalize3(cross3(c.forward, v3(0.0, 1.0, 0.0)));
    c.up = normalize3(cross3(c.right, c.forward));
    c.tanHalfFov = 0.50f;
    c.aspect = static_cast<R>(static_cast<float>(fw) / static_cast<float>(fh));
    return c;
}

// ----------------------------------------------------------------- state ----
struct State {
    int nx = 0, ny = 0, nz = 0, lnx = 0, lny = 0, lnz = 0, fw = 0, fh = 0;
    fs_params prm{};
    std::vector<fs_emitter> emitters;
    std::vector<R> s[10];
    std::vector<R> u, v, w, p, d;
    std::vector<R> light;   // 4 per light cell
    std::vector<R> shadow;  // 1 per light cell
    std::vector<R> dirT;    // directional-light transmittance, 1 per light cell
    float time = 0.0f, renderTime = 0.0f;
    int frameIndex = 0;
    R heatReleaseStep = 0.0, heatReleaseTotal = 0.0;

    size_t cells() const { return size_t(nx) * ny * nz; }
    size_t ucount() const { return size_t(nx + 1) * ny * nz; }
    size_t vcount() const { return size_t(nx) * (ny + 1) * nz; }
    size_t wcount() const { return size_t(nx) * ny * (nz + 1); }
    size_t lcells() const { return size_t(lnx) * lny * lnz; }

    inline size_t si(int x, int y, int z) const { return (size_t(z) * ny + y) * nx + x; }
    inline size_t sic(int x, int y, int z) const {
        x = std::min(std::max(x, 0), nx - 1);
        y = std::min(std::max(y, 0), ny - 1);
        z = std::min(std::max(z, 0), nz - 1);
        return si(x, y, z);
    }
    inline size_t ui(int x, int y, int z) const { return (size_t(z) * ny + y) * (nx + 1) + x; }
    inline size_t vi(int x, int y, int z) const { return (size_t(z) * (ny + 1) + y) * nx + x; }
    inline size_t wi(int x, int y, int z) const { return (size_t(z) * ny + y) * nx + x; }
    inline size_t li(int x, int y, int z) const { return (size_t(z) * lny + y) * lnx + x; }
};

enum { HEAT = 0, FUEL, OXYGEN, SOOT, CHARF, ASH, PYRO, PROGRESS, TURB, SOPT };

static P make_p(const State& st, R time) {
    P p;
    const fs_params& s = st.prm;
    p.nx = st.nx; p.ny = st.ny; p.nz = st.nz;
    p.lnx = st.lnx; p.lny = st.lny; p.lnz = st.lnz;
    p.fw = st.fw; p.fh = st.fh;
    p.dt = s.dt;
    p.time = time;
    p.mouseX = s.mouse_x; p.mouseY = s.mouse_y;
    p.leftDown = s.left_down; p.rightDown = s.right_down;
    p.sceneId = s.scene_id;
    p.windX = s.wind_x; p.windZ = s.wind_z;
    p.turbulence = s.turbulence; p.detail = s.detail; p.smokeGain = s.smoke; p.intensity = s.intensity;
    p.cinematic = s.cinematic_mode; p.raymarchSteps = s.raymarch_steps; p.debugMode = s.render_debug_mode;
    p.exposure = s.exposure; p.reflectionGain = s.reflection_gain; p.smokeDarkness = s.smoke_darkness;
    p.emitterCenterX = s.emitter_center_x; p.emitterCenterZ = s.emitter_center_z;
    p.emitterHeightNorm = s.emitter_height_norm; p.emitterHeightBandNorm = s.emitter_height_band_norm;
    p.emitterRadius = s.emitter_radius;
    p.burnerCount = s.burner_center_count;
    for (int i = 0; i < 4; ++i) {
        p.burnerX[i] = s.burner_center_x[i];
        p.burnerY[i] = s.burner_center_y[i];
        p.burnerZ[i] = s.burner_center_z[i];
    }
    p.omega = s.sor_omega;
    p.iterations = s.sor_iterations;
    p.saltNoise2 = salt_of(s.noise_seed, kSaltNoise2);
    p.saltNoise3 = salt_of(s.noise_seed, kSaltNoise3);
    p.saltJitter = salt_of(s.noise_seed, kSaltJitter);
    p.saltDither = salt_of(s.noise_seed, kSaltDither);
    p.saltGrain = salt_of(s.noise_seed, kSaltGrain);
    p.saltEmber = salt_of(s.noise_seed, kSaltEmber);
    for (int c = 0; c < 3; ++c) p.saltCurl[c] = salt_of(s.noise_seed, 7u + c);
    p.closedTop = s.closed_top;
    p.curlOn = s.noise_forcing;
    p.curlNoise = s.curl_noise;
    p.showGizmos = s.show_gizmos;
    p.activeGizmo = s.active_gizmo;
    p.dirLight = s.dir_light;
    {
        // the light's direction in FP32, normalized as the product normalizes it
        const float dx = s.dir_light_x, dy = s.dir_light_y, dz = s.dir_light_z;
        const float len = std::sqrt(dx * dx + dy * dy + dz * dz);
        const bool ok = len > 1e-6f && std::isfinite(len);
        p.dirLightDir[0] = ok ? dx / len : 0.0f;
        p.dirLightDir[1] = ok ? dy / len : -1.0f;
        p.dirLightDir[2] = ok ? dz / len : 0.0f;
    }
    p.dirLightColor[0] = s.dir_light_r;
    p.dirLightColor[1] = s.dir_light_g;
    p.dirLightColor[2] = s.dir_light_b;
    p.emberCount = s.ember_count;
    p.cam = make_camera(s, st.fw, st.fh);
    p.emitterCount = static_cast<int>(std::min<size_t>(st.emitters.size(), FS_MAX_EMITTERS));
    for (int i = 0; i < p.emitterCount; ++i) {
        const fs_emitter& e = st.emitters[i];
        p.ex[i] = e.cx; p.ey[i] = e.cy; p.ez[i] = e.cz; p.er[i] = e.radius;
        p.efuel[i] = e.fuel_rate;
        p.eheat[i] = (static_cast<R>(e.temperature_k) - 293.0f) / 335.0f;
    }
    return p;
}

// ---------------------------------------------------------- samplers ----
// Clamped manual trilinear sample, src/fire_cuda.cu:489-519 (scalar),
// :521-600 (faces). dims are the array sizes, scale the coordinate multipliers
// (n-1 on collocated axes, n on the face axis).
static inline R sample_grid(const R* f, int dx, int dy, int dz, R sx, R sy, R sz, R u, R v, R w) {
    u = sat(u);
    v = sat(v);
    w = sat(w);
    const R fx = u * sx, fy = v * sy, fz = w * sz;
    const int x0 = static_cast<int>(std::floor(fx));
    const int y0 = static_cast<int>(std::floor(fy));
    const int z0 = static_cast<int>(std::floor(fz));
    const int x1 = std::min(x0 + 1, dx - 1);
    const int y1 = std::min(y0 + 1, dy - 1);
    const int z1 = std::min(z0 + 1, dz - 1);
    const R tx = fx - x0, ty = fy - y0, tz = fz - z0;
    auto at = [&](int x, int y, int z) { return f[(size_t(z) * dy + y) * dx + x]; };
    const R c00 = lerp(at(x0, y0, z0), at(x1, y0, z0), tx);
    const R c10 = lerp(at(x0, y1, z0), at(x1, y1, z0), tx);
    const R c01 = lerp(at(x0, y0, z1), at(x1, y0, z1), tx);
    const R c11 = lerp(at(x0, y1, z1), at(x1, y1, z1), tx);
    return lerp(lerp(c00, c10, ty), lerp(c01, c11, ty), tz);
}

static inline R sample_scalar(const State& st, const R* f, R u, R v, R w) {
    return sample_grid(f, st.nx, st.ny, st.nz, st.nx - 1, st.ny - 1, st.nz - 1, u, v, w);
}
static inline R sample_u(const State& st, const R* f, R u, R v, R w) {
    return sample_grid(f, st.nx + 1, st.ny, st.nz, st.nx, st.ny - 1, st.nz - 1, u, v, w);
}
static inline R sample_v(const State& st, const R* f, R u, R v, R w) {
    return sample_grid(f, st.nx, st.ny + 1, st.nz, st.nx - 1, st.ny, st.nz - 1, u, v, w);
}
static inline R sample_w(const State& st, const R* f, R u, R v, R w) {
    return sample_grid(f, st.nx, st.ny, st.nz + 1, st.nx - 1, st.ny - 1, st.nz, u, v, w);
}
// src/fire_cuda.cu:602-604
static inline V3 sample_velocity(const State& st, const R* uf, const R* vf, const R* wf, R u, R v, R w) {
    return v3(sample_u(st, uf, u, v, w), sample_v(st, vf, u, v, w), sample_w(st, wf, u, v, w));
}

// 8-corner min/max of the scalar sample cell, src/fire_cuda.cu:666-711.
static inline void scalar_bounds(const State& st, const R* f, R u, R v, R w, R& lo, R& hi) {
    u = sat(u);
    v = sat(v);
    w = sat(w);
    const R fx = u * (st.nx - 1), fy = v * (st.ny - 1), fz = w * (st.nz - 1);
    const int x0 = static_cast<int>(std::floor(fx));
    const int y0 = static_cast<int>(std::floor(fy));
    const int z0 = static_cast<int>(std::floor(fz));
    const int x1 = std::min(x0 + 1, st.nx - 1);
    const int y1 = std::min(y0 + 1, st.ny - 1);
    const int z1 = std::min(z0 + 1, st.nz - 1);
    const R c[8] = {f[st.si(x0, y0, z0)], f[st.si(x1, y0, z0)], f[st.si(x0, y1, z0)], f[st.si(x1, y1, z0)],
                    f[st.si(x0, y0, z1)], f[st.si(x1, y0, z1)], f[st.si(x0, y1, z1)], f[st.si(x1, y1, z1)]};
    lo = c[0];
    hi = c[0];
    for (int i = 1; i < 8; ++i) {
        lo = std::min(lo, c[i]);
        hi = std::max(hi, c[i]);
    }
}

// MacCormack with 8-corner limiter, src/fire_cuda.cu:713-732.
static inline R maccormack(const State& st, const P& p, const R* f, R current, V3 prevVel, R pu, R pv, R pw, R blend) {
    const R firstOrder = sample_scalar(st, f, pu, pv, pw);
    const R fu = pu + prevVel.x * p.dt * 0.62f;
    const R fv = pv + prevVel.y * p.dt * 0.62f;
    const R fw = pw + prevVel.z * p.dt * 0.62f;
    const R roundTrip = sample_scalar(st, f, fu, fv, fw);
    R lo, hi;
    scalar_bounds(st, f, pu, pv, pw, lo, hi);
    const R corrected = std::min(hi, std::max(lo, firstOrder + 0.5 * (current - roundTrip)));
    return lerp(firstOrder, corrected, sat(blend));
}

// Light-grid trilinear (float4) and nearest shadow, src/fire_cuda.cu:618-664.
static inline V4 sample_light(const State& st, R u, R v, R w) {
    if (st.lnx <= 1 || st.lny <= 1 || st.lnz <= 1) return V4{0, 0, 0, 0};
    V4 o;
    const R* L = st.light.data();
    R* out[4] = {&o.x, &o.y, &o.z, &o.w};
    u = sat(u);
    v = sat(v);
    w = sat(w);
    const R fx = u * (st.lnx - 1), fy = v * (st.lny - 1), fz = w * (st.lnz - 1);
    const int x0 = static_cast<int>(std::floor(fx));
    const int y0 = static_cast<int>(std::floor(fy));
    const int z0 = static_cast<int>(std::floor(fz));
    const int x1 = std::min(x0 + 1, st.lnx - 1);
    const int y1 = std::min(y0 + 1, st.lny - 1);
    const int z1 = std::min(z0 + 1, st.lnz - 1);
    const R tx = fx - x0, ty = fy - y0, tz = fz - z0;
    for (int c = 0; c < 4; ++c) {
        auto at = [&](int x, int y, int z) { return L[4 * st.li(x, y, z) + c]; };
        const R c00 = lerp(at(x0, y0, z0), at(x1, y0, z0), tx);
        const R c10 = lerp(at(x0, y1, z0), at(x1, y1, z0), tx);
        const R c01 = lerp(at(x0, y0, z1), at(x1, y0, z1), tx);
        const R c11 = lerp(at(x0, y1, z1), at(x1, y1, z1), tx);
        *out[c] = lerp(lerp(c00, c10, ty), lerp(c01, c11, ty), tz);
    }
    return o;
}

static inline R sample_shadow(const State& st, R u, R v, R w) {
    if (st.lnx <= 1 || st.lny <= 1 || st.lnz <= 1) return 1.0;
    u = sat(u);
    v = sat(v);
    w = sat(w);
    const int x = std::min(std::max(static_cast<int>(u * (st.lnx - 1) + 0.5f), 0), st.lnx - 1);
    const int y = std::min(std::max(static_cast<int>(v * (st.lny - 1) + 0.5f), 0), st.lny - 1);
    const int z = std::min(std::max(static_cast<int>(w * (st.lnz - 1) + 0.5f), 0), st.lnz - 1);
    return st.shadow[st.li(x, y, z)];
}

// --------------------------------------------------------- fuel beds ----
// src/fire_cuda.cu:734-744
static inline R capsule_mask(R x, R z, R cx, R cz, R angle, R halfLength, R radius) {
    const R ca = std::cos(angle), sa = std::sin(angle);
    const R px = x - cx, pz = z - cz;
    const R qx = px * ca + pz * sa;
    const R qz = -px * sa + pz * ca;
    const R over = std::max<R>(std::fabs(qx) - halfLength, 0.0);
    const R dist = std::sqrt(over * over + qz * qz);
    return smoothstep(radius, radius * 0.22f, dist);
}

// src/fire_cuda.cu:746-793 (scene 3: no bed).
static R bed_material(R x, R z, const P& p) {
    const uint32_t s2 = p.saltNoise2;
    if (p.sceneId == 3) return 0.0;
    if (p.sceneId == 1) {
        const R logA = capsule_mask(x, z, -0.06f, 0.01f, 0.38f, 0.34f, 0.060f);
        const R logB = capsule_mask(x, z, 0.07f, -0.02f, -0.58f, 0.33f, 0.058f);
        const R logC = capsule_mask(x, z, 0.01f, 0.08f, 1.34f, 0.24f, 0.050f);
        const R contactAB = std::min(logA, logB) * 1.25f;
        const R contactBC = std::min(logB, logC) * 1.10f;
        const R contactCA = std::min(logC, logA) * 0.92f;
        const R coalNoise = fbm2(x * 9.4f + z * 1.7f, z * 11.2f - x * 3.1f, s2);
        const R coalBed = smoothstep(0.38f, 0.08f, std::sqrt(x * x * 1.35f + z * z * 2.40f)) * smoothstep(0.34f, 0.88f, coalNoise);
        const R cracks = smoothstep(0.65f, 0.93f, fbm2(x * 21.0f - z * 6.0f, z * 24.0f + x * 4.0f, s2));
        const R contactPyro = sat(contactAB + contactBC + contactCA);
        return sat((logA * 0.72f + logB * 0.70f + logC * 0.54f + coalBed * 0.84f + contactPyro * 1.18f) * (1.0 - cracks * 0.24f));
    }
    if (p.sceneId == 2) {
        const R sourceRadius = std::max<R>(0.04f, p.emitterRadius);
        R source = 0.0;
        const R cx = p.burnerCount > 0 ? p.burnerX[0] : p.emitterCenterX;
        const R cz = p.burnerCount > 0 ? p.burnerZ[0] : p.emitterCenterZ;
        const R px = x - cx, pz = z - cz;
        const R radius = std::sqrt(px * px + pz * pz);
        const R angle = std::atan2(pz, px);
        const R ring = smoothstep(sourceRadius * 0.072f, sourceRadius * 0.014f, std::fabs(radius - sourceRadius));
        const R innerRing = smoothstep(sourceRadius * 0.052f, sourceRadius * 0.012f, std::fabs(radius - sourceRadius * 0.52f));
        const R ports = 0.5f + 0.5f * std::cos(angle * 32.0f);
        const R portMask = smoothstep(0.80f, 0.992f, ports);
        const R centerPilot = smoothstep(sourceRadius * 0.22f, 0.000f, radius) * 0.045f;
        source = std::max(source, (ring * (0.035f + portMask * 1.72f) + innerRing * 0.055f + centerPilot) *
                                      smoothstep(sourceRadius * 1.28f, sourceRadius * 1.05f, radius));
        return sat(source);
    }
    const R tray = smoothstep(1.08f, 0.030f, std::fabs(x)) * smoothstep(0.56f, 0.026f, std::fabs(z));
    const R emberMat = smoothstep(1.00f, 0.018f, std::fabs(x)) * smoothstep(0.50f, 0.020f, std::fabs(z));
    const R mound = std::exp(-(x * x * 0.70f + z * z * 2.15f));
    const R logMass = fbm2(x * 3.7f + z * 0.9f, z * 5.4f - x * 0.6f, s2);
    const R chips = fbm2(x * 15.0f + z * 4.0f, z * 18.0f - x * 5.0f, s2);
    const R strand = 0.5f + 0.5f * std::sin((x * 19.0f + z * 7.0f) + fbm2(z * 4.0f, x * 3.0f, s2) * 4.2f);
    const R cracks = smoothstep(0.44f, 0.74f, chips) * smoothstep(0.20f, 0.84f, logMass);
    const R voids = smoothstep(0.70f, 0.93f, chips * 0.72f + strand * 0.34f);
    const R clumps = smoothstep(0.22f, 0.86f, logMass * 0.82f + chips * 0.36f + strand * 0.18f);
    const R porousFuel = (0.24f + clumps * 0.94f + emberMat * 0.20f) * (0.64f + mound * 0.46f);
    return sat(tray * porousFuel * (1.0 - cracks * 0.46f) * (1.0 - voids * 0.34f));
}

// src/fire_cuda.cu:795-803
static R bed_source(R x, R z, R h, const P& p) {
    const R material = bed_material(x, z, p);
    if (material == 0.0) return 0.0;
    const R height = p.sceneId == 2
        ? smoothstep(std::max<R>(0.005f, p.emitterHeightBandNorm * 0.34f), std::max<R>(0.0015f, p.emitterHeightBandNorm * 0.055f),
                     std::fabs(h - p.emitterHeightNorm))
        : smoothstep(p.sceneId == 1 ? 0.245f : 0.145f, 0.00f, h);
    const R breathing = 0.84f + 0.16f * fbm2(x * 9.0f + p.time * 0.035f, z * 11.0f - p.time * 0.025f, p.saltNoise2);
    const R exposed = sat(material * breathing);
    return sat(height * exposed);
}

// --------------------------------------------------------------- reset ----
// resetScalarsKernel / resetU/V/WKernel, src/fire_cuda.cu:805-878 and the
// host reset src/fire_cuda.cu:3006-3060.
static void reset(State& st) {
    const P p = make_p(st, 0.0);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < st.nz; ++z) {
        for (int y = 0; y < st.ny; ++y) {
            for (int x = 0; x < st.nx; ++x) {
                const size_t idx = st.si(x, y, z);
                const R u = (x + 0.5) / st.nx, v = (y + 0.5) / st.ny, w = (z + 0.5) / st.nz;
                const R worldX = (u - 0.50f) * 2.10f;
                const R worldZ = (w - 0.50f) * 1.64f;
                const R material = bed_material(worldX, worldZ, p);
                R heat = 0, fuel = 0, soot = 0, ch = 0, ash = 0, pyro = 0, prog = 0, k = 0, so = 0;
                if (material != 0.0) {
                    const int sid = p.sceneId;
                    const R chunk = fbm2(worldX * 18.0f + worldZ * 2.0f, worldZ * 22.0f - worldX * 5.0f, p.saltNoise2);
                    const R bed = bed_source(worldX, worldZ, v, p);
                    const R charScale = sid == 2 ? 0.004f : (sid == 1 ? 1.18f : 1.0f);
                    const R sootScale = sid == 2 ? 0.004f : (sid == 1 ? 0.92f : 1.0f);
                    heat = bed * (sid == 2 ? 1.34f : (sid == 1 ? 0.78f + chunk * 0.26f : 0.86f + chunk * 0.24f));
                    fuel = bed * (sid == 2 ? 1.42f : (sid == 1 ? 1.05f + chunk * 0.30f : 1.08f + chunk * 0.26f));
                    soot = bed * (sid == 2 ? 0.00065f : 0.018f) * sootScale;
                    ch = material * (0.92f + chunk * 0.72f) * charScale;
                    ash = material * (0.018f + smoothstep(0.56f, 0.86f, chunk) * 0.055f) * (sid == 2 ? 0.10f : 1.0f);
                    pyro = bed * (sid == 2 ? 0.018f : (sid == 1 ? 0.060f + chunk * 0.034f : 0.058f + chunk * 0.028f));
                    prog = bed * (sid == 2 ? 0.11f : (sid == 1 ? 0.24f : 0.21f));
                    k = bed * (sid == 2 ? 0.012f : (sid == 1 ? 0.082f : 0.065f));
                    so = bed * (sid == 2 ? 0.00050f : 0.018f) * sootScale;
                }
                st.s[HEAT][idx] = heat;
                st.s[FUEL][idx] = fuel;
                st.s[OXYGEN][idx] = 1.0;
                st.s[SOOT][idx] = soot;
                st.s[CHARF][idx] = ch;
                st.s[ASH][idx] = ash;
                st.s[PYRO][idx] = pyro;
                st.s[PROGRESS][idx] = prog;
                st.s[TURB][idx] = k;
                st.s[SOPT][idx] = so;
                st.p[idx] = 0.0;
                st.d[idx] = 0.0;
            }
        }
    }
    std::fill(st.u.begin(), st.u.end(), 0.0);
    std::fill(st.v.begin(), st.v.end(), 0.0);
    std::fill(st.w.begin(), st.w.end(), 0.0);
    std::fill(st.light.begin(), st.light.end(), 0.0);
    std::fill(st.shadow.begin(), st.shadow.end(), 0.0);
    std::fill(st.dirT.begin(), st.dirT.end(), 0.0);
    st.time = 0.0f;
    st.renderTime = 0.0f;
    st.frameIndex = 0;
    st.heatReleaseStep = 0.0;
    st.heatReleaseTotal = 0.0;
}

// -------------------------------------------------- velocity advection ----
// advectU/V/WKernel, src/fire_cuda.cu:880-932.
static void advect_velocity(State& st, const P& p) {
    std::vector<R> un(st.u.size()), vn(st.v.size()), wn(st.w.size());
    const R* U = st.u.data();
    const R* Vv = st.v.data();
    const R* W = st.w.data();
    const R k = 0.62f;
#pragma omp parallel for schedule(static)
    for (int z = 0; z <= st.nz; ++z) {
        for (int y = 0; y <= st.ny; ++y) {
            for (int x = 0; x <= st.nx; ++x) {
                if (y < st.ny && z < st.nz) {  // u face
                    R out = 0.0;
                    if (x != 0 && x != st.nx) {
                        const R fu = static_cast<R>(x) / st.nx, fv = (y + 0.5) / st.ny, fw = (z + 0.5) / st.nz;
                        const V3 vel = sample_velocity(st, U, Vv, W, fu, fv, fw);
                        out = sample_u(st, U, fu - vel.x * p.dt * k, fv - vel.y * p.dt * k, fw - vel.z * p.dt * k) * 0.995f;
                    }
                    un[st.ui(x, y, z)] = out;
                }
                if (x < st.nx && z < st.nz) {  // v face
                    R out = 0.0;
                    if (y != 0 && !(p.closedTop && y == st.ny)) {
                        const R fu = (x + 0.5) / st.nx, fv = static_cast<R>(y) / st.ny, fw = (z + 0.5) / st.nz;
                        const V3 vel = sample_velocity(st, U, Vv, W, fu, fv, fw);
                        out = sample_v(st, Vv, fu - vel.x * p.dt * k, fv - vel.y * p.dt * k, fw - vel.z * p.dt * k) * 0.996f;
                    }
                    vn[st.vi(x, y, z)] = out;
                }
                if (x < st.nx && y < st.ny) {  // w face
                    R out = 0.0;
                    if (z != 0 && z != st.nz) {
                        const R fu = (x + 0.5) / st.nx, fv = (y + 0.5) / st.ny, fw = static_cast<R>(z) / st.nz;
                        const V3 vel = sample_velocity(st, U, Vv, W, fu, fv, fw);
                        out = sample_w(st, W, fu - vel.x * p.dt * k, fv - vel.y * p.dt * k, fw - vel.z * p.dt * k) * 0.995f;
                    }
                    wn[st.wi(x, y, z)] = out;
                }
            }
        }
    }
    st.u.swap(un);
    st.v.swap(vn);
    st.w.swap(wn);
}

// ------------------------------------------------------ advect + react ----
// advectReactKernel, src/fire_cuda.cu:934-1126, plus fs_emitter sources.
static void advect_react(State& st, const P& p) {
    std::vector<R> out[10];
    for (int f = 0; f < 10; ++f) out[f].resize(st.cells());
    const R* U = st.u.data();
    const R* Vv = st.v.data();
    const R* W = st.w.data();
    const R dt = p.dt;
    const int sid = p.sceneId;
    R heatRelease = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : heatRelease)
    for (int z = 0; z < st.nz; ++z) {
        for (int y = 0; y < st.ny; ++y) {
            for (int x = 0; x < st.nx; ++x) {
                const size_t idx = st.si(x, y, z);
                const R u = (x + 0.5) / st.nx, v = (y + 0.5) / st.ny, w = (z + 0.5) / st.nz;
                const R* in[10];
                for (int f = 0; f < 10; ++f) in[f] = st.s[f].data();
                const R activity = in[HEAT][idx] + in[FUEL][idx] * 0.18f + in[SOOT][idx] * 0.24f + in[PYRO][idx] * 0.20f +
                                   in[PROGRESS][idx] * 0.14f + in[SOPT][idx] * 0.08f + in[TURB][idx] * 0.06f;
                const R sdx = u - p.mouseX, sdy = v - p.mouseY, sdz = w - 0.5f;
                const bool outsideInteractive = p.leftDown == 0 || (sdx * sdx + sdy * sdy + sdz * sdz) > 0.090f;
                bool insideEmitter[FS_MAX_EMITTERS];
                bool anyEmitter = false;
                for (int e = 0; e < p.emitterCount; ++e) {
                    const R ddx = (x + 0.5) - p.ex[e], ddy = (y + 0.5) - p.ey[e], ddz = (z + 0.5) - p.ez[e];
                    insideEmitter[e] = ddx * ddx + ddy * ddy + ddz * ddz <= p.er[e] * p.er[e];
                    anyEmitter = anyEmitter || insideEmitter[e];
                }
                if (v > 0.18f && activity < 0.00035f && in[CHARF][idx] < 0.00010f && in[ASH][idx] < 0.00010f &&
                    outsideInteractive && !anyEmitter && p.rightDown == 0) {
                    const R o = in[OXYGEN][idx];
                    out[HEAT][idx] = 0.0;
                    out[FUEL][idx] = 0.0;
                    out[OXYGEN][idx] = std::min<R>(1.0, o + (1.0 - o) * dt * 0.08f);
                    out[SOOT][idx] = 0.0;
                    out[CHARF][idx] = in[CHARF][idx];
                    out[ASH][idx] = in[ASH][idx];
                    out[PYRO][idx] = 0.0;
                    out[PROGRESS][idx] = 0.0;
                    out[TURB][idx] = in[TURB][idx] * std::max<R>(0.0, 1.0 - dt * 0.54f);
                    out[SOPT][idx] = 0.0;
                    continue;
                }
                const V3 vel = sample_velocity(st, U, Vv, W, u, v, w);
                const R pu = u - vel.x * dt * 0.62f, pv = v - vel.y * dt * 0.62f, pw = w - vel.z * dt * 0.62f;
                const V3 pvel = sample_velocity(st, U, Vv, W, pu, pv, pw);
                const R anti = sat(0.30f + p.detail * 0.34f);
                R heat = maccormack(st, p, in[HEAT], in[HEAT][idx], pvel, pu, pv, pw, anti);
                R fuel = maccormack(st, p, in[FUEL], in[FUEL][idx], pvel, pu, pv, pw, anti * 0.92f);
                R oxygen = maccormack(st, p, in[OXYGEN], in[OXYGEN][idx], pvel, pu, pv, pw, anti * 0.70f);
                R soot = maccormack(st, p, in[SOOT], in[SOOT][idx], pvel, pu, pv, pw, anti);
                R charMass = in[CHARF][idx];
                R ash = in[ASH][idx];
                R pyro = maccormack(st, p, in[PYRO], in[PYRO][idx], pvel, pu, pv, pw, anti * 0.72f) * 0.55f;
                R progress = maccormack(st, p, in[PROGRESS], in[PROGRESS][idx], pvel, pu, pv, pw, anti * 0.82f);
                R turb = maccormack(st, p, in[TURB], in[TURB][idx], pvel, pu, pv, pw, anti * 0.52f);
                R sootOptics = maccormack(st, p, in[SOPT], in[SOPT][idx], pvel, pu, pv, pw, anti * 0.72f);

                const R worldX = (u - 0.50f) * 2.10f;
                const R worldZ = (w - 0.50f) * 1.64f;
                const R material = bed_material(worldX, worldZ, p);
                const R sourceNoise = 0.70f + 0.48f * fbm2(worldX * 13.0f + worldZ * 3.0f + p.time * 0.38f, v * 31.0f + worldZ * 11.0f, p.saltNoise2);
                const R exposedChar = sat(charMass / std::max<R>(0.055f, charMass + ash * 1.75f));
                const R charEdges = smoothstep(0.06f, 0.78f, material) * smoothstep(0.92f, 0.02f, ash);
                const R gasFeed = sid == 2 ? 1.18f : exposedChar * charEdges;
                const R pyroGain = sid == 2 ? 0.42f : (sid == 1 ? 1.34f : 1.0f);
                const R fuelGain = sid == 2 ? 0.46f : (sid == 1 ? 1.18f : 1.0f);
                const R heatGain = sid == 2 ? 1.18f : (sid == 1 ? 1.08f : 1.0f);
                const R sootGain = sid == 2 ? 0.18f : (sid == 1 ? 1.22f : 1.0f);
                const R bed = bed_source(worldX, worldZ, v, p) * sourceNoise * gasFeed * p.intensity;
                const R ambientK = 293.0f;
                R tempK = ambientK + heat * 335.0f;
                const R radFb = smoothstep(430.0f, 980.0f, tempK);
                const R solidO2 = sat(oxygen * 1.28f - soot * 0.030f);
                const R heatFlux = sat((tempK - 405.0f) / 900.0f) + bed * 0.34f + sootOptics * 0.024f;
                const R charRelease = std::min(charMass, charMass * heatFlux * solidO2 * (0.14f + radFb * 1.38f) * dt);
                pyro += (bed * (0.34f + radFb * 1.05f) * dt + charRelease * (2.45f + sourceNoise * 0.72f)) * pyroGain;
                charMass += material * dt * (sid == 2 ? 0.00005f : (sid == 1 ? 0.050f : 0.026f));
                charMass -= charRelease * 0.92f;
                ash += charRelease * (sid == 2 ? static_cast<R>(0.025f) : 0.16f + 0.22f * sat(1.0 - oxygen));
                heat += (bed * (sid == 2 ? 10.40f : 6.20f) * dt * (1.0 + radFb * 1.12f) + charRelease * 3.35f) * heatGain;
                fuel += pyro * (sid == 2 ? 1.34f : 3.70f) * fuelGain;
                soot += pyro * (0.090f + ash * 0.032f + sat(1.0 - oxygen) * 0.13f) * p.smokeGain *
                        (sid == 2 ? 0.010f : (sid == 1 ? 0.92f : 1.0f)) * sootGain;

                if (p.leftDown != 0) {
                    const R mx = p.mouseX, my = p.mouseY;
                    const R source = std::exp(-((u - mx) * (u - mx)) / 0.0180f - ((w - 0.5f) * (w - 0.5f)) / 0.0550f -
                                              ((v - my) * (v - my)) / 0.0020f);
                    heat += source * dt * 2.8f;
                    fuel += source * dt * 1.4f;
                    pyro += source * dt * 0.45f;
                    progress += source * dt * 0.75f;
                    turb += source * dt * 0.28f;
                    oxygen = std::max<R>(oxygen, 0.72f + source * 0.20f);
                }
                if (p.rightDown != 0) {
                    const R source = std::exp(-((u - p.mouseX) * (u - p.mouseX) + (w - 0.5f) * (w - 0.5f)) / 0.0100f -
                                              (v - p.mouseY) * (v - p.mouseY) / 0.0100f);
                    soot += source * dt * 8.0f;
                    heat *= 1.0 - source * dt * 1.2f;
                    fuel *= 1.0 - source * dt * 0.7f;
                }
                for (int e = 0; e < p.emitterCount; ++e) {
                    if (insideEmitter[e]) {
                        fuel += p.efuel[e] * dt;
                        heat = std::max(heat, p.eheat[e]);
                    }
                }

                tempK = ambientK + heat * 335.0f;
                const R thermalActivation = smoothstep(575.0f, 1220.0f, tempK);
                const R arrhenius = std::exp(-3800.0f / std::max<R>(640.0f, tempK));
                const R oxyLim = std::min(fuel, oxygen * 1.42f);
                const R burnMass = std::min(oxyLim, oxyLim * thermalActivation * arrhenius * (118.0f + p.intensity * 34.0f) * dt);
                const R incomplete = sat(1.0 - oxygen * 1.30f + fuel * 0.065f + soot * 0.015f);
                const R release = burnMass * 7.0f;
                const R sootYield = burnMass * (0.035f + incomplete * 0.34f) * p.smokeGain * (sid == 2 ? 0.006f : (sid == 1 ? 0.96f : 1.0f));
                const R sootOxidation = soot * oxygen * smoothstep(760.0f, 1540.0f, tempK) * dt * 0.82f;
                const R frontProduction = burnMass * (9.8f + turb * 3.4f) + pyro * 0.70f + charRelease * 2.10f;
                const R boundaryAir = smoothstep(0.70f, 0.96f, v) + smoothstep(0.11f, 0.03f, u) + smoothstep(0.89f, 0.97f, u) +
                                      smoothstep(0.11f, 0.03f, w) + smoothstep(0.89f, 0.97f, w);
                const R upper = smoothstep(0.52f, 0.98f, v);
                const R bA = sat(boundaryAir);

                heatRelease += release;
                heat += release;
                fuel -= burnMass * 1.20f;
                oxygen -= burnMass * 0.98f;
                soot += sootYield;
                soot -= sootOxidation;
                progress += frontProduction;
                progress *= std::exp(-dt * (0.72f + v * 0.45f + oxygen * 0.16f + bA * 1.05f + upper * 0.26f));
                oxygen = lerp(oxygen, 1.0, bA * dt * 1.7f);
                oxygen += (1.0 - oxygen) * dt * 0.08f;

                tempK = ambientK + heat * 335.0f;
                const R radBase = std::max<R>(0.0, (tempK - ambientK) / 1560.0f);
                const R radSq = radBase * radBase;
                const R radLoss = radSq * radSq;
                heat -= radLoss * dt * (0.88f + soot * 0.085f + sootOptics * 0.018f);
                heat *= std::exp(-dt * (0.31f + v * 0.64f + soot * 0.022f + bA * 0.26f + upper * 0.18f));
                fuel *= std::exp(-dt * (0.58f + heat * 0.20f + v * 0.92f + bA * 0.18f));
                soot *= std::exp(-dt * (0.060f + v * 0.092f + oxygen * 0.028f + bA * 0.96f + upper * 0.22f));
                pyro *= std::exp(-dt * (1.55f + v * 0.58f + bA * 0.18f));

                const R uL = U[st.ui(x, y, z)], uR = U[st.ui(x + 1, y, z)];
                const R vD = Vv[st.vi(x, y, z)], vU = Vv[st.vi(x, y + 1, z)];
                const R wB = W[st.wi(x, y, z)], wF = W[st.wi(x, y, z + 1)];
                const R shear = std::fabs(uR - uL) + std::fabs(vU - vD) + std::fabs(wF - wB);
                const R flameActivity = sat(progress * 0.64f + burnMass * 5.2f + heat * 0.065f);
                turb += dt * (shear * (0.72f + progress * 0.06f) + flameActivity * p.turbulence * 0.62f + frontProduction * 0.085f + soot * 0.018f);
                turb *= std::exp(-dt * (0.54f + v * 0.24f + bA * 0.42f + upper * 0.16f));
                const R particleGrowth = sat(incomplete * 0.36f + ash * 0.13f + soot * 0.055f + smoothstep(0.65f, 1.55f, tempK / 1000.0f) * 0.18f);
                sootOptics = soot * (1.04f + 0.28f * std::sqrt(std::max<R>(0.0, heat)) + 0.24f * ash + 0.46f * particleGrowth);
                sootOptics = lerp(sootOptics, sootOptics * 0.48f, sat(sootOxidation * 0.92f));

                out[HEAT][idx] = std::min<R>(8.8f, std::max<R>(0.0, heat));
                out[FUEL][idx] = std::min<R>(5.5f, std::max<R>(0.0, fuel));
                out[OXYGEN][idx] = std::min<R>(1.0, std::max<R>(0.0, oxygen));
                out[SOOT][idx] = std::min<R>(7.4f, std::max<R>(0.0, soot));
                out[CHARF][idx] = std::min<R>(2.8f, std::max<R>(0.0, charMass));
                out[ASH][idx] = std::min<R>(1.9f, std::max<R>(0.0, ash));
                out[PYRO][idx] = std::min<R>(3.6f, std::max<R>(0.0, pyro));
                out[PROGRESS][idx] = std::min<R>(3.4f, std::max<R>(0.0, progress));
                out[TURB][idx] = std::min<R>(4.2f, std::max<R>(0.0, turb));
                out[SOPT][idx] = std::min<R>(10.0f, std::max<R>(0.0, sootOptics));
            }
        }
    }
    for (int f = 0; f < 10; ++f) st.s[f].swap(out[f]);
    st.heatReleaseStep = heatRelease;
    st.heatReleaseTotal += heatRelease;
}

// ----------------------------------------------------- curl-noise forcing ----
// Extension replacing the fbm noise terms (:1164-1165) when noise_forcing is 1: a
// vector potential on the grid edges of index space (psi_x at (i + 1/2, j, k),
// psi_y at (i, j + 1/2, k), psi_z at (i, j, k + 1/2)), one octave of value noise
// of the edge's normalized position drifting with time, scaled by
// (nx + ny + nz) / (3 * 6); its tangential components are zero on the walls
// (and the top when closed). Each face receives dt * amplitude times the
// index-space curl, whose discrete divergence cancels term by term.
static R curl_edge_potential(const State& st, const P& p, int c, int i, int j, int k) {
    const bool topWall = p.closedTop != 0;
    // walls: x = 0 and nx, z = 0 and nz, the floor, and the top if closed
    if (c == 0 && (j == 0 || (topWall && j == st.ny) || k == 0 || k == st.nz || i >= st.nx)) return 0.0;
    if (c == 1 && (i == 0 || i == st.nx || k == 0 || k == st.nz || j >= st.ny)) return 0.0;
    if (c == 2 && (i == 0 || i == st.nx || j == 0 || (topWall && j == st.ny) || k >= st.nz)) return 0.0;
    const R ex = (c == 0 ? i + 0.5 : static_cast<R>(i)) / st.nx;
    const R ey = (c == 1 ? j + 0.5 : static_cast<R>(j)) / st.ny;
    const R ez = (c == 2 ? k + 0.5 : static_cast<R>(k)) / st.nz;
    const R F = 6.0f;
    const R scale = static_cast<float>(st.nx + st.ny + st.nz) / (3.0f * 6.0f);
    const R t = p.time;
    return scale * (value_noise3(ex * F + t * 0.23f, ey * F - t * 0.17f, ez * F + t * 0.19f, p.saltCurl[c]) - 0.5);
}

// ------------------------------------------- force + boundary + divergence ----
// forceVelocityKernel src/fire_cuda.cu:1128-1183 in gather form, then
// velocityBoundaryKernel :1261-1291, then divergenceKernel :1185-1199.
static void force_boundary_divergence(State& st, const P& p) {
    const size_t n = st.cells();
    std::vector<R> cu(n), cv(n), cw(n);
    const R* H = st.s[HEAT].data();
    const R* S = st.s[SOOT].data();
    const R* K = st.s[TURB].data();
    const R* C = st.s[PROGRESS].data();
    const R* U = st.u.data();
    const R* Vv = st.v.data();
    const R* W = st.w.data();
    const int sid = p.sceneId;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < st.nz; ++z) {
        for (int y = 0; y < st.ny; ++y) {
            for (int x = 0; x < st.nx; ++x) {
                const size_t idx = st.si(x, y, z);
                const R h = H[idx], s = S[idx], k = K[idx], c = C[idx];
                const R u = (x + 0.5) / st.nx, v = (y + 0.5) / st.ny, w = (z + 0.5) / st.nz;
                const R gradX = H[st.sic(x + 1, y, z)] - H[st.sic(x - 1, y, z)];
                const R gradY = H[st.sic(x, y + 1, z)] - H[st.sic(x, y - 1, z)];
                const R gradZ = H[st.sic(x, y, z + 1)] - H[st.sic(x, y, z - 1)];
                const R uL = U[st.ui(x, y, z)], uR = U[st.ui(x + 1, y, z)];
                const R vD = Vv[st.vi(x, y, z)], vU = Vv[st.vi(x, y + 1, z)];
                const R wB = W[st.wi(x, y, z)], wF = W[st.wi(x, y, z + 1)];
                const R curlX = (wF - wB) - (vU - vD);
                const R curlY = (uR - uL) - (wF - wB);
                const R curlZ = (vU - vD) - (uR - uL);
                const R curlMag = std::fabs(curlX) + std::fabs(curlY) + std::fabs(curlZ);
                // with curl noise on, the fbm terms are replaced by the potential's curl on the faces
                const bool curl = p.curlOn != 0;
                const R noiseA = curl ? 0.0 : fbm2(u * 9.0f + w * 3.7f + p.time * 0.22f, v * 13.0f - p.time * 0.31f, p.saltNoise2) - 0.5f;
                const R noiseB = curl ? 0.0 : fbm2(w * 10.0f - p.time * 0.19f, u * 8.0f + v * 5.5f + p.time * 0.27f, p.saltNoise2) - 0.5f;
                const R thermal = smoothstep(0.04f, 4.20f, h + s * 0.28f);
                const R upperDrag = smoothstep(0.26f, 0.94f, v);
                const R lesGain = std::sqrt(std::max<R>(0.0, k)) * (0.30f + c * 0.54f);
                const R curlScale = sid == 2 ? 0.38f : (sid == 1 ? 1.18f : 1.0f);
                const R buoyScale = sid == 2 ? 0.42f : (sid == 1 ? 1.06f : 1.0f);
                const R curlGain = p.turbulence * curlScale * (0.18f + thermal * 0.82f + lesGain * 1.12f);
                const R vortexGain = p.turbulence * curlScale * thermal * sat(curlMag * 0.46f) * 0.68f;
                const R buoyFade = sid == 2 ? smoothstep(0.24f, 0.035f, v) : smoothstep(0.96f, 0.08f, v);
                const R vertVel = 0.5f * (vD + vU);
                const R vertDamp = vertVel * (0.10f + upperDrag * 0.72f + s * 0.020f);
                const R buoyancy = h * buoyScale * (1.62f + p.turbulence * 0.34f + c * 0.16f) * buoyFade - s * 0.095f - gradY * 0.18f - vertDamp;
                const R windProfile = 0.34f + v * 0.96f;
                cv[idx] = p.dt * (buoyancy + curlY * vortexGain * 0.22f);
                cu[idx] = p.dt * (p.windX * windProfile + (-gradZ * 1.36f + noiseA * (0.20f + upperDrag * 0.16f) + curlX * vortexGain) * curlGain);
                cw[idx] = p.dt * (p.windZ * windProfile + (gradX * 1.36f + noiseB * (0.20f + upperDrag * 0.16f) + curlZ * vortexGain) * curlGain);
            }
        }
    }
    // curl-noise potential on the grid edges (zero on the walls)
    const bool curl = p.curlOn != 0;
    auto psi = [&](int c, int i, int j, int k) -> R { return curl_edge_potential(st, p, c, i, j, k); };
    const R kc = p.dt * p.curlNoise;
    // gather onto faces, then the boundary clamp (K10)
#pragma omp parallel for schedule(static)
    for (int z = 0; z <= st.nz; ++z) {
        for (int y = 0; y <= st.ny; ++y) {
            for (int x = 0; x <= st.nx; ++x) {
                if (y < st.ny && z < st.nz) {
                    const size_t f = st.ui(x, y, z);
                    R val = st.u[f];
                    if (x - 1 >= 0) val += cu[st.si(x - 1, y, z)];
                    if (x < st.nx) val += cu[st.si(x, y, z)];
                    if (curl) val += kc * ((psi(2, x, y + 1, z) - psi(2, x, y, z)) - (psi(1, x, y, z + 1) - psi(1, x, y, z)));
                    st.u[f] = (x == 0 || x == st.nx) ? 0.0 : std::min<R>(4.0f, std::max<R>(-4.0f, val));
                }
                if (x < st.nx && z < st.nz) {
                    const size_t f = st.vi(x, y, z);
                    R val = st.v[f];
                    if (y - 1 >= 0) val += cv[st.si(x, y - 1, z)];
                    if (curl) val += kc * ((psi(0, x, y, z + 1) - psi(0, x, y, z)) - (psi(2, x + 1, y, z) - psi(2, x, y, z)));
                    const bool wall = y == 0 || (p.closedTop && y == st.ny);
                    st.v[f] = wall ? 0.0 : std::min<R>(6.4f, std::max<R>(-0.9f, val));
                }
                if (x < st.nx && y < st.ny) {
                    const size_t f = st.wi(x, y, z);
                    R val = st.w[f];
                    if (z - 1 >= 0) val += cw[st.si(x, y, z - 1)];
                    if (z < st.nz) val += cw[st.si(x, y, z)];
                    if (curl) val += kc * ((psi(1, x + 1, y, z) - psi(1, x, y, z)) - (psi(0, x, y + 1, z) - psi(0, x, y, z)));
                    st.w[f] = (z == 0 || z == st.nz) ? 0.0 : std::min<R>(4.0f, std::max<R>(-4.0f, val));
                }
            }
        }
    }
    // divergence (K11): d = -div, p = 0
#pragma omp parallel for schedule(static)
    for (int z = 0; z < st.nz; ++z) {
        for (int y = 0; y < st.ny; ++y) {
            for (int x = 0; x < st.nx; ++x) {
                const R div = st.u[st.ui(x + 1, y, z)] - st.u[st.ui(x, y, z)] + st.v[st.vi(x, y + 1, z)] - st.v[st.vi(x, y, z)] +
                              st.w[st.wi(x, y, z + 1)] - st.w[st.wi(x, y, z)];
                const size_t idx = st.si(x, y, z);
                st.d[idx] = -div;
                st.p[idx] = 0.0;
            }
        }
    }
}

// ------------------------------------------------------------- pressure ----
// Red-black SOR, sorPressureParityKernel src/fire_cuda.cu:1201-1227 and the
// host loop :3280-3284: parity 0 then 1, in place, Neumann walls by index
// clamp (a missing neighbour reads the cell itself).
static void sor_half(State& st, int parity, R omega) {
    R* Pp = st.p.data();
    const R* D = st.d.data();
#pragma omp parallel for schedule(static)
    for (int z = 0; z < st.nz; ++z) {
        for (int y = 0; y < st.ny; ++y) {
            const int first = (parity ^ y ^ z) & 1;
            for (int x = first; x < st.nx; x += 2) {
                const size_t idx = st.si(x, y, z);
                const R pL = Pp[st.sic(x - 1, y, z)], pR = Pp[st.sic(x + 1, y, z)];
                const R pD = Pp[st.sic(x, y - 1, z)], pU = Pp[st.sic(x, y + 1, z)];
                const R pB = Pp[st.sic(x, y, z - 1)], pF = Pp[st.sic(x, y, z + 1)];
                const R projected = (D[idx] + pL + pR + pD + pU + pB + pF) * (1.0f / 6.0f);
                Pp[idx] = Pp[idx] + omega * (projected - Pp[idx]);
            }
        }
    }
}

static void pressure(State& st, const P& p) {
    for (int i = 0; i < p.iterations; ++i) {
        sor_half(st, 0, p.omega);
        sor_half(st, 1, p.omega);
    }
}

// subtractPressureKernel :1229-1245 then velocityWallBoundaryKernel :1293-1317.
static void project(State& st) {
    const R* Pp = st.p.data();
#pragma omp parallel for schedule(static)
    for (int z = 0; z < st.nz; ++z) {
        for (int y = 0; y < st.ny; ++y) {
            for (int x = 0; x < st.nx; ++x) {
                const R pc = Pp[st.si(x, y, z)];
                if (x > 0) st.u[st.ui(x, y, z)] -= pc - Pp[st.si(x - 1, y, z)];
                if (y > 0) st.v[st.vi(x, y, z)] -= pc - Pp[st.si(x, y - 1, z)];
                if (z > 0) st.w[st.wi(x, y, z)] -= pc - Pp[st.si(x, y, z - 1)];
            }
        }
    }
    for (int z = 0; z < st.nz; ++z)
        for (int y = 0; y < st.ny; ++y) {
            st.u[st.ui(0, y, z)] = 0.0;
            st.u[st.ui(st.nx, y, z)] = 0.0;
        }
    for (int z = 0; z < st.nz; ++z)
        for (int x = 0; x < st.nx; ++x) st.v[st.vi(x, 0, z)] = 0.0;
    for (int y = 0; y < st.ny; ++y)
        for (int x = 0; x < st.nx; ++x) {
            st.w[st.wi(x, y, 0)] = 0.0;
            st.w[st.wi(x, y, st.nz)] = 0.0;
        }
}

// ------------------------------------------------------------------ light ----
// src/fire_cuda.cu:1420-1443
static V3 blackbody(R tempK) {
    const R t = std::max<R>(1000.0f, std::min<R>(12000.0f, tempK)) / 100.0f;
    R r, g, b;
    if (t <= 66.0f) {
        r = 1.0;
        g = sat((99.4708025861f * std::log(t) - 161.1195681661f) / 255.0f);
    } else {
        r = sat(329.698727446f * std::pow(t - 60.0f, static_cast<R>(-0.1332047592f)) / 255.0f);
        g = sat(288.1221695283f * std::pow(t - 60.0f, static_cast<R>(-0.0755148492f)) / 255.0f);
    }
    if (t >= 66.0f) {
        b = 1.0;
    } else if (t <= 19.0f) {
        b = 0.0;
    } else {
        b = sat((138.5177312231f * std::log(t - 10.0f) - 305.0447927307f) / 255.0f);
    }
    return v3(r, g, b);
}

// src/fire_cuda.cu:359-368
static inline R domain_fade(R u, R v, R w) {
    const R side = smoothstep(0.000f, 0.095f, u) * smoothstep(1.000f, 0.905f, u) * smoothstep(0.000f, 0.100f, w) *
                   smoothstep(1.000f, 0.900f, w);
    const R floorFade = smoothstep(0.000f, 0.018f, v);
    const R topFade = smoothstep(1.000f, 0.520f, v);
    return sat(side * floorFade * topFade);
}

// buildSceneLightKernel :1445-1525 then propagateSceneLightKernel ×2 :1527-1565
// with (0.48, 0.92) then (0.34, 0.96) (:3313-3314).
// buildSceneLightKernel :1445-1525: radiance + extinction into L, shadow into st.shadow.
static void light_build(State& st, const P& p, std::vector<R>& L) {
    const R* F[10];
    for (int f = 0; f < 10; ++f) F[f] = st.s[f].data();
#pragma omp parallel for schedule(static)
    for (int lz = 0; lz < st.lnz; ++lz) {
        for (int ly = 0; ly < st.lny; ++ly) {
            for (int lx = 0; lx < st.lnx; ++lx) {
                const R u = (lx + 0.5) / st.lnx, v = (ly + 0.5) / st.lny, w = (lz + 0.5) / st.lnz;
                const R fade = domain_fade(u, v, w);
                const R heat = sample_scalar(st, F[HEAT], u, v, w);
                const R fuel = sample_scalar(st, F[FUEL], u, v, w);
                const R oxygen = sample_scalar(st, F[OXYGEN], u, v, w);
                const R soot = sample_scalar(st, F[SOOT], u, v, w);
                const R charMass = sample_scalar(st, F[CHARF], u, v, w);
                const R ash = sample_scalar(st, F[ASH], u, v, w);
                const R pyro = sample_scalar(st, F[PYRO], u, v, w);
                const R progress = sample_scalar(st, F[PROGRESS], u, v, w);
                const R sootOptics = sample_scalar(st, F[SOPT], u, v, w);
                const R cellX = 1.0f / static_cast<float>(st.nx);
                const R cellY = 1.0f / static_cast<float>(st.ny);
                const R cellZ = 1.0f / static_cast<float>(st.nz);
                const R gx = sample_scalar(st, F[HEAT], u + cellX, v, w) - sample_scalar(st, F[HEAT], u - cellX, v, w);
                const R gy = sample_scalar(st, F[HEAT], u, v + cellY, w) - sample_scalar(st, F[HEAT], u, v - cellY, w);
                const R gz = sample_scalar(st, F[HEAT], u, v, w + cellZ) - sample_scalar(st, F[HEAT], u, v, w - cellZ);
                const R front = std::sqrt(gx * gx + gy * gy + gz * gz) + fuel * oxygen * 0.035f + progress * 0.16f;
                const R t = p.time;
                const R fine = fbm3(u * 34.0f + t * 0.45f, v * 42.0f - t * 0.82f, w * 34.0f + t * 0.22f, p.saltNoise3);
                const R shear = fbm3(u * 12.0f - t * 0.38f, v * 11.0f - t * 0.62f, w * 13.0f + t * 0.20f, p.saltNoise3);
                const R reactionFront = smoothstep(0.035f, 0.54f, front + pyro * 0.10f);
                const R heightFade = smoothstep(0.70f, 0.070f, v);
                const R lowerShadow = smoothstep(0.62f, 0.06f, v);
                const R plumeShadow = smoothstep(0.12f, 0.88f, v);
                const R sheetGate = smoothstep(0.56f, 0.94f, fine * 0.48f + shear * 0.42f + reactionFront * 0.26f + heat * 0.018f) *
                                    smoothstep(0.04f, 0.86f, reactionFront) * heightFade;
                const R sheetCore = sat(sheetGate * (0.42f + progress * 0.22f + pyro * 0.34f));
                const R tempK = 293.0f + heat * 360.0f + fuel * 44.0f + pyro * 120.0f + progress * 70.0f + sheetCore * 1020.0f;
                const R flameEmitter = fade * smoothstep(820.0f, 1880.0f, tempK) * smoothstep(0.010f, 0.68f, fuel + pyro * 0.30f + heat * 0.040f) *
                                       smoothstep(0.030f, 0.82f, oxygen) * reactionFront * sheetGate * (0.24f + sheetCore * 1.42f);
                const R emberEmitter = smoothstep(0.002f, 0.22f, charMass + ash * 0.18f) * smoothstep(0.09f, 0.00f, v) *
                                       (0.10f + heat * 0.012f + pyro * 0.050f);
                const R sootAtt = std::exp(-(sootOptics * 0.25f + soot * 0.090f + ash * 0.035f));
                const V3 flame = mul3(blackbody(tempK), flameEmitter * sootAtt * 3.40f);
                const V3 coal = mul3(v3(1.0, 0.22f, 0.040f), emberEmitter * p.reflectionGain * 0.20f);
                const R extinction = sat(sootOptics * (0.26f + plumeShadow * 0.42f) + soot * (0.065f + plumeShadow * 0.20f) + ash * 0.050f +
                                         progress * lowerShadow * 0.020f);
                const size_t idx = st.li(lx, ly, lz);
                L[4 * idx + 0] = flame.x + coal.x;
                L[4 * idx + 1] = flame.y + coal.y;
                L[4 * idx + 2] = flame.z + coal.z;
                L[4 * idx + 3] = extinction;
                const R vertDepth = v * (1.0f + plumeShadow * 0.70f);
                const R radial = std::exp(-((u - 0.50f) * (u - 0.50f) * 5.5f + (w - 0.50f) * (w - 0.50f) * 7.0f));
                const R optical = sootOptics * (0.62f + plumeShadow * 1.15f) + soot * (0.22f + plumeShadow * 0.46f) + ash * 0.072f;
                st.shadow[idx] = std::exp(-(optical * (0.72f + vertDepth * 0.42f) + radial * optical * 0.34f + progress * 0.035f));
            }
        }
    }
}

// propagateSceneLightKernel :1527-1565, in -> out.
static void light_propagate(const State& st, const std::vector<R>& in, std::vector<R>& out, R diffusion, R decay) {
    {
#pragma omp parallel for schedule(static)
        for (int lz = 0; lz < st.lnz; ++lz) {
            for (int ly = 0; ly < st.lny; ++ly) {
                for (int lx = 0; lx < st.lnx; ++lx) {
                    const size_t idx = st.li(lx, ly, lz);
                    const int dx[6] = {-1, 1, 0, 0, 0, 0};
                    const int dy[6] = {0, 0, -1, 1, 0, 0};
                    const int dz[6] = {0, 0, 0, 0, -1, 1};
                    V3 nb = v3(0, 0, 0);
                    R weight = 0.0;
                    for (int i = 0; i < 6; ++i) {
                        const int sx = std::min(std::max(lx + dx[i], 0), st.lnx - 1);
                        const int sy = std::min(std::max(ly + dy[i], 0), st.lny - 1);
                        const int sz = std::min(std::max(lz + dz[i], 0), st.lnz - 1);
                        const size_t j = st.li(sx, sy, sz);
                        const R occ = std::exp(-in[4 * j + 3] * 0.42f);
                        nb = add3(nb, mul3(v3(in[4 * j], in[4 * j + 1], in[4 * j + 2]), occ));
                        weight += occ;
                    }
                    nb = mul3(nb, 1.0 / std::max<R>(0.0001f, weight));
                    const R shadow = st.shadow[idx];
                    const V3 direct = v3(in[4 * idx], in[4 * idx + 1], in[4 * idx + 2]);
                    const V3 bounced = lerp3(direct, nb, diffusion);
                    const R floorLift = smoothstep(0.10f, 0.00f, (ly + 0.5) / st.lny);
                    const V3 contact = mul3(v3(1.0, 0.40f, 0.12f), floorLift * luminance3(direct) * 0.18f);
                    const V3 next = add3(mul3(bounced, decay * (0.66f + shadow * 0.34f)), contact);
                    out[4 * idx + 0] = next.x;
                    out[4 * idx + 1] = next.y;
                    out[4 * idx + 2] = next.z;
                    out[4 * idx + 3] = in[4 * idx + 3];
                }
            }
        }
    }
}

// Directional light (an extension, no reference counterpart): Beer-Lambert
// transmittance toward the light from every light-cell centre. Samples at
// s_k = (k + 1/2) ds with weights sat(1/2 + (len - s_k)/ds), len = the ray's
// length inside the volume box (slab exits); sum_k ds w_k = len exactly.
static const R kDirKappa = 3.0f, kDirScatter = 2.0f, kDirDiffuse = 0.10f;
static inline R dir_step(const State& st) {
    return 0.5f * std::min(std::min(static_cast<R>(2.10f) / st.lnx, static_cast<R>(2.03f) / st.lny), static_cast<R>(1.64f) / st.lnz);
}
static inline R slab_exit(R x0, R d, R lo, R hi) {
    if (d > 1.0e-6f) return (hi - x0) / d;
    if (d < -1.0e-6f) return (lo - x0) / d;
    return 1.0e30f;
}
static void dir_transmittance(State& st, const P& p) {
    const R ds = dir_step(st);
    const R dx = -p.dirLightDir[0], dy = -p.dirLightDir[1], dz = -p.dirLightDir[2];
    std::vector<R> ext(st.lcells());
    for (size_t i = 0; i < ext.size(); ++i) ext[i] = st.light[4 * i + 3];
#pragma omp parallel for collapse(2) schedule(static)
    for (int z = 0; z < st.lnz; ++z)
        for (int y = 0; y < st.lny; ++y)
            for (int x = 0; x < st.lnx; ++x) {
                const R x0 = (x + 0.5) / st.lnx * 2.10f - 1.05f;
                const R y0 = (y + 0.5) / st.lny * 2.03f + 0.02f;
                const R z0 = (z + 0.5) / st.lnz * 1.64f - 0.82f;
                const R len = std::min(std::min(slab_exit(x0, dx, -1.05f, 1.05f), slab_exit(y0, dy, 0.02f, 2.05f)), slab_exit(z0, dz, -0.82f, 0.82f));
                R tau = 0.0;
                for (int k = 0;; ++k) {
                    const R s = ds * (k + 0.5);
                    const R wk = std::min<R>(1.0, 0.5 + (len - s) / ds);
                    if (wk <= 0.0) break;
                    const R u = (x0 + dx * s + 1.05f) / 2.10f, v = (y0 + dy * s - 0.02f) / 2.03f, w = (z0 + dz * s + 0.82f) / 1.64f;
                    tau += wk * sample_grid(ext.data(), st.lnx, st.lny, st.lnz, st.lnx - 1, st.lny - 1, st.lnz - 1, u, v, w);
                }
                st.dirT[st.li(x, y, z)] = std::exp(-kDirKappa * ds * tau);
            }
}
static inline R sample_dir_trans(const State& st, R u, R v, R w) {
    return sample_grid(st.dirT.data(), st.lnx, st.lny, st.lnz, st.lnx - 1, st.lny - 1, st.lnz - 1, u, v, w);
}

// Build, then two propagations (0.48, 0.92) and (0.34, 0.96) (:3313-3314).
static void light_pass(State& st, const P& p) {
    const size_t lc = st.lcells();
    std::vector<R> L(4 * lc), tmp(4 * lc);
    light_build(st, p, L);
    light_propagate(st, L, tmp, 0.48f, 0.92f);
    light_propagate(st, tmp, st.light, 0.34f, 0.96f);
    if (p.dirLight) dir_transmittance(st, p);
}

// -------------------------------------------------------------- render ----
// src/fire_cuda.cu:1575-1587
static inline V3 volume_uv_from_world(V3 p) {
    return v3(sat((p.x + 1.05f) / 2.10f), sat((p.y - 0.02f) / 2.03f), sat((p.z + 0.82f) / 1.64f));
}
static inline V3 volume_world_from_uv(V3 uv) {
    return v3(uv.x * 2.10f - 1.05f, uv.y * 2.03f + 0.02f, uv.z * 1.64f - 0.82f);
}

// src/fire_cuda.cu:1589-1626
static V3 gather_irradiance(const State& st, V3 hit, V3 normal) {
    V3 irr = v3(0, 0, 0);
    for (int i = 0; i < 3; ++i) {
        V3 suv = v3(0.50f, 0.16f, 0.50f);
        if (i == 1) suv = v3(0.34f, 0.26f, 0.43f);
        else if (i == 2) suv = v3(0.66f, 0.26f, 0.57f);
        const V4 light = sample_light(st, suv.x, suv.y, suv.z);
        const V3 radiance = v3(light.x, light.y, light.z);
        const R lum = luminance3(radiance);
        if (lum <= 0.00001f) continue;
        const V3 source = volume_world_from_uv(suv);
        const V3 toLight = sub3(source, hit);
        const R dist2 = std::max<R>(0.045f, dot3(toLight, toLight));
        const V3 wi = mul3(toLight, 1.0 / std::sqrt(dist2));
        const R ndotl = std::max<R>(0.0, dot3(normal, wi));
        const R heightOcc = std::exp(-std::max<R>(0.0, hit.y - source.y) * light.w * 1.15f);
        const R vis = sample_shadow(st, suv.x, suv.y, suv.z) * std::exp(-light.w * 0.92f) * heightOcc;
        const R geometric = (0.12f + ndotl * 0.88f) * std::exp(-dist2 * 0.22f) / (0.12f + dist2);
        irr = add3(irr, mul3(radiance, vis * geometric));
    }
    const R grazing = 1.0f + smoothstep(0.0f, 0.45f, 1.0 - std::max<R>(0.0, normal.y)) * 0.36f;
    return mul3(irr, 4.20f * grazing);
}

// src/fire_cuda.cu:1628-1645
static R volume_shadow_ray(const State& st, V3 hit) {
    const V3 target = volume_uv_from_world(hit);
    const V3 source = v3(0.50f, 0.12f, 0.50f);
    R od = 0.0;
    for (int i = 1; i <= 5; ++i) {
        const R t = static_cast<R>(static_cast<float>(i) / 6.0f);
        const V3 uv = lerp3(source, target, t);
        const V4 light = sample_light(st, uv.x, uv.y, uv.z);
        const R shadow = sample_shadow(st, uv.x, uv.y, uv.z);
        od += light.w * (0.18f + t * 0.22f) + (1.0 - shadow) * (0.08f + t * 0.16f);
    }
    return std::exp(-od);
}

// src/fire_cuda.cu:1647-1659
static inline R rect_mask(R px, R py, R cx, R cy, R hx, R hy, R feather) {
    const R dx = std::fabs(px - cx) - hx, dy = std::fabs(py - cy) - hy;
    return smoothstep(feather, -feather, std::max(dx, dy));
}
static inline R line_mask(R px, R py, R ax, R ay, R bx, R by, R thickness) {
    const R pax = px - ax, pay = py - ay, bax = bx - ax, bay = by - ay;
    const R h = sat((pax * bax + pay * bay) / std::max<R>(0.000001f, bax * bax + bay * bay));
    const R dx = pax - bax * h, dy = pay - bay * h;
    return smoothstep(thickness, thickness * 0.25f, std::sqrt(dx * dx + dy * dy));
}

// src/fire_cuda.cu:348-357
static V3 project_point(const Cam& cam, V3 point) {
    const V3 view = sub3(point, cam.eye);
    const R depth = dot3(view, cam.forward);
    if (depth <= 0.02f) return v3(-10.0f, -10.0f, -1.0f);
    const R px = dot3(view, cam.right) / (depth * cam.tanHalfFov * cam.aspect);
    const R py = dot3(view, cam.up) / (depth * cam.tanHalfFov);
    return v3(0.5f + px * 0.5f, 0.5f - py * 0.5f, depth);
}

// src/fire_cuda.cu:1661-1681
static void source_model_overlay(R uvx, R uvy, const Cam& cam, const P& p, V3& color) {
    if (uvx < 0.24f || uvx > 0.76f || uvy < 0.42f || uvy > 0.88f) return;
    if (p.sceneId == 1) return;
    const bool burner = p.sceneId == 2 && p.burnerCount > 0;
    const R sx = burner ? p.burnerX[0] : p.emitterCenterX;
    const R sy = burner ? p.burnerY[0] : p.emitterHeightNorm * 2.03f + 0.02f;
    const R sz = burner ? p.burnerZ[0] : p.emitterCenterZ;
    const V3 c = project_point(cam, v3(sx, sy, sz));
    const R tray = smoothstep(0.135f, 0.105f, std::fabs(uvx - c.x)) * smoothstep(0.050f, 0.036f, std::fabs(uvy - c.y));
    const R inner = smoothstep(0.112f, 0.086f, std::fabs(uvx - c.x)) * smoothstep(0.039f, 0.027f, std::fabs(uvy - c.y));
    const R alpha = sat((tray - inner * 0.55f) * 0.65f);
    const V3 model = lerp3(v3(0.015f, 0.013f, 0.011f), v3(0.13f, 0.12f, 0.10f), alpha);
    if (alpha > 0.001f) color = lerp3(color, model, sat(alpha));
}

// roomBackgroundRay src/fire_cuda.cu:1683-1865. floorGrid (:1567-1573) is
// identically 0, so every term it scales is dropped (lerp by 0 and +0 are exact).
static V3 room_background(const State& st, const P& p, R uvx, R uvy, V3 rd, R glow) {
    const Cam& cam = p.cam;
    const R roomHalf = 2.55f, ceilingY = 2.18f;
    const bool cinematic = p.cinematic != 0;
    R bestT = 1.0e6f;
    V3 hit = v3(0, 0, 0);
    V3 normal = v3(0, 1, 0);
    int surface = 0;
    if (rd.y < -0.0001f) {
        const R t = -cam.eye.y / rd.y;
        const V3 q = add3(cam.eye, mul3(rd, t));
        if (t > 0.0 && std::fabs(q.x) <= roomHalf && std::fabs(q.z) <= roomHalf && t < bestT) {
            bestT = t; hit = q; normal = v3(0, 1, 0); surface = 1;
        }
    }
    if (rd.y > 0.0001f) {
        const R t = (ceilingY - cam.eye.y) / rd.y;
        const V3 q = add3(cam.eye, mul3(rd, t));
        if (t > 0.0 && std::fabs(q.x) <= roomHalf && std::fabs(q.z) <= roomHalf && t < bestT) {
            bestT = t; hit = q; normal = v3(0, -1, 0); surface = 2;
        }
    }
    if (std::fabs(rd.x) > 0.0001f) {
        const R planeX = rd.x > 0.0 ? roomHalf : -roomHalf;
        const R t = (planeX - cam.eye.x) / rd.x;
        const V3 q = add3(cam.eye, mul3(rd, t));
        if (t > 0.0 && q.y >= 0.0 && q.y <= ceilingY && std::fabs(q.z) <= roomHalf && t < bestT) {
            bestT = t; hit = q; normal = v3(planeX > 0.0 ? -1.0 : 1.0, 0, 0); surface = 3;
        }
    }
    if (std::fabs(rd.z) > 0.0001f) {
        const R planeZ = rd.z > 0.0 ? roomHalf : -roomHalf;
        const R t = (planeZ - cam.eye.z) / rd.z;
        const V3 q = add3(cam.eye, mul3(rd, t));
        if (t > 0.0 && q.y >= 0.0 && q.y <= ceilingY && std::fabs(q.x) <= roomHalf && t < bestT) {
            bestT = t; hit = q; normal = v3(0, 0, planeZ > 0.0 ? -1.0 : 1.0); surface = 4;
        }
    }
    V3 color = v3(0.014f, 0.014f, 0.014f);
    V3 albedo = v3(0.26f, 0.26f, 0.25f);
    R giScale = 0.24f;
    if (surface == 1) {
        const R slabTone = std::sin(hit.x * 2.20f + hit.z * 1.40f) * 0.50f + std::sin(hit.z * 3.10f - hit.x * 0.80f) * 0.50f;
        color = cinematic ? v3(0.016f, 0.016f, 0.015f) : v3(0.014f, 0.014f, 0.013f);
        color = add3(color, mul3(v3(0.014f, 0.014f, 0.013f), slabTone * 0.018f));
        albedo = v3(0.26f, 0.26f, 0.25f);
        giScale = 0.22f;
        if (p.sceneId == 1) {
            const R logMat = bed_material(hit.x, hit.z, p);
            const R coal = std::exp(-(hit.x * hit.x * 1.25f + hit.z * hit.z * 2.45f)) *
                           smoothstep(0.28f, 0.90f, fbm2(hit.x * 13.0f, hit.z * 15.0f, p.saltNoise2));
            const R scorch = std::exp(-(hit.x * hit.x * 0.66f + hit.z * hit.z * 1.15f));
            color = lerp3(color, v3(0.018f, 0.015f, 0.012f), sat(logMat * 0.78f + scorch * 0.20f));
            color = add3(color, mul3(v3(0.090f, 0.020f, 0.004f), coal * glow * 0.026f));
            color = add3(color, mul3(v3(0.17f, 0.11f, 0.055f), logMat * 0.050f));
            albedo = lerp3(albedo, v3(0.040f, 0.030f, 0.022f), logMat * 0.86f);
            giScale = lerp(giScale, 0.44f, logMat * 0.80f);
        } else if (p.sceneId == 2) {
            const R radius = std::sqrt(hit.x * hit.x + hit.z * hit.z);
            const R plate = smoothstep(0.78f, 0.70f, radius);
            const R burner = bed_material(hit.x, hit.z, p);
            color = lerp3(color, v3(0.018f, 0.019f, 0.020f), plate * 0.70f);
            color = lerp3(color, v3(0.005f, 0.005f, 0.005f), burner * 0.72f);
            color = add3(color, mul3(v3(0.090f, 0.032f, 0.006f), burner * glow * 0.012f));
            albedo = lerp3(albedo, v3(0.032f, 0.033f, 0.034f), plate * 0.82f);
            giScale = lerp(giScale, 0.42f, plate * 0.72f);
        } else {
            const R trayBody = rect_mask(hit.x, hit.z, 0.0, 0.0, 1.08f, 0.58f, 0.022f);
            const R trayInner = rect_mask(hit.x, hit.z, 0.0, 0.0, 0.96f, 0.47f, 0.020f);
            const R trayRim = sat(trayBody - trayInner * 0.74f);
            const R emberBed = trayInner * std::exp(-(hit.x * hit.x * 0.74f + hit.z * hit.z * 2.20f));
            const R trayContact = trayInner * std::exp(-(hit.x * hit.x * 1.35f + hit.z * hit.z * 3.35f));
            const R scorch = std::exp(-(hit.x * hit.x * 0.82f + hit.z * hit.z * 1.75f));
            color = lerp3(color, v3(0.017f, 0.014f, 0.012f), trayBody * 0.82f);
            color = lerp3(color, v3(0.020f, 0.019f, 0.018f), scorch * (1.0 - trayBody) * 0.16f);
            color = lerp3(color, v3(0.010f, 0.009f, 0.008f), trayContact * 0.22f);
            color = add3(color, mul3(v3(0.056f, 0.014f, 0.003f), emberBed * glow * 0.26f));
            color = add3(color, mul3(v3(0.30f, 0.28f, 0.24f), trayRim * 0.10f));
            albedo = lerp3(albedo, v3(0.030f, 0.024f, 0.020f), trayBody * 0.75f);
            giScale = lerp(giScale, 0.48f, trayBody * 0.85f);
        }
        const R reflectCore = std::exp(-(hit.x * hit.x * 2.45f + (hit.z + 0.18f) * (hit.z + 0.18f) * 1.95f)) * glow;
        color = add3(color, mul3(v3(0.55f, 0.11f, 0.023f), reflectCore * p.reflectionGain * 0.48f));
    } else if (surface == 2) {
        color = cinematic ? v3(0.012f, 0.012f, 0.012f) : v3(0.011f, 0.011f, 0.010f);
        albedo = v3(0.24f, 0.24f, 0.23f);
        giScale = 0.20f;
        const R panelA = line_mask(hit.x, hit.z, -1.2f, -2.0f, -1.2f, 2.0f, 0.010f);
        const R panelB = line_mask(hit.x, hit.z, 0.0f, -2.0f, 0.0f, 2.0f, 0.008f);
        const R fixL = std::exp(-((hit.x + 1.50f) * (hit.x + 1.50f) + (hit.z - 0.92f) * (hit.z - 0.92f)) * 120.0f);
        const R fixR = std::exp(-((hit.x - 1.72f) * (hit.x - 1.72f) + (hit.z - 0.80f) * (hit.z - 0.80f)) * 120.0f);
        color = add3(color, mul3(v3(0.10f, 0.096f, 0.088f), (panelA + panelB) * 0.08f));
        color = add3(color, mul3(v3(0.26f, 0.20f, 0.13f), (fixL + fixR) * 0.10f));
    } else {
        color = cinematic ? v3(0.015f, 0.015f, 0.014f) : color;
        albedo = v3(0.25f, 0.25f, 0.24f);
        giScale = 0.24f;
        const R baseShadow = smoothstep(0.34f, 0.02f, hit.y);
        const R crownShadow = smoothstep(ceilingY - 0.32f, ceilingY - 0.02f, hit.y);
        color = lerp3(color, v3(0.012f, 0.012f, 0.011f), sat(baseShadow * 0.30f + crownShadow * 0.20f));
        if (surface == 4 && hit.z > 0.0) {
            const R recess = rect_mask(hit.x, hit.y, 0.0, 1.00f, 0.92f, 0.72f, 0.055f);
            const R recessEdge = line_mask(hit.x, hit.y, -0.92f, 0.28f, -0.92f, 1.72f, 0.010f) +
                                 line_mask(hit.x, hit.y, 0.92f, 0.28f, 0.92f, 1.72f, 0.010f);
            color = lerp3(color, v3(0.015f, 0.015f, 0.014f), recess * 0.42f);
            color = add3(color, mul3(v3(0.050f, 0.046f, 0.038f), recessEdge * 0.030f));
        }
        const R soot = std::exp(-(hit.x * hit.x * 0.82f + hit.z * hit.z * 1.12f)) * smoothstep(0.36f, 1.92f, hit.y);
        color = lerp3(color, v3(0.010f, 0.011f, 0.012f), soot * 0.82f * p.smokeDarkness);
        albedo = lerp3(albedo, v3(0.030f, 0.032f, 0.034f), soot * 0.75f);
        const R wallGlow = std::exp(-(hit.x * hit.x * 1.1f + hit.z * hit.z * 1.3f + (hit.y - 0.70f) * (hit.y - 0.70f) * 1.0f)) * glow;
        color = add3(color, mul3(v3(0.15f, 0.050f, 0.018f), wallGlow * 0.090f));
        if (surface == 3 && hit.x < 0.0) {
            const R window = rect_mask(hit.z, hit.y, -1.28f, 0.92f, 0.06f, 0.62f, 0.030f);
            const R glass = 0.55f + 0.25f * std::sin(hit.y * 23.0f + hit.z * 15.0f);
            color = lerp3(color, v3(0.009f, 0.011f, 0.012f), window * 0.92f);
            color = add3(color, mul3(v3(0.10f, 0.12f, 0.13f), window * glass * 0.16f));
        }
    }
    const V3 vuv = volume_uv_from_world(hit);
    const V4 gi = sample_light(st, vuv.x, vuv.y, vuv.z);
    const R sceneShadow = sample_shadow(st, vuv.x, vuv.y, vuv.z);
    const R volVis = volume_shadow_ray(st, hit);
    const V3 gathered = gather_irradiance(st, hit, normal);
    const V3 indirect = mul3(add3(mul3(v3(gi.x, gi.y, gi.z), 0.42f), gathered), volVis);
    const R grazingFalloff = surface == 2 ? 0.62f : 1.0f;
    const R contactOcc = smoothstep(0.005f, 0.42f, sceneShadow) * std::exp(-gi.w * 0.72f) * (0.30f + volVis * 0.70f);
    const R lift = 0.66f + smoothstep(0.006f, 0.13f, luminance3(indirect)) * 0.72f;
    const R warmth = surface == 1 ? 1.72f : (surface == 2 ? 0.88f : 1.16f);
    color = add3(color, mul3(mul3v(indirect, albedo), giScale * grazingFalloff * contactOcc * lift * warmth));
    if (p.dirLight) {
        const R ndl = std::max<R>(0.0, -(normal.x * p.dirLightDir[0] + normal.y * p.dirLightDir[1] + normal.z * p.dirLightDir[2]));
        const R td = sample_dir_trans(st, vuv.x, vuv.y, vuv.z);
        color = add3(color, mul3(mul3v(albedo, v3(p.dirLightColor[0], p.dirLightColor[1], p.dirLightColor[2])), ndl * td * kDirDiffuse));
    }
    const R shadowedSoot = (1.0 - volVis) * smoothstep(0.18f, 0.82f, 1.0 - sceneShadow);
    color = lerp3(color, v3(0.006f, 0.007f, 0.008f), shadowedSoot * (surface == 2 ? 0.38f : 0.26f));
    const R wallEdge = surface == 3 ? roomHalf - std::fabs(hit.z) : roomHalf - std::fabs(hit.x);
    const R corner = (surface == 1 || surface == 2) ? std::min(roomHalf - std::fabs(hit.x), roomHalf - std::fabs(hit.z))
                                                    : std::min(wallEdge, std::min(hit.y, ceilingY - hit.y));
    color = mul3(color, 0.84f + smoothstep(0.0f, 0.54f, corner) * 0.16f);
    const R gx = std::floor(hit.x * 19.0f + hit.z * 13.0f), gy = std::floor(hit.y * 15.0f + p.time * 0.08f);
    const R grain = (hash2(static_cast<int32_t>(gx), static_cast<int32_t>(gy), p.saltGrain) - 0.5f) * 0.006f;
    color = add3(color, v3(grain, grain, grain));
    source_model_overlay(uvx, uvy, cam, p, color);
    const R fog = smoothstep(3.8f, 0.2f, bestT);
    const R vignette = smoothstep(0.78f, 0.20f, std::sqrt((uvx - 0.50f) * (uvx - 0.50f) + (uvy - 0.46f) * (uvy - 0.46f)));
    return mul3(color, (0.12f + fog * 0.56f) * (0.58f + vignette * 0.30f));
}

// src/fire_cuda.cu:1867-1907
static bool intersect_box(V3 ro, V3 rd, V3 bmin, V3 bmax, R& tn, R& tf) {
    R tNear = 0.0, tFar = 1.0e6f;
    auto inv = [](R d) { return 1.0 / (std::fabs(d) < 0.00001f ? std::copysign(static_cast<R>(0.00001f), d) : d); };
    const R ro3[3] = {ro.x, ro.y, ro.z}, inv3[3] = {inv(rd.x), inv(rd.y), inv(rd.z)};
    const R lo3[3] = {bmin.x, bmin.y, bmin.z}, hi3[3] = {bmax.x, bmax.y, bmax.z};
    for (int a = 0; a < 3; ++a) {
        R t0 = (lo3[a] - ro3[a]) * inv3[a];
        R t1 = (hi3[a] - ro3[a]) * inv3[a];
        if (t0 > t1) std::swap(t0, t1);
        tNear = std::max(tNear, t0);
        tFar = std::min(tFar, t1);
    }
    tn = tNear;
    tf = tFar;
    return tFar > tNear;
}

// src/fire_cuda.cu:370-379
static V3 debug_heat_color(R value) {
    value = sat(value);
    const R red = smoothstep(0.05f, 0.45f, value);
    const R orange = smoothstep(0.24f, 0.78f, value);
    const R white = smoothstep(0.72f, 1.0f, value);
    return v3(sat(red + white * 0.70f), sat(orange * 0.58f + white * 0.78f), sat(white * 0.62f));
}

// renderKernel src/fire_cuda.cu:1953-2244 for one pixel.
static V4 render_pixel(const State& st, const P& p, int x, int y) {
    const R* F[10];
    for (int f = 0; f < 10; ++f) F[f] = st.s[f].data();
    const Cam& cam = p.cam;
    const R uvx = (x + 0.5) / p.fw, uvy = (y + 0.5) / p.fh;
    const R sx = (uvx * 2.0 - 1.0) * cam.aspect * cam.tanHalfFov;
    const R sy = (1.0 - uvy * 2.0) * cam.tanHalfFov;
    const V3 rd = normalize3(add3(add3(cam.forward, mul3(cam.right, sx)), mul3(cam.up, sy)));
    const R baseGlow = sat(sample_scalar(st, F[HEAT], 0.50f, 0.06f, 0.50f) * 0.42f);
    R boxNear = 0, boxFar = 0;
    V3 accum = v3(0, 0, 0);
    R trans = 1.0, flameGlow = 0.0, smokeOcc = 0.0;
    R dbgFlame = 0, dbgOD = 0, dbgTemp = 0, dbgFuel = 0, dbgVel = 0;
    const int sid = p.sceneId;
    const R t = p.time;
    if (intersect_box(cam.eye, rd, v3(-1.05f, 0.02f, -0.82f), v3(1.05f, 2.05f, 0.82f), boxNear, boxFar)) {
        const int steps = std::min(std::max(p.raymarchSteps, 24), 256);
        const R t0 = std::max<R>(0.0, boxNear);
        const R t1 = std::min(boxFar, t0 + 3.05f);
        const R stepT = (t1 - t0) / steps;
        const R jitter = hash2(x, y, p.saltJitter);
        for (int i = 0; i < steps && trans > 0.018f; ++i) {
            const R tt = t0 + (i + jitter) * stepT;
            const V3 wp = add3(cam.eye, mul3(rd, tt));
            const R fu = sat((wp.x + 1.05f) / 2.10f), fv = sat((wp.y - 0.02f) / 2.03f), fw = sat((wp.z + 0.82f) / 1.64f);
            const R fade = domain_fade(fu, fv, fw);
            if (fade <= 0.001f) continue;
            const R cHeat = sample_scalar(st, F[HEAT], fu, fv, fw);
            const R cFuel = sample_scalar(st, F[FUEL], fu, fv, fw);
            const R cSoot = sample_scalar(st, F[SOOT], fu, fv, fw);
            const R cProg = sample_scalar(st, F[PROGRESS], fu, fv, fw);
            const R coarse = (cHeat + cFuel * 0.42f + cSoot * 0.34f + cProg * 0.32f) * fade;
            if (coarse <= 0.00150f) {
                ++i;
                continue;
            }
            const R warpA = fbm3(fu * 7.0f + t * 0.18f, fv * 5.0f - t * 0.11f, fw * 9.0f + t * 0.07f, p.saltNoise3) - 0.5f;
            const R warpB = fbm3(fw * 8.0f - t * 0.14f, fu * 8.0f + t * 0.10f, fv * 6.0f - t * 0.08f, p.saltNoise3) - 0.5f;
            const R warpGain = 0.028f + fv * 0.060f;
            const R wu = sat(fu + warpA * warpGain), wv = sat(fv + (warpA - warpB) * 0.020f), ww = sat(fw + warpB * warpGain);
            const R heat = sample_scalar(st, F[HEAT], wu, wv, ww);
            const R fuel = sample_scalar(st, F[FUEL], wu, wv, ww);
            const R oxygen = sample_scalar(st, F[OXYGEN], wu, wv, ww);
            const R soot = sample_scalar(st, F[SOOT], wu, wv, ww);
            const R charMass = sample_scalar(st, F[CHARF], wu, wv, ww);
            const R ash = sample_scalar(st, F[ASH], wu, wv, ww);
            const R pyro = sample_scalar(st, F[PYRO], wu, wv, ww);
            const R progress = sample_scalar(st, F[PROGRESS], wu, wv, ww);
            const R turb = sample_scalar(st, F[TURB], wu, wv, ww);
            const R sootOptics = sample_scalar(st, F[SOPT], wu, wv, ww);
            const V4 sceneLight = sample_light(st, wu, wv, ww);
            const R sceneShadow = sample_shadow(st, wu, wv, ww);
            V3 vel = v3(0, 0, 0);
            if (p.debugMode == 6) vel = sample_velocity(st, st.u.data(), st.v.data(), st.w.data(), wu, wv, ww);
            const R active = (heat + fuel * 0.42f + soot * 0.34f + progress * 0.32f + pyro * 0.18f) * fade;
            if (active <= 0.0025f) continue;
            dbgFuel = std::max(dbgFuel, fade * sat(fuel * 0.18f + charMass * 0.22f + pyro * 0.30f));
            if (p.debugMode == 6) dbgVel = std::max(dbgVel, sat(std::sqrt(dot3(vel, vel)) * 0.22f));
            const R cellX = 1.0f / static_cast<float>(st.nx);
            const R cellY = 1.0f / static_cast<float>(st.ny);
            const R cellZ = 1.0f / static_cast<float>(st.nz);
            const R gx = sample_scalar(st, F[HEAT], wu + cellX, wv, ww) - sample_scalar(st, F[HEAT], wu - cellX, wv, ww);
            const R gy = sample_scalar(st, F[HEAT], wu, wv + cellY, ww) - sample_scalar(st, F[HEAT], wu, wv - cellY, ww);
            const R gz = sample_scalar(st, F[HEAT], wu, wv, ww + cellZ) - sample_scalar(st, F[HEAT], wu, wv, ww - cellZ);
            const R frontGradient = std::sqrt(gx * gx + gy * gy + gz * gz) + std::fabs(fuel - cFuel) * 0.62f;
            const R reactionFront = smoothstep(0.020f, 0.46f, frontGradient * 1.35f + fuel * oxygen * 0.040f + progress * 0.18f + pyro * 0.080f);
            const R fine = fbm3(wu * 42.0f + t * 0.72f, wv * 54.0f - t * 1.28f, ww * 43.0f + t * 0.38f, p.saltNoise3);
            const R shear = fbm3(wu * 15.0f - t * 0.72f, wv * 13.0f - t * 1.08f, ww * 17.0f + t * 0.32f, p.saltNoise3);
            const R plume = sat(fine * 0.58f + shear * 0.42f);
            const R hole = sat((1.0 - fine) * 0.36f + shear * 0.64f);
            const R les = sat(turb * 0.48f + progress * 0.28f);
            const R ragged = smoothstep(0.18f, 0.92f, fine * (0.46f + les * 0.30f) + shear * 0.38f + plume * 0.32f + sat(active * 0.08f));
            const R fieldEdge = smoothstep(0.05f, 2.30f, heat + fuel * 0.32f) * (0.10f + reactionFront * 0.90f);
            const R filament = smoothstep(0.60f, 1.0f, fine + les * 0.10f) * smoothstep(0.38f, 0.92f, shear + (0.72f - fv) * 0.12f + reactionFront * 0.10f) *
                               smoothstep(1.0f, 0.02f, fv) * fieldEdge;
            const R srcV = sid == 2 ? std::max<R>(0.0, fv - p.emitterHeightNorm) : fv;
            const R lowJet = sid == 2
                ? smoothstep(std::max<R>(0.012f, p.emitterHeightBandNorm * 0.68f), std::max<R>(0.003f, p.emitterHeightBandNorm * 0.10f),
                             std::fabs(fv - p.emitterHeightNorm)) * smoothstep(0.05f, 0.58f, fuel + heat * 0.025f) * smoothstep(0.50f, 1.0f, oxygen)
                : 0.0;
            const R portJet = sid == 2 ? lowJet * smoothstep(0.68f, 0.98f, fine + shear * 0.18f) : 0.0;
            const R campBias = sid == 1 ? smoothstep(0.035f, 0.72f, fuel + pyro * 0.50f + charMass * 0.12f) * smoothstep(0.76f, 0.035f, fv) : 0.0;
            const R traySheet = sid == 0 ? smoothstep(0.018f, 0.42f, fuel + pyro * 0.72f + charMass * 0.08f) * smoothstep(0.30f, 0.015f, fv) : 0.0;
            const R convective = smoothstep(0.42f, 0.88f, shear + heat * 0.040f + plume * 0.14f + les * 0.10f + portJet * 0.56f + campBias * 0.20f + traySheet * 0.24f) *
                                 (sid == 2 ? smoothstep(p.emitterHeightNorm + p.emitterHeightBandNorm * 1.20f, p.emitterHeightNorm - p.emitterHeightBandNorm * 0.08f, fv)
                                           : smoothstep(sid == 1 ? 0.82f : 0.88f, 0.030f, fv)) *
                                 fieldEdge;
            const R breakup = smoothstep(0.34f, 0.88f, fine * 0.48f + shear * 0.42f + hole * 0.24f + heat * 0.026f);
            const R verticalFade = smoothstep(0.96f, 0.025f, fv);
            const R heightFade = sid == 2 ? smoothstep(p.emitterHeightNorm + p.emitterHeightBandNorm * 0.72f, p.emitterHeightNorm - p.emitterHeightBandNorm * 0.05f, fv)
                                          : smoothstep(sid == 1 ? 0.72f : 0.76f, 0.10f, fv);
            const R lowerWhite = smoothstep(0.060f, 0.00f, srcV) * smoothstep(0.05f, 1.30f, pyro + fuel * 0.20f);
            const R tempK = 293.0f + heat * 360.0f + fuel * 44.0f + pyro * 90.0f + filament * 560.0f + convective * 420.0f + lowerWhite * 110.0f + progress * 64.0f;
            const R combustion = smoothstep(740.0f, 1630.0f, tempK) * smoothstep(0.016f, 0.62f, fuel + heat * 0.055f) * smoothstep(0.045f, 0.62f, oxygen) *
                                 verticalFade * heightFade;
            const R thinFront = smoothstep(0.04f, 0.56f, frontGradient + progress * 0.10f) * smoothstep(0.02f, 0.92f, reactionFront);
            const R tear = smoothstep(0.52f, 0.88f, hole + shear * 0.18f + les * 0.30f + sat(1.0 - oxygen) * 0.12f);
            const R sheetHole = smoothstep(0.54f, 0.90f, hole + les * 0.26f + sat(1.0 - oxygen) * 0.12f);
            const R flameSheet = std::max(
                smoothstep(0.55f, 0.94f, fine * 0.46f + shear * 0.44f + plume * 0.18f + reactionFront * 0.15f + heat * 0.014f - fv * 0.050f) *
                    (1.0 - sheetHole * (0.74f + les * 0.12f)) * (1.0 - tear * 0.18f),
                smoothstep(0.18f, 0.030f, srcV) * smoothstep(0.75f, 2.0f, heat) * thinFront * 0.030f);
            const R coherent = sat(filament * 0.86f + convective * 0.55f + thinFront * 0.74f);
            const R confinement = sat(coherent + reactionFront * 0.08f);
            const R coverage = sid == 0 ? (0.72f + traySheet * 1.24f) : 1.0;
            const R flameDensity = fade * combustion * reactionFront * flameSheet * (0.010f + filament * 0.92f + convective * 0.24f + thinFront * 1.22f) *
                                   (0.052f + breakup * 0.78f + ragged * 0.56f + les * 0.34f) * (0.18f + thinFront * 2.55f) *
                                   (0.060f + confinement * 0.94f + portJet * 0.72f + campBias * 0.34f) * heightFade * coverage;
            const R plumeVoid = smoothstep(0.54f, 0.90f, hole + fine * 0.26f + shear * 0.22f + fv * 0.20f);
            const R topDissolve = smoothstep(sid == 1 ? 0.82f : 0.88f, sid == 1 ? 0.34f : 0.42f, fv);
            const R raggedPlume = 1.0 - plumeVoid * smoothstep(0.18f, 0.74f, fv) * 0.96f;
            const R upperPlume = smoothstep(0.14f, 0.48f, fv) * smoothstep(0.030f, 1.20f, sootOptics + soot * 0.42f) *
                                 (0.025f + fine * 0.035f + shear * 0.034f + plume * 0.34f + ragged * 0.12f + turb * 0.035f) * topDissolve * raggedPlume;
            const R smokeScale = sid == 2 ? 0.035f : (sid == 1 ? 0.56f : 1.0f);
            const R rawSoot = fade * smokeScale *
                              (smoothstep(0.020f, 1.05f, sootOptics + soot * 0.32f) * smoothstep(0.16f, 0.42f, fv) *
                                   (0.028f + fine * 0.040f + shear * 0.052f + plume * 0.38f + ragged * 0.16f + ash * 0.016f) +
                               upperPlume * (0.22f + plume * 0.30f) + smoothstep(0.004f, 0.34f, sootOptics) * (0.020f + plume * 0.034f));
            const R frontMask = sat(flameDensity * 4.8f + combustion * reactionFront * flameSheet * (0.08f + thinFront * 0.56f + filament * 0.18f));
            const R emissive = sat(frontMask);
            const R resolved = smoothstep(0.030f, 0.32f, emissive) * smoothstep(0.035f, 0.72f, oxygen);
            const R smokeOnly = 1.0 - smoothstep(0.025f, 0.20f, resolved);
            const R scatterSep = sat(smokeOnly * (1.0 - smoothstep(0.02f, 0.16f, emissive) * 0.58f));
            const R cleanout = 1.0 - smoothstep(0.025f, 0.34f, resolved) * 0.82f;
            const R smokeOnlyHeight = smoothstep(0.20f, 0.70f, fv) * (1.0 - smoothstep(0.02f, 0.20f, emissive));
            const R neutrality = smoothstep(0.18f, 0.78f, fv) * scatterSep;
            const R absorption = rawSoot * cleanout * (1.10f + smokeOnlyHeight * 2.20f + neutrality * 2.40f);
            const R scatterDensity = rawSoot * scatterSep * smoothstep(0.10f, 0.55f, fv + sootOptics * 0.10f) * (0.24f + (1.0 - smokeOnlyHeight) * 0.34f);
            const R radiusP = sat(0.08f + sootOptics * 0.028f + ash * 0.085f + sat(1.0 - oxygen) * 0.12f);
            const R sootAbs = absorption * (3.10f + radiusP * 4.60f + sootOptics * 0.70f + neutrality * 2.85f) * p.smokeDarkness;
            const R sootScat = scatterDensity * (0.018f + (1.0 - radiusP) * 0.050f) * (0.20f + p.smokeGain * 0.10f);
            const R sootExt = sootAbs + sootScat;
            const R flameExt = flameDensity * (0.12f + (1.0 - resolved) * 0.12f);
            const R extinction = (sootExt + flameExt) * stepT;
            const R smokeAlpha = 1.0 - std::exp(-sootAbs * stepT);
            const R scatterAlpha = 1.0 - std::exp(-sootScat * stepT);
            const R lightVis = sceneShadow * (0.35f + sceneShadow * 0.65f);
            const V3 localIrr = v3(sceneLight.x, sceneLight.y, sceneLight.z);
            const R sootLoad = sat(sootOptics * 0.74f + soot * 0.20f + upperPlume * 0.46f);
            const R whiteCore = smoothstep(2450.0f, 3450.0f, tempK) * combustion * smoothstep(0.32f, 0.90f, oxygen) * smoothstep(0.72f, 0.02f, fv) *
                                (0.02f + thinFront * 0.48f + filament * 0.18f) * (1.0 - smoothstep(0.16f, 0.74f, sootLoad) * 0.68f);
            const R radiantBase = sat((tempK - 760.0f) / 2000.0f);
            const R radiantCurve = radiantBase * radiantBase * (0.82f + radiantBase * 0.18f);
            const R radiantPower = radiantCurve * (1.42f + filament * 3.15f + convective * 1.60f + progress * 0.26f + whiteCore * 0.92f);
            const R coreStructure = sat(whiteCore * (0.080f + filament * 0.28f + thinFront * 0.32f));
            const R whiteFil = coreStructure * (0.55f + fine * 0.30f + shear * 0.18f);
            const R orangeEdge = (1.0 - whiteFil * 0.72f) * thinFront * reactionFront * flameSheet * (0.52f + breakup * 0.95f) * smoothstep(0.08f, 0.80f, oxygen);
            V3 flameColor = lerp3(v3(1.32f, 0.27f, 0.038f), blackbody(tempK), sat(whiteCore * 0.92f + whiteFil * 0.82f));
            if (sid == 2) {
                const R blueBase = portJet * smoothstep(0.010f, 0.11f, flameDensity + combustion * 0.18f);
                flameColor = lerp3(v3(0.020f, 0.18f, 2.40f), v3(0.36f, 0.58f, 1.95f), sat(whiteCore * 0.34f + blueBase * 0.20f));
            } else if (sid == 1) {
                flameColor = lerp3(flameColor, v3(1.52f, 0.24f, 0.028f), sat(campBias * (1.0 - whiteCore) * 0.58f));
            }
            V3 emission = mul3(flameColor, flameDensity * radiantPower * stepT * (1.10f + whiteCore * 0.38f));
            if (sid == 2) {
                emission = mul3(emission, 0.42f);
                emission = add3(emission, mul3(v3(0.012f, 0.28f, 5.20f), portJet * combustion * stepT * 4.40f));
            }
            emission = add3(emission, mul3(v3(1.10f, 1.02f, 0.86f), whiteFil * whiteFil * flameDensity * radiantPower * stepT * 0.82f));
            emission = add3(emission, mul3(v3(1.70f, 0.30f, 0.040f),
                                           orangeEdge * flameDensity * radiantPower * stepT * (sid == 2 ? 0.06f : (sid == 1 ? 3.20f : 3.10f))));
            emission = add3(emission, mul3(v3(1.36f, 0.070f, 0.008f), filament * (1.0f + tear * 0.55f) * flameDensity * radiantPower * stepT *
                                                                         (sid == 2 ? 0.035f : (sid == 1 ? 1.55f : 1.10f)) * (1.0 - whiteCore * 0.62f)));
            const R ashVeil = sat(ash * 0.014f + smoothstep(0.30f, 0.88f, fv) * (1.0 - sootLoad) * 0.003f);
            V3 smokeColor = lerp3(v3(0.008f, 0.0092f, 0.0115f), v3(0.0018f, 0.0020f, 0.0024f), sootLoad);
            smokeColor = lerp3(smokeColor, v3(0.014f, 0.0165f, 0.0210f), ashVeil);
            smokeColor = mul3(smokeColor, smokeOnly);
            const R irrLuma = luminance3(localIrr);
            const R warmDamp = 1.0 - neutrality * 0.985f;
            const R warmLeak = (1.0 - neutrality) * (1.0 - smokeOnlyHeight) * 0.18f;
            const R warmScatter = sid == 2 ? 0.010f : (sid == 1 ? 0.18f : 0.30f);
            const R backScatter = scatterAlpha * lightVis * (0.00085f + irrLuma * 0.0011f * warmDamp + baseGlow * 0.00010f * warmScatter) *
                                  (0.018f + (1.0 - radiusP) * 0.018f) * scatterSep;
            smokeColor = add3(smokeColor, mul3(v3(0.010f, 0.012f, 0.016f), backScatter));
            const R scatterGain = scatterAlpha * (0.00046f + lightVis * 0.00072f * warmDamp + flameDensity * 0.00008f * warmScatter + turb * 0.00018f) *
                                  (0.038f + (1.0 - radiusP) * 0.045f) * scatterSep;
            const R sceneScatter = scatterAlpha * scatterSep * (0.0065f + (1.0 - radiusP) * 0.014f) * lightVis;
            const V3 neutralLight = v3(irrLuma * 0.052f, irrLuma * 0.064f, irrLuma * 0.086f);
            const V3 smokeLight = lerp3(neutralLight, localIrr, warmLeak);
            const R coalGlow = smoothstep(0.004f, 0.18f, charMass) * smoothstep(0.54f, 0.02f, fv) * (0.070f + pyro * 0.12f + heat * 0.009f);
            dbgFlame = std::max(dbgFlame, sat(flameDensity * radiantPower * 0.26f));
            dbgOD = sat(dbgOD + sootExt * stepT);
            dbgTemp = std::max(dbgTemp, sat((tempK - 520.0f) / 1900.0f));
            accum = add3(accum, mul3(emission, trans));
            accum = add3(accum, mul3(v3(1.0, 0.24f, 0.035f), trans * coalGlow * stepT));
            accum = add3(accum, mul3(smokeColor, trans * scatterGain * smokeOnly));
            accum = add3(accum, mul3(smokeLight, trans * sceneScatter * (0.18f + warmDamp * 0.42f)));
            if (p.dirLight) {
                const R td = sample_dir_trans(st, wu, wv, ww);
                accum = add3(accum, mul3(v3(p.dirLightColor[0], p.dirLightColor[1], p.dirLightColor[2]), trans * scatterAlpha * td * kDirScatter));
            }
            flameGlow = std::max(flameGlow, flameDensity * radiantPower * (0.12f + filament * 0.18f + whiteFil * 0.20f));
            smokeOcc = sat(smokeOcc + smokeAlpha * trans);
            trans *= std::exp(-extinction);
        }
    }
    if (p.debugMode != 0) {
        V3 c = v3(0, 0, 0);
        if (p.debugMode == 1) c = debug_heat_color(dbgFlame);
        else if (p.debugMode == 2) { const R od = sat(dbgOD); c = v3(od, od * 0.90f, od * 0.78f); }
        else if (p.debugMode == 3) c = v3(trans, trans, trans);
        else if (p.debugMode == 4) c = debug_heat_color(dbgTemp);
        else if (p.debugMode == 5) c = v3(dbgFuel * 0.82f, dbgFuel * 0.48f, dbgFuel * 0.18f);
        else c = v3(dbgVel * 0.20f, dbgVel * 0.72f, dbgVel);
        return V4{c.x, c.y, c.z, 1.0};
    }
    V3 color = accum;
    if (trans > 0.006f) {
        const V3 room = room_background(st, p, uvx, uvy, rd, baseGlow);
        color = add3(mul3(room, trans), accum);
    }
    const R bloom = sat(flameGlow * 0.10f) * (1.0 - sat(smokeOcc * 0.95f));
    color = add3(color, mul3(v3(1.0, 0.25f, 0.045f), bloom * bloom * 0.017f * p.reflectionGain));
    return V4{std::max<R>(0.0, color.x), std::max<R>(0.0, color.y), std::max<R>(0.0, color.z), 1.0};
}

// emberHdrKernel src/fire_cuda.cu:2323-2452: sparks splatted additively into
// the HDR frame, spark after spark. The per-spark hashes are hash2(i, k) with
// k = 0..3 under the ember salt (the reference hashes (i, 3.7 / 9.1 / 19.4 /
// 31.6) with its sin hash); everything downstream of them is the reference's.
static void splat_embers(const State& st, const P& p, R* out) {
    const int count = std::min(std::max(p.emberCount, 0), 192);
    if (count == 0 || p.debugMode != 0) return;
    const Cam& cam = p.cam;
    const R baseGlow = sat(sample_scalar(st, st.s[HEAT].data(), 0.50f, 0.06f, 0.50f) * 0.42f);
    const R sparkGate = smoothstep(0.02f, 0.30f, baseGlow);
    if (sparkGate <= 0.0001f) return;
    const int sid = p.sceneId;
    for (int i = 0; i < count; ++i) {
        const R h0 = hash2(i, 0, p.saltEmber), h1 = hash2(i, 1, p.saltEmber);
        const R h2 = hash2(i, 2, p.saltEmber), h3 = hash2(i, 3, p.saltEmber);
        if (sid == 1 && h2 > 0.34f) continue;
        if (sid == 2 && h0 > 0.14f) continue;
        const R lt = p.time * (0.20f + h2 * 0.18f + (sid == 2 ? 0.36f : 0.0f)) + h0;
        const R life = lt - std::floor(lt);
        const R angle = h0 * 6.2831853f;
        R startX = 0.0, startZ = 0.0;
        if (sid == 2) {
            const R radius = std::max<R>(0.04f, p.emitterRadius);
            const R cx = p.burnerCount > 0 ? p.burnerX[0] : p.emitterCenterX;
            const R cz = p.burnerCount > 0 ? p.burnerZ[0] : p.emitterCenterZ;
            startX = cx + std::cos(angle) * radius * (0.82f + h1 * 0.18f);
            startZ = cz + std::sin(angle) * radius * (0.82f + h1 * 0.18f);
        } else if (sid == 1) {
            const R logA = capsule_mask(std::cos(angle) * 0.34f, std::sin(angle) * 0.25f, -0.06f, 0.01f, 0.38f, 0.34f, 0.095f);
            const R logB =
                capsule_mask(std::cos(angle + 1.7f) * 0.35f, std::sin(angle + 1.7f) * 0.25f, 0.07f, -0.02f, -0.58f, 0.33f, 0.092f);
            const R contact = sat(logA + logB + h3 * 0.35f);
            const R radius = lerp(0.16f, 0.46f, std::sqrt(h1));
            startX = std::cos(angle) * radius * (0.64f + contact * 0.36f);
            startZ = std::sin(angle) * radius * (0.46f + contact * 0.26f);
        } else {
            const R ring = std::sqrt(h1);
            startX = std::cos(angle) * 0.48f * ring;
            startZ = std::sin(angle) * 0.48f * ring;
        }
        const R su = sat(startX / 2.10f + 0.50f), sw = sat(startZ / 1.64f + 0.50f);
        const R sv = sid == 2 ? p.emitterHeightNorm : (sid == 1 ? 0.035f + h2 * 0.055f : 0.045f);
        const R heat = sample_scalar(st, st.s[HEAT].data(), su, sv, sw);
        const R charMass = sample_scalar(st, st.s[CHARF].data(), su, sv, sw);
        const R pyro = sample_scalar(st, st.s[PYRO].data(), su, sv, sw);
        const R turb = sample_scalar(st, st.s[TURB].data(), su, sv, sw);
        const V3 vel = sample_velocity(st, st.u.data(), st.v.data(), st.w.data(), su, sv, sw);
        const R materialGate = smoothstep(0.018f, 0.40f, charMass + pyro * 0.34f + heat * 0.075f);
        if (materialGate <= 0.0001f) continue;
        const R tempK = 620.0f + heat * 330.0f + pyro * 210.0f + charMass * 42.0f;
        const R lift = sid == 1 ? 0.86f : (sid == 2 ? 0.16f : 0.80f);
        const R t = life * (sid == 1 ? 1.24f : (sid == 2 ? 0.34f : 1.10f));
        const R dragK = sid == 2 ? 4.8f : 1.35f;
        const R rise = 0.26f + h3 * 0.62f + turb * 0.11f;
        const R fall = 0.22f + h2 * 0.16f;
        const R driftX = vel.x * 0.34f + (h1 - 0.5f) * 0.20f + p.windX * 0.26f;
        const R driftZ = vel.z * 0.28f + (h2 - 0.5f) * 0.18f;
        const R drag = std::exp(-t * dragK);
        const V3 world = v3(startX + driftX * t * drag, sv * 2.03f + 0.02f + rise * t * lift - fall * t * t, startZ + driftZ * t * drag);
        const R pt = std::max<R>(0.0, t - 0.045f);
        const R pdrag = std::exp(-pt * dragK);
        const V3 prev = v3(startX + driftX * pt * pdrag, sv * 2.03f + 0.02f + rise * pt * lift - fall * pt * pt, startZ + driftZ * pt * pdrag);
        const V3 sc = project_point(cam, world), ps = project_point(cam, prev);
        if (sc.z <= 0.0 || sc.x < -0.05f || sc.x > 1.05f || sc.y < -0.05f || sc.y > 1.05f) continue;
        const R size = smoothstep(720.0f, 1580.0f, tempK) * materialGate;
        const R radius = ((sid == 2 ? 0.00042f : 0.00086f) + h3 * (sid == 2 ? 0.00042f : 0.00145f)) * (0.72f + size * 0.90f) /
                         std::max<R>(0.55f, sc.z);
        const R cxp = sc.x * st.fw - 0.5f, cyp = sc.y * st.fh - 0.5f;
        const R trx = sc.x - ps.x, trY = sc.y - ps.y;
        const R trailLen = std::sqrt(trx * trx + trY * trY);
        const R capped = std::min<R>(trailLen, sid == 1 ? 0.016f : 0.005f);
        const bool hasDir = trailLen > 0.000001f;
        const R tdx = hasDir ? trx / trailLen : 0.0, tdy = hasDir ? trY / trailLen : -1.0;
        const int pad = std::max(2, static_cast<int>(std::ceil((radius + capped) * std::max(st.fw, st.fh) * 3.0f)));
        const int x0 = std::max(0, static_cast<int>(std::floor(cxp)) - pad), x1 = std::min(st.fw - 1, static_cast<int>(std::ceil(cxp)) + pad);
        const int y0 = std::max(0, static_cast<int>(std::floor(cyp)) - pad), y1 = std::min(st.fh - 1, static_cast<int>(std::ceil(cyp)) + pad);
        const V3 hot = sid == 2 ? v3(0.36f, 0.58f, 1.04f) : v3(1.0, 0.42f, 0.090f);
        const V3 cool = sid == 2 ? v3(0.055f, 0.10f, 0.18f) : v3(0.18f, 0.026f, 0.008f);
        const R cooling = smoothstep(0.16f, 0.92f, life) * (1.0 - sat(heat * 0.045f));
        const V3 color = lerp3(hot, cool, cooling);
        const R fade = smoothstep(1.0, 0.08f, life) * smoothstep(0.0, 0.16f, life) * sparkGate * materialGate * size *
                       (sid == 2 ? 0.055f : (sid == 1 ? 0.58f : 0.82f));
        const R streak = radius * (1.0 + capped * 130.0f);
        for (int y = y0; y <= y1; ++y) {
            const R uy = (y + 0.5) / st.fh;
            for (int x = x0; x <= x1; ++x) {
                const R ux = (x + 0.5) / st.fw;
                const R dx = ux - sc.x, dy = uy - sc.y;
                const R along = dx * tdx + dy * tdy;
                const R ax = dx - tdx * along, ay = dy - tdy * along;
                const R core = std::exp(-((ax * ax + ay * ay) / std::max<R>(0.0000002f, radius * radius) +
                                          along * along / std::max<R>(0.0000002f, streak * streak)));
                const R halo = std::exp(-(dx * dx + dy * dy) / std::max<R>(0.0000002f, radius * radius * 8.0f)) * 0.16f;
                const R spark = (core + halo) * fade;
                if (spark > 0.00001f) {
                    R* o = out + 4 * (size_t(y) * st.fw + x);
                    o[0] += color.x * spark;
                    o[1] += color.y * spark;
                    o[2] += color.z * spark;
                }
            }
        }
    }
}

static void render_hdr(const State& st, R* out) {
    const P p = make_p(st, st.renderTime);
#pragma omp parallel for schedule(dynamic, 4)
    for (int y = 0; y < st.fh; ++y) {
        for (int x = 0; x < st.fw; ++x) {
            const V4 c = render_pixel(st, p, x, y);
            R* o = out + 4 * (size_t(y) * st.fw + x);
            o[0] = c.x;
            o[1] = c.y;
            o[2] = c.z;
            o[3] = c.w;
        }
    }
    splat_embers(st, p, out);
}

// ------------------------------------------------------------- tonemap ----
// src/fire_cuda.cu:2246-2274
static inline R aces_curve(R v) {
    const R a = 2.51f, b = 0.03f, c = 2.43f, d = 0.59f, e = 0.14f;
    return sat((v * (a * v + b)) / (v * (c * v + d) + e));
}
static V3 camera_response(V3 radiance, R exposure) {
    V3 c = mul3(v3(std::max<R>(0.0, radiance.x), std::max<R>(0.0, radiance.y), std::max<R>(0.0, radiance.z)), exposure);
    const R l0 = std::max<R>(0.000001f, luminance3(c));
    c = mul3(c, aces_curve(l0) / l0);
    const R luma = luminance3(c);
    const R toe = smoothstep(0.000f, 0.055f, luma);
    const R shoulder = smoothstep(0.78f, 1.08f, luma);
    c = lerp3(mul3(c, 0.82f), c, toe);
    c = lerp3(c, mul3(c, 1.0 / std::max<R>(1.0, luma / 0.94f)), shoulder * 0.10f);
    const R lift = smoothstep(0.56f, 0.90f, luma);
    c = mul3(c, 1.0f + lift * 0.28f);
    const R g = 1.0f / 2.2f;
    return v3(std::pow(sat(c.x), g), std::pow(sat(c.y), g), std::pow(sat(c.z), g));
}

// tonemapKernel src/fire_cuda.cu:2276-2297 (dither from the pixel hash).
// drawGizmos src/fire_cuda.cu:1909-1951 over the unpacked 8-bit colour
// (overlayKernel :2468-2494, after the tonemap :3419-3425).
static inline R segment_alpha(R px, R py, R ax, R ay, R bx, R by, R thickness) {
    const R pax = px - ax, pay = py - ay, bax = bx - ax, bay = by - ay;
    const R h = sat((pax * bax + pay * bay) / std::max<R>(0.000001f, bax * bax + bay * bay));
    const R dx = pax - bax * h, dy = pay - bay * h;
    return smoothstep(thickness, thickness * 0.35f, std::sqrt(dx * dx + dy * dy));
}
static void draw_gizmos(V3& color, R uvx, R uvy, const P& p) {
    const Cam& cam = p.cam;
    const bool burner = p.sceneId == 2 && p.burnerCount > 0;
    const R sx = burner ? p.burnerX[0] : p.emitterCenterX;
    const R sy = burner ? p.burnerY[0] : p.emitterHeightNorm * 2.03f + 0.02f;
    const R sz = burner ? p.burnerZ[0] : p.emitterCenterZ;
    const V3 c = project_point(cam, v3(sx, sy, sz));
    if (c.z > 0.0) {
        const R dist = std::sqrt((uvx - c.x) * (uvx - c.x) + (uvy - c.y) * (uvy - c.y));
        const R a = smoothstep(0.006f, 0.006f * 0.25f, std::fabs(dist - 0.125f / c.z)) + smoothstep(0.012f, 0.012f * 0.45f, dist);
        const V3 tint = p.activeGizmo == 1 ? v3(1.0f, 0.34f, 0.055f) : v3(0.45f, 0.18f, 0.055f);
        color = lerp3(color, tint, sat(a * 0.85f));
    }
    const V3 ex = v3(1, 0, 0), ey = v3(0, 1, 0), ez = v3(0, 0, 1);
    const R bx = 0.90f, by = 0.84f;
    const V3 axes[3] = {ex, ey, ez};
    const V3 tints[3] = {v3(1.0f, 0.12f, 0.08f), v3(0.22f, 1.0f, 0.30f), v3(0.20f, 0.48f, 1.0f)};
    for (int k = 0; k < 3; ++k) {
        const R tx = bx + dot3(axes[k], cam.right) * 0.070f, ty = by - dot3(axes[k], cam.up) * 0.070f;
        color = lerp3(color, tints[k], sat(segment_alpha(uvx, uvy, bx, by, tx, ty, 0.005f)));
    }
}

static void tonemap(const State& st, const R* hdr, uint32_t* out) {
    const P p = make_p(st, st.renderTime);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < st.fh; ++y) {
        for (int x = 0; x < st.fw; ++x) {
            const size_t i = size_t(y) * st.fw + x;
            V3 c = camera_response(v3(hdr[4 * i], hdr[4 * i + 1], hdr[4 * i + 2]), p.exposure);
            const R dither = (hash2(x, y, p.saltDither) - 0.5f) / 255.0f;
            c = add3(c, v3(dither, dither, dither));
            uint32_t r = static_cast<uint32_t>(sat(c.x) * 255.0f);
            uint32_t g = static_cast<uint32_t>(sat(c.y) * 255.0f);
            uint32_t b = static_cast<uint32_t>(sat(c.z) * 255.0f);
            if (p.showGizmos && p.debugMode == 0) {
                V3 g8 = v3(r / 255.0f, g / 255.0f, b / 255.0f);
                draw_gizmos(g8, (x + 0.5) / st.fw, (y + 0.5) / st.fh, p);
                r = static_cast<uint32_t>(sat(g8.x) * 255.0f);
                g = static_cast<uint32_t>(sat(g8.y) * 255.0f);
                b = static_cast<uint32_t>(sat(g8.z) * 255.0f);
            }
            out[i] = 0xff000000u | (r << 16) | (g << 8) | b;
        }
    }
}

// -------------------------------------------------------------- metrics ----
// divergenceOnlyKernel :1247-1259 + accumulateMetricsKernel :1319-1418 + host
// :3495-3544, in FP64; plus heat release and the SOR residual.
static void metrics(const State& st, fs_metrics& m) {
    std::memset(&m, 0, sizeof(m));
    const P p = make_p(st, st.time);
    m.frame_index = st.frameIndex;
    m.grid_x = st.nx;
    m.grid_y = st.ny;
    m.grid_z = st.nz;
    m.pressure_iterations = p.iterations;
    m.time_seconds = st.time;
    const R* F[10];
    for (int f = 0; f < 10; ++f) F[f] = st.s[f].data();
    double acc[24] = {0};
    double mx[9] = {0};
    long invalid = 0;
    for (int z = 0; z < st.nz; ++z) {
        for (int y = 0; y < st.ny; ++y) {
            for (int x = 0; x < st.nx; ++x) {
                const size_t idx = st.si(x, y, z);
                const R heat = F[HEAT][idx], fuel = F[FUEL][idx], oxygen = F[OXYGEN][idx], soot = F[SOOT][idx];
                const R ch = F[CHARF][idx], ash = F[ASH][idx], pyro = F[PYRO][idx], prog = F[PROGRESS][idx];
                const R turb = F[TURB][idx], so = F[SOPT][idx];
                const R divBefore = std::fabs(st.d[idx]);
                const R div = st.u[st.ui(x + 1, y, z)] - st.u[st.ui(x, y, z)] + st.v[st.vi(x, y + 1, z)] - st.v[st.vi(x, y, z)] +
                              st.w[st.wi(x, y, z + 1)] - st.w[st.wi(x, y, z)];
                const R divAfter = std::fabs(-div);
                const R pc = st.p[idx];
                const R resid = st.d[idx] + st.p[st.sic(x - 1, y, z)] + st.p[st.sic(x + 1, y, z)] + st.p[st.sic(x, y - 1, z)] +
                                st.p[st.sic(x, y + 1, z)] + st.p[st.sic(x, y, z - 1)] + st.p[st.sic(x, y, z + 1)] - 6.0 * pc;
                const R u = (x + 0.5) / st.nx, v = (y + 0.5) / st.ny, w = (z + 0.5) / st.nz;
                const V4 L = sample_light(st, u, v, w);
                const R shadow = sample_shadow(st, u, v, w);
                const R vals[12] = {heat, fuel, oxygen, soot, ch, ash, pyro, prog, turb, so, divBefore, divAfter};
                bool bad = false;
                for (R q : vals) bad = bad || !std::isfinite(q);
                if (bad) {
                    ++invalid;
                    continue;
                }
                const R tempK = 293.0f + heat * 335.0f;
                const R hrr = (fuel + pyro * 0.42f) * oxygen * smoothstep(610.0f, 1280.0f, tempK) * std::exp(-5200.0f / std::max<R>(650.0f, tempK));
                const R od = so + soot * (0.24f + 0.05f * std::sqrt(std::max<R>(0.0, heat)));
                const R nh = (y + 0.5) / st.ny;
                const R flameMass = smoothstep(0.72f, 2.10f, heat + prog * 0.62f + pyro * 0.24f) * smoothstep(0.035f, 0.66f, oxygen) *
                                    smoothstep(0.012f, 0.52f, fuel + pyro * 0.36f);
                const R smokeMass = smoothstep(0.035f, 1.16f, so + soot * 0.38f + ash * 0.12f) * smoothstep(0.08f, 0.88f, nh);
                const R volShadow = shadow * std::exp(-L.w * 0.80f);
                const R lum = luminance3(v3(L.x, L.y, L.z));
                const R roomIrr = lum * volShadow;
                const R marker = heat + fuel * 0.24f + prog * 0.55f;
                const R add[22] = {heat, fuel, oxygen, soot, ch, ash, pyro, prog, turb, so, lum, shadow, flameMass, smokeMass,
                                   flameMass * smokeMass, volShadow, roomIrr, od, hrr, divBefore * divBefore, divAfter * divAfter, resid * resid};
                for (int i = 0; i < 22; ++i) acc[i] += add[i];
                const R mxv[8] = {heat, fuel, soot, pyro, prog, turb, divBefore, divAfter};
                for (int i = 0; i < 8; ++i)
                    if (mxv[i] > 0.0) mx[i] = std::max(mx[i], mxv[i]);
                if (marker > 0.74f) {
                    const R h = nh * 2.03f;
                    if (h > 0.0) mx[8] = std::max(mx[8], h);
                }
            }
        }
    }
    const double cells = std::max<double>(1.0, double(st.cells()));
    m.invalid_cells = static_cast<int32_t>(invalid);
    m.heat_sum = acc[0]; m.fuel_sum = acc[1]; m.oxygen_sum = acc[2]; m.soot_sum = acc[3];
    m.char_sum = acc[4]; m.ash_sum = acc[5]; m.pyrolysis_sum = acc[6]; m.progress_sum = acc[7];
    m.turbulence_energy_sum = acc[8]; m.soot_optical_depth_sum = acc[9];
    m.mean_scene_light = acc[10] / cells;
    m.mean_scene_shadow = acc[11] / cells;
    m.flame_mass_proxy = acc[12];
    m.smoke_mass_proxy = acc[13];
    m.flame_smoke_overlap_proxy = acc[14];
    m.mean_volumetric_shadow = acc[15] / cells;
    m.mean_room_irradiance = acc[16] / cells;
    m.mean_optical_depth = acc[17] / cells;
    m.heat_release_proxy = acc[18];
    m.divergence_before_l2 = std::sqrt(acc[19] / cells);
    m.divergence_after_l2 = std::sqrt(acc[20] / cells);
    m.sor_residual_l2 = std::sqrt(acc[21] / cells);
    m.max_heat = mx[0]; m.max_fuel = mx[1]; m.max_soot = mx[2]; m.max_pyrolysis = mx[3];
    m.max_progress = mx[4]; m.max_turbulence_energy = mx[5];
    m.divergence_before_max = mx[6];
    m.divergence_after_max = mx[7];
    m.flame_height_meters = mx[8];
    m.divergence_reduction = m.divergence_before_l2 > 0.000001 ? 1.0 - m.divergence_after_l2 / m.divergence_before_l2 : 0.0;
    m.heat_release_step = st.heatReleaseStep;
    m.heat_release_total = st.heatReleaseTotal;
}

// ---------------------------------------------------------------- step ----
static void run_stage(State& st, int stage) {
    const P p = make_p(st, st.time);
    switch (stage) {
        case FS_STAGE_ADVECT_VELOCITY: advect_velocity(st, p); break;
        case FS_STAGE_ADVECT_REACT: advect_react(st, p); break;
        case FS_STAGE_FORCE: force_boundary_divergence(st, p); break;
        case FS_STAGE_PRESSURE: pressure(st, p); break;
        case FS_STAGE_PROJECT: project(st); break;
        case FS_STAGE_LIGHT: light_pass(st, p); break;
        default: break;
    }
}

// One step: time bookkeeping in FP32 (src/fire_cuda.cu:3135-3142), then the
// passes in the order of src/fire_cuda.cu:3196-3318.
static void step(State& st) {
    const float dt = st.prm.dt;
    st.renderTime += dt;
    st.time += dt;
    for (int s = 0; s < FS_STAGE_COUNT; ++s) run_stage(st, s);
    ++st.frameIndex;
}

static int light_dim(int n, int lo, int hi) { return std::max(lo, std::min(hi, (n + 3) / 4)); }

}  // namespace

// ================================================================ C API ====
extern "C" {

void* or_create(int nx, int ny, int nz, int fw, int fh) {
    if (nx < 2 || ny < 2 || nz < 2 || fw < 1 || fh < 1) return nullptr;
    State* st = new State();
    st->nx = nx; st->ny = ny; st->nz = nz;
    st->fw = fw; st->fh = fh;
    st->lnx = light_dim(nx, 18, 48);
    st->lny = light_dim(ny, 20, 56);
    st->lnz = light_dim(nz, 16, 36);
    for (auto& f : st->s) f.assign(st->cells(), 0.0);
    st->u.assign(st->ucount(), 0.0);
    st->v.assign(st->vcount(), 0.0);
    st->w.assign(st->wcount(), 0.0);
    st->p.assign(st->cells(), 0.0);
    st->d.assign(st->cells(), 0.0);
    st->light.assign(4 * st->lcells(), 0.0);
    st->shadow.assign(st->lcells(), 0.0);
    st->dirT.assign(st->lcells(), 0.0);
    return st;
}

void or_destroy(void* h) { delete static_cast<State*>(h); }

int or_clamp_params(const fs_params* in, fs_params* out) {
    if (!in || !out) return 1;
    fs_params tmp;
    clamp_params(*in, tmp);
    *out = tmp;
    return 0;
}

void or_set_params(void* h, const fs_params* prm) {
    State* st = static_cast<State*>(h);
    clamp_params(*prm, st->prm);
}
void or_get_params(void* h, fs_params* out) { *out = static_cast<State*>(h)->prm; }

void or_set_emitters(void* h, const fs_emitter* e, int n) {
    State* st = static_cast<State*>(h);
    st->emitters.assign(e, e + std::max(0, std::min(n, FS_MAX_EMITTERS)));
}

void or_reset(void* h) { reset(*static_cast<State*>(h)); }

void or_set_time(void* h, float t, float rt) {
    State* st = static_cast<State*>(h);
    st->time = t;
    st->renderTime = rt;
}
void or_get_time(void* h, float* t, float* rt, int* frame) {
    State* st = static_cast<State*>(h);
    *t = st->time;
    *rt = st->renderTime;
    *frame = st->frameIndex;
}

void or_light_dims(void* h, int* lnx, int* lny, int* lnz) {
    State* st = static_cast<State*>(h);
    *lnx = st->lnx; *lny = st->lny; *lnz = st->lnz;
}

static std::vector<R>* field_ptr(State* st, int f) {
    if (f >= 0 && f < 10) return &st->s[f];
    switch (f) {
        case FS_FIELD_U: return &st->u;
        case FS_FIELD_V: return &st->v;
        case FS_FIELD_W: return &st->w;
        case FS_FIELD_PRESSURE: return &st->p;
        case FS_FIELD_DIVERGENCE: return &st->d;
        case FS_FIELD_LIGHT: return &st->light;
        case FS_FIELD_SHADOW: return &st->shadow;
        case FS_FIELD_DIR_TRANSMITTANCE: return &st->dirT;
        default: return nullptr;
    }
}

long or_field_count(void* h, int f) {
    std::vector<R>* v = field_ptr(static_cast<State*>(h), f);
    return v ? static_cast<long>(v->size()) : -1;
}
int or_set_field(void* h, int f, const double* src) {
    std::vector<R>* v = field_ptr(static_cast<State*>(h), f);
    if (!v) return 1;
    std::copy(src, src + v->size(), v->begin());
    return 0;
}
int or_get_field(void* h, int f, double* dst) {
    std::vector<R>* v = field_ptr(static_cast<State*>(h), f);
    if (!v) return 1;
    std::copy(v->begin(), v->end(), dst);
    return 0;
}

void or_run_stage(void* h, int stage) { run_stage(*static_cast<State*>(h), stage); }
void or_step(void* h, int n) {
    State* st = static_cast<State*>(h);
    for (int i = 0; i < n; ++i) step(*st);
}
void or_render_hdr(void* h, double* out) { render_hdr(*static_cast<State*>(h), out); }
void or_tonemap(void* h, const double* hdr, uint32_t* out) { tonemap(*static_cast<State*>(h), hdr, out); }
void or_metrics(void* h, fs_metrics* m) { metrics(*static_cast<State*>(h), *m); }

int or_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void or_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// ---- primitives exposed for the pinning tests ----
double or_sample_grid(const double* f, int dx, int dy, int dz, double sx, double sy, double sz, double u, double v, double w) {
    return sample_grid(f, dx, dy, dz, sx, sy, sz, u, v, w);
}
double or_smoothstep(double a, double b, double x) { return smoothstep(a, b, x); }
uint32_t or_pcg(uint32_t v) { return pcg(v); }
double or_hash2(int x, int y, uint32_t salt) { return hash2(x, y, salt); }
double or_hash3(int x, int y, int z, uint32_t salt) { return hash3(x, y, z, salt); }
double or_value_noise2(double x, double y, uint32_t salt) { return value_noise2(x, y, salt); }
double or_value_noise3(double x, double y, double z, uint32_t salt) { return value_noise3(x, y, z, salt); }
double or_fbm2(double x, double y, uint32_t salt) { return fbm2(x, y, salt); }
double or_fbm3(double x, double y, double z, uint32_t salt) { return fbm3(x, y, z, salt); }
uint32_t or_salt(uint32_t seed, uint32_t k) { return salt_of(seed, k); }
void or_blackbody(double tempK, double* rgb) {
    const V3 c = blackbody(tempK);
    rgb[0] = c.x; rgb[1] = c.y; rgb[2] = c.z;
}
double or_aces(double v) { return aces_curve(v); }
void or_camera_response(const double* rgb, double exposure, double* out) {
    const V3 c = camera_response(v3(rgb[0], rgb[1], rgb[2]), exposure);
    out[0] = c.x; out[1] = c.y; out[2] = c.z;
}
double or_domain_fade(double u, double v, double w) { return domain_fade(u, v, w); }

// MacCormack sample of a (nz, ny, nx) scalar array f at backtrace (pu, pv, pw)
// with the previous-position velocity pvel (3) and timestep dt.
double or_maccormack(const double* f, int nx, int ny, int nz, double dt, double current, const double* pvel, double pu, double pv, double pw,
                     double blend) {
    State st;
    st.nx = nx; st.ny = ny; st.nz = nz;
    P p;
    p.dt = dt;
    return maccormack(st, p, f, current, v3(pvel[0], pvel[1], pvel[2]), pu, pv, pw, blend);
}
double or_capsule_mask(double x, double z, double cx, double cz, double angle, double halfLength, double radius) {
    return capsule_mask(x, z, cx, cz, angle, halfLength, radius);
}
// Bed material and source of the scene's current parameters at world (x, z).
double or_bed_material(void* h, double x, double z) {
    const State& st = *static_cast<State*>(h);
    return bed_material(x, z, make_p(st, st.time));
}
double or_bed_source(void* h, double x, double z, double height, double time) {
    const State& st = *static_cast<State*>(h);
    return bed_source(x, z, height, make_p(st, time));
}
// The light build alone (light <- radiance + extinction, shadow), at the scene time.
void or_light_build(void* h) {
    State& st = *static_cast<State*>(h);
    std::vector<R> L(4 * st.lcells());
    light_build(st, make_p(st, st.time), L);
    st.light.swap(L);
}
// One light propagation of the current light volume, in place.
void or_light_propagate(void* h, double diffusion, double decay) {
    State& st = *static_cast<State*>(h);
    std::vector<R> out(4 * st.lcells());
    light_propagate(st, st.light, out, diffusion, decay);
    st.light.swap(out);
}
// Directional-light transmittance of the current light volume (any dir_light).
void or_dir_transmittance(void* h) {
    State& st = *static_cast<State*>(h);
    dir_transmittance(st, make_p(st, st.time));
}
// Room colour along ray rd from the scene camera, for pixel uv (render time clock).
void or_room_background(void* h, double uvx, double uvy, const double* rd, double glow, double* rgb) {
    const State& st = *static_cast<State*>(h);
    const V3 c = room_background(st, make_p(st, st.renderTime), uvx, uvy, v3(rd[0], rd[1], rd[2]), glow);
    rgb[0] = c.x; rgb[1] = c.y; rgb[2] = c.z;
}
// The scene camera's eye position (the origin of or_room_background rays).
void or_camera_eye(void* h, double* eye) {
    const State& st = *static_cast<State*>(h);
    const Cam c = make_camera(st.prm, st.fw, st.fh);
    eye[0] = c.eye.x; eye[1] = c.eye.y; eye[2] = c.eye.z;
}

}  // extern "C"
