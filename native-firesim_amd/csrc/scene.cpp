// scene.cpp - per-scene handles and the C ABI of include/firesim.h.
//
// The reference keeps one simulation in process globals on the legacy stream
// (src/fire_cuda.cu:116-195, orchestration :3106-3549). Here every scene owns
// one device allocation (all fields, light volume, frame, reduction scratch)
// and one non-blocking HIP stream; steps are enqueued asynchronously and only
// the read/render/metrics calls synchronise.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "scene_impl.hpp"

namespace {

thread_local std::string g_globalError = "no error";

int global_fail(int status, const std::string& msg) {
    g_globalError = msg;
    return status;
}

float clampf(float v, float lo, float hi) { return std::max(lo, std::min(hi, v)); }
int32_t clampi(int32_t v, int32_t lo, int32_t hi) { return std::max(lo, std::min(hi, v)); }

// Reference clamps src/fire_cuda.cu:3135-3179 with the extensions documented
// on fs_params.
void clamp_params(const fs_params& in, fs_params& o) {
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
    o.ember_count = clampi(in.ember_count, 0, fs::kMaxEmbers);
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

uint32_t pcg_host(uint32_t v) {
    const uint32_t state = v * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

int light_dim(int n, int lo, int hi) { return std::max(lo, std::min(hi, (n + 3) / 4)); }

size_t align256(size_t n) { return (n + 255) & ~static_cast<size_t>(255); }

struct Layout {
    size_t scalarBytes[fs::NSCALAR];
    size_t total = 0;
};


// Byte size of each sub-buffer, in allocation order.
std::vector<size_t> buffer_sizes(const Dims& d) {
    std::vector<size_t> b;
    const size_t c = d.cells();
    for (int k = 0; k < 2; ++k)
        for (int i = 0; i < fs::NSCALAR; ++i) {
            const bool narrow = d.bf16 && (i == fs::HEAT || i == fs::FUEL || i == fs::SOOT);
            b.push_back(c * (narrow ? 2 : 4));
        }
    const size_t uc = static_cast<size_t>(d.nx + 1) * d.ny * d.nz;
    const size_t vc = static_cast<size_t>(d.nx) * (d.ny + 1) * d.nz;
    const size_t wc = static_cast<size_t>(d.nx) * d.ny * (d.nz + 1);
    for (int k = 0; k < 2; ++k) {
        b.push_back(uc * 4);
        b.push_back(vc * 4);
        b.push_back(wc * 4);
    }
    for (int k = 0; k < 6; ++k) b.push_back(c * 4);  // p, p2, d, cu, cv, cw
    const size_t edges = static_cast<size_t>(d.nx) * (d.ny + 1) * d.nz;
    for (int k = 0; k < 3; ++k) b.push_back(edges * 4);  // curl-noise potential
    b.push_back(d.lcells() * 16);
    b.push_back(d.lcells() * 16);
    b.push_back(d.lcells() * 4);
    b.push_back(d.lcells() * 4);  // directional-light transmittance
    b.push_back(d.pixels() * 16);
    b.push_back(d.pixels() * 4);
    b.push_back(d.pixels() * 8);
    b.push_back(d.partials() * 8);
    b.push_back(2 * 8);
    b.push_back(fs::kMetricSlots * 8);
    b.push_back(2 * 4);  // device clock
    b.push_back(fs::kMaxEmbers * sizeof(fs::Ember));
    return b;
}

size_t total_bytes(const Dims& d) {
    size_t t = 0;
    for (size_t s : buffer_sizes(d)) t += align256(s);
    return t;
}

bool valid_desc(const fs_scene_desc* d, std::string& why) {
    if (!d) {
        why = "scene descriptor is null";
        return false;
    }
    if (d->nx < 4 || d->ny < 4 || d->nz < 4) {
        why = "grid dimensions must each be >= 4";
        return false;
    }
    // Kernels index every field with 32-bit byte offsets: no array may hold
    // more than 2^30 values (the face arrays hold up to (nx+1)(ny+1)(nz+1)).
    const double values = static_cast<double>(d->nx + 1) * (d->ny + 1) * (d->nz + 1);
    if (values > fs::kMaxFieldValues) {
        why = "grid too large: (nx+1)(ny+1)(nz+1) must not exceed 2^30";
        return false;
    }
    if (d->frame_w < 1 || d->frame_h < 1 || static_cast<double>(d->frame_w) * d->frame_h > 1.0e8) {
        why = "frame dimensions must be positive and at most 1e8 pixels";
        return false;
    }
    if (d->storage != FS_STORAGE_FP32 && d->storage != FS_STORAGE_BF16) {
        why = "unknown storage mode";
        return false;
    }
    return true;
}

Dims dims_of(const fs_scene_desc& d) {
    Dims m;
    m.nx = d.nx; m.ny = d.ny; m.nz = d.nz;
    m.lnx = light_dim(d.nx, 18, 48);
    m.lny = light_dim(d.ny, 20, 56);
    m.lnz = light_dim(d.nz, 16, 36);
    m.fw = d.frame_w; m.fh = d.frame_h;
    m.bf16 = d.storage == FS_STORAGE_BF16;
    return m;
}

}  // namespace


namespace {

void camera_of(const fs_params& s, int fw, int fh, DevParams& p) {
    // computeCameraFromSettings src/fire_cuda.cu:315-329
    const float tx = 0.0f, ty = 0.72f, tz = 0.0f;
    const float cp = std::cos(s.camera_pitch);
    const float ex = std::sin(s.camera_yaw) * s.camera_distance * cp;
    const float ey = ty + std::sin(s.camera_pitch) * s.camera_distance;
    const float ez = std::cos(s.camera_yaw) * s.camera_distance * cp;
    auto norm = [](float& x, float& y, float& z) {
        const float len = std::sqrt(std::max(0.000001f, x * x + y * y + z * z));
        x /= len; y /= len; z /= len;
    };
    float fx = tx - ex, fy = ty - ey, fz = tz - ez;
    norm(fx, fy, fz);
    // right = normalize(cross(forward, (0,1,0)))
    float rx = fy * 0.0f - fz * 1.0f, ry = fz * 0.0f - fx * 0.0f, rz = fx * 1.0f - fy * 0.0f;
    norm(rx, ry, rz);
    // up = normalize(cross(right, forward))
    float ux = ry * fz - rz * fy, uy = rz * fx - rx * fz, uz = rx * fy - ry * fx;
    norm(ux, uy, uz);
    p.camEye[0] = ex; p.camEye[1] = ey; p.camEye[2] = ez;
    p.camForward[0] = fx; p.camForward[1] = fy; p.camForward[2] = fz;
    p.camRight[0] = rx; p.camRight[1] = ry; p.camRight[2] = rz;
    p.camUp[0] = ux; p.camUp[1] = uy; p.camUp[2] = uz;
    p.tanHalfFov = 0.50f;
    p.aspect = static_cast<float>(fw) / static_cast<float>(fh);
}

DevParams dev_params(const fs_scene& sc) {
    DevParams p;
    std::memset(&p, 0, sizeof(p));
    const fs_params& s = sc.prm;
    const Dims& d = sc.dims;
    p.nx = d.nx; p.ny = d.ny; p.nz = d.nz;
    p.invN[0] = 1.0f / static_cast<float>(d.nx);
    p.invN[1] = 1.0f / static_cast<float>(d.ny);
    p.invN[2] = 1.0f / static_cast<float>(d.nz);
    p.lnx = d.lnx; p.lny = d.lny; p.lnz = d.lnz;
    p.fw = d.fw; p.fh = d.fh;
    p.dt = s.dt;
    p.clock = sc.ds.clock;
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
    auto key = [&](uint32_t k) { return fs::NoiseKey{pcg_host(s.noise_seed + k * 0x9E3779B9u), s.noise_mode}; };
    p.saltNoise2 = key(1);
    p.saltNoise3 = key(2);
    p.saltJitter = key(3);
    p.saltDither = key(4);
    p.saltGrain = key(5);
    p.saltEmber = key(6);
    for (int c = 0; c < 3; ++c) p.saltCurl[c] = fs::NoiseKey{pcg_host(s.noise_seed + (7u + c) * 0x9E3779B9u), 0};
    p.emberCount = s.ember_count;
    p.closedTop = s.closed_top;
    p.curlOn = s.noise_forcing;
    p.curlNoise = s.curl_noise;
    p.showGizmos = s.show_gizmos;
    p.activeGizmo = s.active_gizmo;
    p.dirLight = s.dir_light;
    {
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
    camera_of(s, d.fw, d.fh, p);
    p.emitterCount = static_cast<int>(std::min<size_t>(sc.emitters.size(), fs::kMaxEmitters));
    for (int i = 0; i < p.emitterCount; ++i) {
        const fs_emitter& e = sc.emitters[i];
        p.emitter[i][0] = e.cx;
        p.emitter[i][1] = e.cy;
        p.emitter[i][2] = e.cz;
        p.emitter[i][3] = e.radius;
        p.emitter[i][4] = e.fuel_rate;
        p.emitter[i][5] = (e.temperature_k - 293.0f) / 335.0f;
    }
    return p;
}

int set_device(fs_scene* s) {
    const hipError_t e = hipSetDevice(s->device);
    return e == hipSuccess ? FS_OK : s->hip_fail("hipSetDevice", e);
}

// fused: the step's form, where the force launch also sums the react
// partials (a pass run alone, or timed alone, finishes its own work).
// inStep: the pass runs inside a whole step; with a tile SOR kernel the
// pressure pass then starts from p = 0 without reading p, so the force pass
// does not store it. A pass run alone keeps the reference's form: force
// stores p = 0 and pressure starts from p.
int run_stage(fs_scene* s, int stage, const DevParams& p, bool fused = false, bool inStep = false) {
    hipError_t e = hipSuccess;
    const char* what = "";
    const bool zeroGuess = inStep && s->ds.sorKernel != 0;
    switch (stage) {
        case FS_STAGE_ADVECT_VELOCITY: e = fs::launch_advect_velocity(s->ds, p, s->stream); what = "advect velocity"; break;
        case FS_STAGE_ADVECT_REACT: e = fs::launch_advect_react(s->ds, p, s->stream, !fused); what = "advect/react"; break;
        case FS_STAGE_FORCE: e = fs::launch_force(s->ds, p, s->stream, fused, !zeroGuess); what = "force/boundary/divergence"; break;
        case FS_STAGE_PRESSURE: e = fs::launch_pressure(s->ds, p, s->stream, zeroGuess); what = "pressure SOR"; break;
        case FS_STAGE_PROJECT: e = fs::launch_project(s->ds, p, s->stream); what = "pressure subtract"; break;
        case FS_STAGE_LIGHT: e = fs::launch_light(s->ds, p, s->stream); what = "scene lighting"; break;
        default: return s->fail(FS_ERR_INVALID_ARGUMENT, "unknown stage");
    }
    return e == hipSuccess ? FS_OK : s->hip_fail(what, e);
}

// One step's device work: advance the device clock, then the six passes (the
// clock advances inside the velocity-advection launch, which does not read it).
int enqueue_step(fs_scene* s, const DevParams& p) {
    static_assert(FS_STAGE_ADVECT_VELOCITY == 0, "velocity advection is the first pass");
    const hipError_t e = fs::launch_advect_velocity(s->ds, p, s->stream, true);
    if (e != hipSuccess) return s->hip_fail("advect velocity", e);
    static_assert(FS_STAGE_FORCE == FS_STAGE_ADVECT_REACT + 1, "force follows advect-react");
    for (int stage = FS_STAGE_ADVECT_VELOCITY + 1; stage < FS_STAGE_COUNT; ++stage)
        if (int st = run_stage(s, stage, p, true, true)) return st;
    return FS_OK;
}

// The graphs may have launches queued or running on the scene stream: they go
// to the graveyard and are destroyed once the stream has passed them.
void drop_graphs(fs_scene* s) {
    for (auto& g : s->exec) {
        s->graveyard.retire(g, s->stream);
        g = nullptr;
    }
    s->graphsValid = false;
    s->graveyard.reap(false);
}

bool same_buffers(const DeviceScene& a, const DeviceScene& b) {
    for (int i = 0; i < fs::NSCALAR; ++i)
        if (a.s[i] != b.s[i]) return false;
    return a.u == b.u && a.v == b.v && a.w == b.w && a.p == b.p && a.light == b.light;
}

// Captures the step from the current buffer assignment (graph 0) and from the
// assignment one step later (graph 1); two steps restore the assignment.
int capture_graphs(fs_scene* s) {
    drop_graphs(s);
    const DevParams p = dev_params(*s);
    const DeviceScene base = s->ds;
    for (int g = 0; g < 2; ++g) {
        s->state[g] = s->ds;
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            s->ds = base;
            return s->hip_fail("hipStreamBeginCapture", e);
        }
        const int st = enqueue_step(s, p);
        e = hipStreamEndCapture(s->stream, &graph);
        if (st != FS_OK || e != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            s->ds = base;
            drop_graphs(s);
            return st != FS_OK ? st : s->hip_fail("hipStreamEndCapture", e);
        }
        e = hipGraphInstantiate(&s->exec[g], graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            s->ds = base;
            drop_graphs(s);
            return s->hip_fail("hipGraphInstantiate", e);
        }
    }
    if (!same_buffers(s->ds, base)) {
        s->ds = base;
        drop_graphs(s);
        return s->fail(FS_ERR_HIP, "step graph capture: two steps did not restore the buffer assignment");
    }
    s->ds = base;
    s->parity = 0;
    s->graphsValid = true;
    return FS_OK;
}

// Enqueues one step (graph replay or direct launches) and advances the host clock.
int one_step(fs_scene* s) {
    s->renderTime += s->prm.dt;
    s->time += s->prm.dt;
    int st = FS_OK;
    if (s->useGraph) {
        if (!s->graphsValid && (st = capture_graphs(s)) != FS_OK) return st;
        const hipError_t e = hipGraphLaunch(s->exec[s->parity], s->stream);
        if (e != hipSuccess) return s->hip_fail("hipGraphLaunch", e);
        s->ds = s->state[s->parity ^ 1];
        s->parity ^= 1;
    } else {
        st = enqueue_step(s, dev_params(*s));
    }
    ++s->frameIndex;
    return st;
}

// Element count of a field, and whether it is a scalar.
bool field_info(const fs_scene* s, int f, size_t& count) {
    const Dims& d = s->dims;
    if (f >= 0 && f < fs::NSCALAR) { count = d.cells(); return true; }
    switch (f) {
        case FS_FIELD_U: count = static_cast<size_t>(d.nx + 1) * d.ny * d.nz; return true;
        case FS_FIELD_V: count = static_cast<size_t>(d.nx) * (d.ny + 1) * d.nz; return true;
        case FS_FIELD_W: count = static_cast<size_t>(d.nx) * d.ny * (d.nz + 1); return true;
        case FS_FIELD_PRESSURE:
        case FS_FIELD_DIVERGENCE: count = d.cells(); return true;
        case FS_FIELD_LIGHT: count = 4 * d.lcells(); return true;
        case FS_FIELD_SHADOW:
        case FS_FIELD_DIR_TRANSMITTANCE: count = d.lcells(); return true;
        default: return false;
    }
}

void* field_ptr(fs_scene* s, int f) {
    if (f >= 0 && f < fs::NSCALAR) return s->ds.s[f];
    switch (f) {
        case FS_FIELD_U: return s->ds.u;
        case FS_FIELD_V: return s->ds.v;
        case FS_FIELD_W: return s->ds.w;
        case FS_FIELD_PRESSURE: return s->ds.p;
        case FS_FIELD_DIVERGENCE: return s->ds.d;
        case FS_FIELD_LIGHT: return s->ds.light;
        case FS_FIELD_SHADOW: return s->ds.shadow;
        case FS_FIELD_DIR_TRANSMITTANCE: return s->ds.dirT;
        default: return nullptr;
    }
}

bool narrow_field(const fs_scene* s, int f) {
    return s->dims.bf16 && (f == FS_FIELD_HEAT || f == FS_FIELD_FUEL || f == FS_FIELD_SOOT);
}

}  // namespace



// ====================================================================== ABI
extern "C" {

int fs_abi_version(void) { return FS_ABI_VERSION; }

int fs_copy_bandwidth(int32_t device, uint64_t bytes, int32_t reps, double* gbps) {
    if (!gbps) return global_fail(FS_ERR_INVALID_ARGUMENT, "null output");
    *gbps = 0.0;
    if (bytes < (1ull << 20) || bytes > (64ull << 30) || bytes % 16 != 0 || reps < 1 || reps > 1000)
        return global_fail(FS_ERR_INVALID_ARGUMENT, "bytes must be a multiple of 16 in [1 MiB, 64 GiB] and reps in [1, 1000]");
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) return global_fail(FS_ERR_HIP, "no HIP device");
    if (device < 0 || device >= count) return global_fail(FS_ERR_INVALID_ARGUMENT, "device ordinal out of range");
    e = hipSetDevice(device);
    void *a = nullptr, *b = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (e == hipSuccess) e = hipMalloc(&a, bytes);
    if (e == hipSuccess) e = hipMalloc(&b, bytes);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemsetAsync(a, 0, bytes, st);
    if (e == hipSuccess) e = fs::launch_copy4(a, b, bytes, st);  // warm
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    for (int r = 0; e == hipSuccess && r < reps; ++r) e = fs::launch_copy4(a, b, bytes, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (st) (void)hipStreamSynchronize(st);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (st) (void)hipStreamDestroy(st);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return global_fail(e == hipErrorOutOfMemory ? FS_ERR_OUT_OF_MEMORY : FS_ERR_HIP, std::string("copy bandwidth: ") + hipGetErrorString(e));
    }
    if (ms <= 0.0f) return global_fail(FS_ERR_HIP, "copy bandwidth: no elapsed time");
    *gbps = 2.0 * static_cast<double>(bytes) * reps / (static_cast<double>(ms) * 1e-3) / 1e9;
    return FS_OK;
}

void fs_default_params(fs_params* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    // FireSettings defaults, src/fire_cuda.h:65-101
    o->dt = 1.0f / 60.0f;
    o->mouse_x = 0.5f;
    o->mouse_y = 0.12f;
    o->turbulence = 0.72f;
    o->detail = 0.88f;
    o->smoke = 0.92f;
    o->intensity = 1.0f;
    o->camera_pitch = 0.18f;
    o->camera_distance = 3.35f;
    o->raymarch_steps = 56;
    o->ember_count = 96;
    o->exposure = 1.0f;
    o->reflection_gain = 1.0f;
    o->smoke_darkness = 1.0f;
    o->emitter_height_band_norm = 0.06f;
    o->emitter_radius = 0.48f;
    // kPressureOmega / kPressureIterations, src/fire_cuda.cu:104-105
    o->sor_omega = 1.52f;
    o->sor_iterations = 40;
}

int fs_clamp_params(const fs_params* in, fs_params* out) {
    if (!in || !out) return global_fail(FS_ERR_INVALID_ARGUMENT, "null params");
    fs_params tmp;
    clamp_params(*in, tmp);
    *out = tmp;
    return FS_OK;
}

int fs_light_dims(int32_t nx, int32_t ny, int32_t nz, int32_t* lnx, int32_t* lny, int32_t* lnz) {
    if (!lnx || !lny || !lnz) return global_fail(FS_ERR_INVALID_ARGUMENT, "null output");
    *lnx = light_dim(nx, 18, 48);
    *lny = light_dim(ny, 20, 56);
    *lnz = light_dim(nz, 16, 36);
    return FS_OK;
}

int fs_scene_bytes(const fs_scene_desc* desc, uint64_t* bytes) {
    std::string why;
    if (!bytes) return global_fail(FS_ERR_INVALID_ARGUMENT, "null output");
    if (!valid_desc(desc, why)) return global_fail(FS_ERR_INVALID_ARGUMENT, why);
    *bytes = total_bytes(dims_of(*desc));
    return FS_OK;
}

const char* fs_last_global_error(void) { return g_globalError.c_str(); }

namespace {

// fs_create / fs_create_in: `memory` is null (the scene allocates its pool)
// or caller-owned device memory of at least total_bytes(dims) bytes.
int create_scene(const fs_scene_desc* desc, void* memory, uint64_t memoryBytes, fs_scene** out, bool reset = true) {
    std::string why;
    if (!out) return global_fail(FS_ERR_INVALID_ARGUMENT, "null output handle");
    *out = nullptr;
    if (!valid_desc(desc, why)) return global_fail(FS_ERR_INVALID_ARGUMENT, why);
    const Dims dims = dims_of(*desc);
    if (memory && (memoryBytes < total_bytes(dims) || (reinterpret_cast<uintptr_t>(memory) & 255u) != 0))
        return global_fail(FS_ERR_INVALID_ARGUMENT, "scene memory must be 256-byte aligned and at least fs_scene_bytes long");
    fs_scene* s = new (std::nothrow) fs_scene();
    if (!s) return global_fail(FS_ERR_OUT_OF_MEMORY, "host allocation failed");
    s->device = desc->device;
    s->dims = dims;
    hipError_t e = hipSetDevice(s->device);
    if (e != hipSuccess) {
        delete s;
        return global_fail(FS_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    }
    const std::vector<size_t> sizes = buffer_sizes(s->dims);
    s->poolBytes = total_bytes(s->dims);
    if (memory) {
        s->pool = memory;
        s->ownsPool = false;
    } else {
        e = hipMalloc(&s->pool, s->poolBytes);
        if (e != hipSuccess) {
            delete s;
            (void)hipGetLastError();
            return global_fail(e == hipErrorOutOfMemory ? FS_ERR_OUT_OF_MEMORY : FS_ERR_HIP,
                               std::string("hipMalloc scene: ") + hipGetErrorString(e));
        }
    }
    e = hipStreamCreateWithFlags(&s->ownStream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        if (s->ownsPool) (void)hipFree(s->pool);
        delete s;
        return global_fail(FS_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    s->stream = s->ownStream;
    char* base = static_cast<char*>(s->pool);
    size_t off = 0, k = 0;
    auto take = [&]() {
        void* ptr = base + off;
        off += align256(sizes[k++]);
        return ptr;
    };
    DeviceScene& ds = s->ds;
    ds.nx = s->dims.nx; ds.ny = s->dims.ny; ds.nz = s->dims.nz;
    ds.lnx = s->dims.lnx; ds.lny = s->dims.lny; ds.lnz = s->dims.lnz;
    ds.fw = s->dims.fw; ds.fh = s->dims.fh;
    ds.bf16 = s->dims.bf16;
    for (int i = 0; i < fs::NSCALAR; ++i) ds.s[i] = take();
    for (int i = 0; i < fs::NSCALAR; ++i) ds.sNext[i] = take();
    ds.u = static_cast<float*>(take()); ds.v = static_cast<float*>(take()); ds.w = static_cast<float*>(take());
    ds.uN = static_cast<float*>(take()); ds.vN = static_cast<float*>(take()); ds.wN = static_cast<float*>(take());
    ds.p = static_cast<float*>(take()); ds.p2 = static_cast<float*>(take()); ds.d = static_cast<float*>(take());
    ds.cu = static_cast<float*>(take()); ds.cv = static_cast<float*>(take()); ds.cw = static_cast<float*>(take());
    for (int k = 0; k < 3; ++k) ds.psi[k] = static_cast<float*>(take());
    ds.light = static_cast<float4*>(take()); ds.lightNext = static_cast<float4*>(take());
    ds.shadow = static_cast<float*>(take());
    ds.dirT = static_cast<float*>(take());
    ds.hdr = static_cast<float4*>(take());
    ds.bgra = static_cast<uint32_t*>(take());
    ds.rgba16 = static_cast<uint16_t*>(take());
    ds.partials = static_cast<double*>(take());
    ds.partialCapacity = s->dims.partials();
    ds.acc = static_cast<double*>(take());
    ds.metricsOut = static_cast<double*>(take());
    ds.clock = static_cast<float*>(take());
    ds.embers = static_cast<fs::Ember*>(take());
    fs_default_params(&s->prm);
    const int st = reset ? fs_reset(s) : FS_OK;
    if (st != FS_OK) {
        g_globalError = s->lastError;
        fs_destroy(s);
        return st;
    }
    *out = s;
    return FS_OK;
}

}  // namespace

int fs_create(const fs_scene_desc* desc, fs_scene** out) { return create_scene(desc, nullptr, 0, out); }

int fs_create_in(const fs_scene_desc* desc, void* device_memory, uint64_t bytes, fs_scene** out) {
    if (!device_memory) return global_fail(FS_ERR_INVALID_ARGUMENT, "null scene memory");
    return create_scene(desc, device_memory, bytes, out);
}

int fs_destroy(fs_scene* s) {
    if (!s) return FS_OK;
    (void)hipSetDevice(s->device);
    if (s->batch) (void)fs_batch_remove(s->batch, s);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    drop_graphs(s);
    s->graveyard.reap(true);
    if (s->ownStream) (void)hipStreamDestroy(s->ownStream);
    if (s->pool && s->ownsPool) (void)hipFree(s->pool);
    delete s;
    return FS_OK;
}

// resetSimulationFields src/fire_cuda.cu:3006-3060: all fields to their
// initial state (the bed evaluated at time 0), light and shadow zero, clocks 0.
int fs_reset(fs_scene* s) {
    if (int st = fsi::reset_async(s)) return st;
    const hipError_t e = hipStreamSynchronize(s->stream);
    return e == hipSuccess ? FS_OK : s->hip_fail("reset sync", e);
}

int fs_set_params(fs_scene* s, const fs_params* prm) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    if (!prm) return s->fail(FS_ERR_INVALID_ARGUMENT, "null params");
    clamp_params(*prm, s->prm);
    drop_graphs(s);
    return FS_OK;
}

int fs_get_desc(const fs_scene* s, fs_scene_desc* out) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    if (!out) return global_fail(FS_ERR_INVALID_ARGUMENT, "null output");
    out->nx = s->dims.nx;
    out->ny = s->dims.ny;
    out->nz = s->dims.nz;
    out->frame_w = s->dims.fw;
    out->frame_h = s->dims.fh;
    out->storage = s->dims.bf16 ? FS_STORAGE_BF16 : FS_STORAGE_FP32;
    out->device = s->device;
    return FS_OK;
}

int fs_get_params(const fs_scene* s, fs_params* out) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    if (!out) return global_fail(FS_ERR_INVALID_ARGUMENT, "null output");
    *out = s->prm;
    return FS_OK;
}

int fs_set_emitters(fs_scene* s, const fs_emitter* e, int32_t count) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    if (count < 0 || count > FS_MAX_EMITTERS) return s->fail(FS_ERR_INVALID_ARGUMENT, "emitter count out of range [0, 16]");
    if (count > 0 && !e) return s->fail(FS_ERR_INVALID_ARGUMENT, "null emitter array");
    for (int i = 0; i < count; ++i) {
        if (!(e[i].radius >= 0.0f) || !std::isfinite(e[i].cx) || !std::isfinite(e[i].cy) || !std::isfinite(e[i].cz) ||
            !std::isfinite(e[i].fuel_rate) || !std::isfinite(e[i].temperature_k))
            return s->fail(FS_ERR_INVALID_ARGUMENT, "emitter values must be finite with radius >= 0");
    }
    s->emitters.assign(e, e + count);
    drop_graphs(s);
    return FS_OK;
}

int fs_set_tuning(fs_scene* s, const fs_tuning* t) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    if (!t) return s->fail(FS_ERR_INVALID_ARGUMENT, "null tuning");
    if (t->sor_kernel < 0 || t->sor_kernel >= FS_SOR_KERNEL_COUNT) return s->fail(FS_ERR_INVALID_ARGUMENT, "unknown SOR kernel");
    s->ds.sorKernel = t->sor_kernel;
    s->useGraph = t->use_graph != 0;
    drop_graphs(s);
    return FS_OK;
}

int fs_get_tuning(const fs_scene* s, fs_tuning* out) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    if (!out) return global_fail(FS_ERR_INVALID_ARGUMENT, "null output");
    out->sor_kernel = s->ds.sorKernel;
    out->use_graph = s->useGraph ? 1 : 0;
    return FS_OK;
}

int fs_autotune(fs_scene* s, int32_t reps, fs_autotune_report* out) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    if (!out || reps < 1 || reps > 100) return s->fail(FS_ERR_INVALID_ARGUMENT, "null report or reps out of range [1, 100]");
    if (int st = set_device(s)) return st;
    drop_graphs(s);
    const size_t bytes = s->dims.cells() * sizeof(float);
    // keep the pressure field in the force scratch buffer while candidates run
    hipError_t e = hipMemcpyAsync(s->ds.cu, s->ds.p, bytes, hipMemcpyDeviceToDevice, s->stream);
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    const DevParams p = dev_params(*s);
    const int previous = s->ds.sorKernel;
    int best = previous;
    double bestMs = 1e30;
    for (int k = 0; e == hipSuccess && k < FS_SOR_KERNEL_COUNT; ++k) {
        s->ds.sorKernel = k;
        e = fs::launch_pressure(s->ds, p, s->stream);  // warm
        if (e == hipSuccess) e = hipEventRecord(t0, s->stream);
        for (int r = 0; e == hipSuccess && r < reps; ++r) e = fs::launch_pressure(s->ds, p, s->stream);
        if (e == hipSuccess) e = hipEventRecord(t1, s->stream);
        if (e == hipSuccess) e = hipEventSynchronize(t1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
        out->sor_ms[k] = ms / reps;
        if (e == hipSuccess && out->sor_ms[k] < bestMs) {
            bestMs = out->sor_ms[k];
            best = k;
        }
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    s->ds.sorKernel = e == hipSuccess ? best : previous;
    if (e == hipSuccess) e = hipMemcpyAsync(s->ds.p, s->ds.cu, bytes, hipMemcpyDeviceToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return s->hip_fail("fs_autotune", e);
    out->chosen = best;
    return FS_OK;
}

const char* fs_last_error(const fs_scene* s) { return s ? s->lastError.c_str() : g_globalError.c_str(); }

// One step = FP32 clock update (src/fire_cuda.cu:3135-3142), then the passes
// in the reference order (:3196-3318).
int fs_step(fs_scene* s, int32_t nsteps) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    if (nsteps < 0) return s->fail(FS_ERR_INVALID_ARGUMENT, "negative step count");
    if (int st = set_device(s)) return st;
    for (int i = 0; i < nsteps; ++i)
        if (int st = one_step(s)) return st;
    return FS_OK;
}

int fs_step_latency(fs_scene* s, int32_t nsteps, double* step_ms) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    if (!step_ms) return s->fail(FS_ERR_INVALID_ARGUMENT, "null latency output");
    if (nsteps < 0 || nsteps > 4096) return s->fail(FS_ERR_INVALID_ARGUMENT, "step count out of range [0, 4096]");
    if (int st = set_device(s)) return st;
    std::vector<hipEvent_t> ev(static_cast<size_t>(nsteps) + 1, nullptr);
    int status = FS_OK;
    for (auto& e : ev) {
        const hipError_t err = hipEventCreate(&e);
        if (err != hipSuccess) {
            status = s->hip_fail("hipEventCreate", err);
            break;
        }
    }
    if (status == FS_OK && s->useGraph && !s->graphsValid) status = capture_graphs(s);
    for (int i = 0; status == FS_OK && i < nsteps; ++i) {
        hipError_t err = hipEventRecord(ev[i], s->stream);
        if (err != hipSuccess) status = s->hip_fail("hipEventRecord", err);
        if (status == FS_OK) status = one_step(s);
    }
    if (status == FS_OK && nsteps > 0) {
        hipError_t err = hipEventRecord(ev[nsteps], s->stream);
        if (err == hipSuccess) err = hipStreamSynchronize(s->stream);
        if (err != hipSuccess) status = s->hip_fail("latency sync", err);
    }
    for (int i = 0; status == FS_OK && i < nsteps; ++i) {
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
        step_ms[i] = ms;
    }
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    return status;
}

int fs_step_timed(fs_scene* s, int32_t nsteps, fs_stage_times* out, double* step_ms) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    if (!out) return s->fail(FS_ERR_INVALID_ARGUMENT, "null timing output");
    if (nsteps < 0 || nsteps > 4096) return s->fail(FS_ERR_INVALID_ARGUMENT, "step count out of range [0, 4096]");
    if (int st = set_device(s)) return st;
    std::memset(out, 0, sizeof(*out));
    const int per = FS_STAGE_COUNT + 1;
    std::vector<hipEvent_t> ev(static_cast<size_t>(nsteps) * per, nullptr);
    int status = FS_OK;
    for (auto& e : ev) {
        const hipError_t err = hipEventCreate(&e);
        if (err != hipSuccess) {
            status = s->hip_fail("hipEventCreate", err);
            break;
        }
    }
    drop_graphs(s);  // direct launches change the buffer assignment
    for (int i = 0; status == FS_OK && i < nsteps; ++i) {
        s->renderTime += s->prm.dt;
        s->time += s->prm.dt;
        const DevParams p = dev_params(*s);
        hipEvent_t* e = &ev[static_cast<size_t>(i) * per];
        hipError_t err = fs::launch_advance_clock(s->ds, p, s->stream);
        if (err == hipSuccess) err = hipEventRecord(e[0], s->stream);
        for (int stage = 0; err == hipSuccess && status == FS_OK && stage < FS_STAGE_COUNT; ++stage) {
            status = run_stage(s, stage, p, false, true);
            if (status == FS_OK) err = hipEventRecord(e[stage + 1], s->stream);
        }
        if (err != hipSuccess) status = s->hip_fail("hipEventRecord", err);
        ++s->frameIndex;
    }
    if (status == FS_OK) {
        const hipError_t err = hipStreamSynchronize(s->stream);
        if (err != hipSuccess) status = s->hip_fail("timed step sync", err);
    }
    for (int i = 0; status == FS_OK && i < nsteps; ++i) {
        hipEvent_t* e = &ev[static_cast<size_t>(i) * per];
        for (int stage = 0; stage < FS_STAGE_COUNT; ++stage) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e[stage], e[stage + 1]);
            out->stage_ms[stage] += ms;
        }
        float stepMs = 0.0f;
        (void)hipEventElapsedTime(&stepMs, e[0], e[FS_STAGE_COUNT]);
        if (step_ms) step_ms[i] = stepMs;
    }
    if (status == FS_OK && nsteps > 0) {
        float totalMs = 0.0f;
        (void)hipEventElapsedTime(&totalMs, ev[0], ev[static_cast<size_t>(nsteps) * per - 1]);
        out->total_ms = totalMs;
    }
    out->steps = nsteps;
    out->pressure_launches = fs::pressure_launch_count(s->ds, s->prm.sor_iterations);
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    return status;
}

int fs_run_stage(fs_scene* s, int32_t stage) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    if (stage < 0 || stage >= FS_STAGE_COUNT) return s->fail(FS_ERR_INVALID_ARGUMENT, "unknown stage");
    if (int st = set_device(s)) return st;
    drop_graphs(s);  // direct launches change the buffer assignment
    return run_stage(s, stage, dev_params(*s));
}

int fs_synchronize(fs_scene* s) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    if (int st = set_device(s)) return st;
    const hipError_t e = hipStreamSynchronize(s->stream);
    return e == hipSuccess ? FS_OK : s->hip_fail("hipStreamSynchronize", e);
}

int fs_set_time(fs_scene* s, float time, float renderTime) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    if (int st = set_device(s)) return st;
    const float clock[2] = {time, renderTime};
    hipError_t e = hipMemcpyAsync(s->ds.clock, clock, sizeof(clock), hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return s->hip_fail("fs_set_time", e);
    s->time = time;
    s->renderTime = renderTime;
    return FS_OK;
}

int fs_get_time(const fs_scene* s, float* time, float* renderTime, int32_t* frameIndex) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    if (time) *time = s->time;
    if (renderTime) *renderTime = s->renderTime;
    if (frameIndex) *frameIndex = s->frameIndex;
    return FS_OK;
}

int fs_render_hdr(fs_scene* s, float* rgba) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    if (!rgba) return s->fail(FS_ERR_INVALID_ARGUMENT, "output pixel pointer was null");
    if (int st = set_device(s)) return st;
    const DevParams p = dev_params(*s);
    hipError_t e = fs::launch_render(s->ds, p, s->stream);
    if (e != hipSuccess) return s->hip_fail("render launch", e);
    e = hipMemcpyAsync(rgba, s->ds.hdr, s->dims.pixels() * 16, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return e == hipSuccess ? FS_OK : s->hip_fail("render readback", e);
}

int fs_render_bgra8_async(fs_scene* s, uint32_t* bgra) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    if (!bgra) return s->fail(FS_ERR_INVALID_ARGUMENT, "output pixel pointer was null");
    if (int st = set_device(s)) return st;
    const DevParams p = dev_params(*s);
    hipError_t e = fs::launch_render(s->ds, p, s->stream);
    if (e == hipSuccess) e = fs::launch_tonemap(s->ds, p, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(bgra, s->ds.bgra, s->dims.pixels() * 4, hipMemcpyDeviceToHost, s->stream);
    return e == hipSuccess ? FS_OK : s->hip_fail("render launch", e);
}

int fs_render_bgra8(fs_scene* s, uint32_t* bgra) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    if (!bgra) return s->fail(FS_ERR_INVALID_ARGUMENT, "output pixel pointer was null");
    if (int st = set_device(s)) return st;
    const DevParams p = dev_params(*s);
    hipError_t e = fs::launch_render(s->ds, p, s->stream);
    if (e == hipSuccess) e = fs::launch_tonemap(s->ds, p, s->stream);
    if (e != hipSuccess) return s->hip_fail("render launch", e);
    e = hipMemcpyAsync(bgra, s->ds.bgra, s->dims.pixels() * 4, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return e == hipSuccess ? FS_OK : s->hip_fail("frame readback", e);
}

int fs_render_rgba16f(fs_scene* s, uint16_t* rgba16) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    if (!rgba16) return s->fail(FS_ERR_INVALID_ARGUMENT, "output pixel pointer was null");
    if (int st = set_device(s)) return st;
    const DevParams p = dev_params(*s);
    hipError_t e = fs::launch_render(s->ds, p, s->stream);
    if (e == hipSuccess) e = fs::launch_pack_fp16(s->ds, p, s->stream);
    if (e != hipSuccess) return s->hip_fail("render launch", e);
    e = hipMemcpyAsync(rgba16, s->ds.rgba16, s->dims.pixels() * 8, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return e == hipSuccess ? FS_OK : s->hip_fail("frame readback", e);
}

int fs_tonemap(fs_scene* s, const float* rgba, uint32_t* bgra) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    if (!rgba || !bgra) return s->fail(FS_ERR_INVALID_ARGUMENT, "null pixel buffer");
    if (int st = set_device(s)) return st;
    const DevParams p = dev_params(*s);
    hipError_t e = hipMemcpyAsync(s->ds.hdr, rgba, s->dims.pixels() * 16, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess) e = fs::launch_tonemap(s->ds, p, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(bgra, s->ds.bgra, s->dims.pixels() * 4, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return e == hipSuccess ? FS_OK : s->hip_fail("tonemap", e);
}

int fs_field_count(const fs_scene* s, int32_t field, uint64_t* count) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    size_t n = 0;
    if (!count || !field_info(s, field, n)) return global_fail(FS_ERR_INVALID_ARGUMENT, "unknown field");
    *count = n;
    return FS_OK;
}

int fs_read_field(fs_scene* s, int32_t field, float* host) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    size_t n = 0;
    if (!host || !field_info(s, field, n)) return s->fail(FS_ERR_INVALID_ARGUMENT, "unknown field or null buffer");
    if (int st = set_device(s)) return st;
    hipError_t e = hipSuccess;
    if (narrow_field(s, field)) {
        // widen into the force scratch buffer, then copy out
        e = fs::launch_unpack_scalar(s->ds, field, s->ds.cu, s->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(host, s->ds.cu, n * 4, hipMemcpyDeviceToHost, s->stream);
    } else {
        e = hipMemcpyAsync(host, field_ptr(s, field), n * 4, hipMemcpyDeviceToHost, s->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return e == hipSuccess ? FS_OK : s->hip_fail("fs_read_field", e);
}

int fs_write_field(fs_scene* s, int32_t field, const float* host) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    size_t n = 0;
    if (!host || !field_info(s, field, n)) return s->fail(FS_ERR_INVALID_ARGUMENT, "unknown field or null buffer");
    if (int st = set_device(s)) return st;
    hipError_t e = hipSuccess;
    if (narrow_field(s, field)) {
        e = hipMemcpyAsync(s->ds.cu, host, n * 4, hipMemcpyHostToDevice, s->stream);
        if (e == hipSuccess) e = fs::launch_pack_scalar(s->ds, field, s->ds.cu, s->stream);
    } else {
        e = hipMemcpyAsync(field_ptr(s, field), host, n * 4, hipMemcpyHostToDevice, s->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    return e == hipSuccess ? FS_OK : s->hip_fail("fs_write_field", e);
}

int fs_metrics_collect(fs_scene* s, fs_metrics* m) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    if (!m) return s->fail(FS_ERR_INVALID_ARGUMENT, "metrics output pointer was null");
    if (int st = set_device(s)) return st;
    const DevParams p = dev_params(*s);
    hipError_t e = fs::launch_metrics(s->ds, p, s->stream);
    double slots[fs::kMetricSlots];
    double acc[2];
    if (e == hipSuccess) e = hipMemcpyAsync(slots, s->ds.metricsOut, sizeof(slots), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(acc, s->ds.acc, sizeof(acc), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e != hipSuccess) return s->hip_fail("metrics", e);
    std::memset(m, 0, sizeof(*m));
    const double cells = std::max(1.0, static_cast<double>(s->dims.cells()));
    m->frame_index = s->frameIndex;
    m->grid_x = s->dims.nx;
    m->grid_y = s->dims.ny;
    m->grid_z = s->dims.nz;
    m->pressure_iterations = s->prm.sor_iterations;
    m->invalid_cells = static_cast<int32_t>(slots[31]);
    m->time_seconds = s->time;
    m->heat_sum = slots[0]; m->fuel_sum = slots[1]; m->oxygen_sum = slots[2]; m->soot_sum = slots[3];
    m->char_sum = slots[4]; m->ash_sum = slots[5]; m->pyrolysis_sum = slots[6]; m->progress_sum = slots[7];
    m->turbulence_energy_sum = slots[8]; m->soot_optical_depth_sum = slots[9];
    m->mean_scene_light = slots[10] / cells;
    m->mean_scene_shadow = slots[11] / cells;
    m->flame_mass_proxy = slots[12];
    m->smoke_mass_proxy = slots[13];
    m->flame_smoke_overlap_proxy = slots[14];
    m->mean_volumetric_shadow = slots[15] / cells;
    m->mean_room_irradiance = slots[16] / cells;
    m->mean_optical_depth = slots[17] / cells;
    m->heat_release_proxy = slots[18];
    m->divergence_before_l2 = std::sqrt(slots[19] / cells);
    m->divergence_after_l2 = std::sqrt(slots[20] / cells);
    m->sor_residual_l2 = std::sqrt(slots[21] / cells);
    m->max_heat = slots[22]; m->max_fuel = slots[23]; m->max_soot = slots[24]; m->max_pyrolysis = slots[25];
    m->max_progress = slots[26]; m->max_turbulence_energy = slots[27];
    m->divergence_before_max = slots[28];
    m->divergence_after_max = slots[29];
    m->flame_height_meters = slots[30];
    m->divergence_reduction = m->divergence_before_l2 > 0.000001 ? 1.0 - m->divergence_after_l2 / m->divergence_before_l2 : 0.0;
    m->heat_release_step = acc[0];
    m->heat_release_total = acc[1];
    return FS_OK;
}

void* fs_stream(fs_scene* s) { return s ? static_cast<void*>(s->stream) : nullptr; }

}  // extern "C"

namespace fsi {
int create_unreset(const fs_scene_desc* desc, void* memory, uint64_t bytes, fs_scene** out) {
    return create_scene(desc, memory, bytes, out, false);
}

int reset_async(fs_scene* s) {
    if (!s) return global_fail(FS_ERR_NOT_INITIALIZED, "null scene");
    if (int st = set_device(s)) return st;
    // Everything except the scalars is zero at reset; clear the whole pool.
    hipError_t e = hipMemsetAsync(s->pool, 0, s->poolBytes, s->stream);
    if (e != hipSuccess) return s->hip_fail("reset memset", e);
    const DevParams p = dev_params(*s);
    e = fs::launch_reset(s->ds, p, s->stream);
    if (e != hipSuccess) return s->hip_fail("reset kernel", e);
    s->time = 0.0f;
    s->renderTime = 0.0f;
    s->frameIndex = 0;
    return FS_OK;
}

DevParams dev_params(const fs_scene& sc) { return ::dev_params(sc); }
void drop_graphs(fs_scene* s) { ::drop_graphs(s); }
int global_fail(int status, const std::string& msg) { return ::global_fail(status, msg); }
}  // namespace fsi
