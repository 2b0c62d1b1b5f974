// fire_api.cpp - the reference's process-wide step interface (fire_api.h)
// over the per-scene C ABI.
//
// Behaviour follows the reference host path stepAndRenderInternal
// (src/fire_cuda.cu:3106-3549): the grid is derived from the requested size
// (:2916-2922), settings are clamped per call (:3135-3179), an in-call reset
// precedes the step, the step advances both clocks and a render-only call
// advances only the render clock, and every call returns false with a message
// in fireCudaLastError() on failure. The pressure solve keeps the reference
// constants (omega 1.52, 40 iterations, :104-105).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "../../include/firesim.h"
#include "../fire_api.h"

namespace {

fs_scene* g_scene = nullptr;
int g_frameIndex = 0;
fs_params g_applied{};
bool g_haveApplied = false;
char g_lastError[512] = "No error.";

bool fail(const char* msg) {
    std::snprintf(g_lastError, sizeof(g_lastError), "%s", msg);
    return false;
}

bool fail_fs(const char* what) {
    std::snprintf(g_lastError, sizeof(g_lastError), "%s: %s", what, fs_last_error(g_scene));
    return false;
}

int clampi(int v, int lo, int hi) { return std::max(lo, std::min(hi, v)); }
float clampf(float v, float lo, float hi) { return std::max(lo, std::min(hi, v)); }

// FireSettings -> fs_params with the reference's per-call clamps.
fs_params to_params(const FireSettings& s) {
    fs_params p;
    fs_default_params(&p);
    p.dt = clampf(s.dt, 0.001f, 1.0f / 30.0f);
    p.mouse_x = clampf(s.mouseX, 0.0f, 1.0f);
    p.mouse_y = clampf(s.mouseY, 0.0f, 1.0f);
    p.left_down = s.leftDown;
    p.right_down = s.rightDown;
    p.scene_id = clampi(s.sceneId, 0, 2);
    p.wind_x = clampf(s.wind, -1.0f, 1.0f);
    p.wind_z = 0.0f;
    p.turbulence = clampf(s.turbulence, 0.05f, 1.8f);
    p.detail = clampf(s.detail, 0.1f, 1.5f);
    p.smoke = clampf(s.smoke, 0.0f, 1.6f);
    p.intensity = clampf(s.intensity, 0.2f, 2.4f);
    p.camera_yaw = s.cameraYaw;
    p.camera_pitch = clampf(s.cameraPitch, -0.75f, 0.75f);
    p.camera_distance = clampf(s.cameraDistance, 1.5f, 6.0f);
    p.cinematic_mode = s.cinematicMode != 0 ? 1 : 0;
    p.raymarch_steps = clampi(s.raymarchSteps, 24, 120);
    p.ember_count = clampi(s.emberCount, 0, 192);
    p.render_debug_mode = clampi(s.renderDebugMode, 0, 6);
    p.show_gizmos = s.showGizmos != 0 ? 1 : 0;
    p.active_gizmo = s.activeGizmo;
    p.exposure = clampf(s.exposure, 0.45f, 2.40f);
    p.reflection_gain = clampf(s.reflectionGain, 0.0f, 2.2f);
    p.smoke_darkness = clampf(s.smokeDarkness, 0.35f, 2.4f);
    p.emitter_center_x = s.emitterCenterX;
    p.emitter_center_z = s.emitterCenterZ;
    p.emitter_height_norm = s.emitterHeightNorm;
    p.emitter_height_band_norm = s.emitterHeightBandNorm;
    p.emitter_radius = s.emitterRadius;
    p.burner_center_count = s.burnerCenterCount;
    for (int i = 0; i < 4; ++i) {
        p.burner_center_x[i] = s.burnerCenterX[i];
        p.burner_center_y[i] = s.burnerCenterY[i];
        p.burner_center_z[i] = s.burnerCenterZ[i];
    }
    p.sor_omega = 1.52f;
    p.sor_iterations = 40;
    fs_clamp_params(&p, &p);
    return p;
}

bool apply(const FireSettings& s) {
    const fs_params p = to_params(s);
    // Re-applying identical parameters would throw away the captured step graph.
    if (g_haveApplied && std::memcmp(&p, &g_applied, sizeof(p)) == 0) return true;
    if (fs_set_params(g_scene, &p) != FS_OK) return fail_fs("settings");
    g_applied = p;
    g_haveApplied = true;
    return true;
}

bool step_and_render(std::uint32_t* bgra, const FireSettings& settings, FireCudaFrameMetrics* metrics, bool advance, bool render,
                     bool d3d) {
    if (!g_scene) return fail("Renderer is not initialized.");
    if (render && !d3d && bgra == nullptr) return fail("Output pixel pointer was null.");
    if (render && d3d) return fail("D3D11 interop is not available in this library; render to a host buffer with fireCudaStepAndRender.");
    if (!render && !advance) return fail("Step-only path requires simulation advancement.");
    if (!apply(settings)) return false;
    // In-call reset (:3136-3143, :3181, :3053-3055): the reference advances
    // its clocks, captures the step time, then the reset zeroes the clocks.
    // So the frame that resets still steps and renders at the pre-reset time
    // plus dt, and reports time 0; the next frame steps at time dt.
    const bool resetNow = settings.reset != 0;
    if (resetNow) {
        float t0 = 0.0f, rt0 = 0.0f;
        int32_t frame = 0;
        fs_get_time(g_scene, &t0, &rt0, &frame);
        if (fs_reset(g_scene) != FS_OK) return fail_fs("reset");
        if (fs_set_time(g_scene, t0, rt0) != FS_OK) return fail_fs("reset clock");
        g_frameIndex = 0;
    }
    fs_stage_times times{};
    if (advance) {
        const int st = metrics ? fs_step_timed(g_scene, 1, &times, nullptr) : fs_step(g_scene, 1);
        if (st != FS_OK) return fail_fs("step");
    } else {
        float t = 0.0f, rt = 0.0f;
        int32_t frame = 0;
        fs_get_time(g_scene, &t, &rt, &frame);
        if (fs_set_time(g_scene, t, rt + g_applied.dt) != FS_OK) return fail_fs("render clock");
    }
    double renderMs = 0.0;
    if (render) {
        const auto t0 = std::chrono::steady_clock::now();
        if (fs_render_bgra8(g_scene, bgra) != FS_OK) return fail_fs("render");
        renderMs = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    if (resetNow && fs_set_time(g_scene, 0.0f, 0.0f) != FS_OK) return fail_fs("reset clock");
    if (metrics) {
        fs_metrics m{};
        if (fs_metrics_collect(g_scene, &m) != FS_OK) return fail_fs("metrics");
        FireCudaFrameMetrics out{};
        out.frameIndex = g_frameIndex;
        out.gridX = m.grid_x;
        out.gridY = m.grid_y;
        out.gridZ = m.grid_z;
        out.pressureIterations = m.pressure_iterations;
        out.invalidCells = m.invalid_cells;
        out.timeSeconds = static_cast<float>(m.time_seconds);
        out.gpuVelocityMs = static_cast<float>(times.stage_ms[FS_STAGE_ADVECT_VELOCITY]);
        out.gpuReactionMs = static_cast<float>(times.stage_ms[FS_STAGE_ADVECT_REACT] + times.stage_ms[FS_STAGE_FORCE]);
        out.gpuProjectionMs = static_cast<float>(times.stage_ms[FS_STAGE_PRESSURE] + times.stage_ms[FS_STAGE_PROJECT]);
        out.gpuLightingMs = static_cast<float>(times.stage_ms[FS_STAGE_LIGHT]);
        out.gpuSolveMs = out.gpuVelocityMs + out.gpuReactionMs + out.gpuProjectionMs;
        out.gpuRaymarchMs = static_cast<float>(renderMs);
        out.gpuRenderMs = static_cast<float>(renderMs);
        out.heatSum = static_cast<float>(m.heat_sum);
        out.fuelSum = static_cast<float>(m.fuel_sum);
        out.oxygenSum = static_cast<float>(m.oxygen_sum);
        out.sootSum = static_cast<float>(m.soot_sum);
        out.charSum = static_cast<float>(m.char_sum);
        out.ashSum = static_cast<float>(m.ash_sum);
        out.pyrolysisSum = static_cast<float>(m.pyrolysis_sum);
        out.progressSum = static_cast<float>(m.progress_sum);
        out.turbulenceEnergySum = static_cast<float>(m.turbulence_energy_sum);
        out.sootOpticalDepthSum = static_cast<float>(m.soot_optical_depth_sum);
        out.maxHeat = static_cast<float>(m.max_heat);
        out.maxFuel = static_cast<float>(m.max_fuel);
        out.maxSoot = static_cast<float>(m.max_soot);
        out.maxPyrolysis = static_cast<float>(m.max_pyrolysis);
        out.maxProgress = static_cast<float>(m.max_progress);
        out.maxTurbulenceEnergy = static_cast<float>(m.max_turbulence_energy);
        out.flameHeightMeters = static_cast<float>(m.flame_height_meters);
        out.meanOpticalDepth = static_cast<float>(m.mean_optical_depth);
        out.meanSceneLight = static_cast<float>(m.mean_scene_light);
        out.meanSceneShadow = static_cast<float>(m.mean_scene_shadow);
        out.flameMassProxy = static_cast<float>(m.flame_mass_proxy);
        out.smokeMassProxy = static_cast<float>(m.smoke_mass_proxy);
        out.flameSmokeOverlapProxy = static_cast<float>(m.flame_smoke_overlap_proxy);
        out.meanVolumetricShadow = static_cast<float>(m.mean_volumetric_shadow);
        out.meanRoomIrradiance = static_cast<float>(m.mean_room_irradiance);
        out.heatReleaseProxy = static_cast<float>(m.heat_release_proxy);
        out.divergenceBeforeL2 = static_cast<float>(m.divergence_before_l2);
        out.divergenceBeforeMax = static_cast<float>(m.divergence_before_max);
        out.divergenceAfterL2 = static_cast<float>(m.divergence_after_l2);
        out.divergenceAfterMax = static_cast<float>(m.divergence_after_max);
        out.divergenceReduction = static_cast<float>(m.divergence_reduction);
        *metrics = out;
    }
    ++g_frameIndex;
    return true;
}

}  // namespace

// src/fire_cuda.cu:2912-3004
bool fireCudaInitialize(int frameWidth, int frameHeight, int gridWidth, int gridHeight) {
    fireCudaShutdown();
    if (frameWidth <= 0 || frameHeight <= 0) return fail("Frame dimensions must be positive.");
    const bool highQuality = gridWidth >= 320 || gridHeight >= 220;
    fs_scene_desc d{};
    d.nx = clampi(gridWidth / 2, 72, highQuality ? 176 : 104);
    d.ny = clampi(gridHeight - 32, 80, highQuality ? 208 : 112);
    d.nz = clampi(gridWidth / 3, 48, highQuality ? 132 : 72);
    d.frame_w = frameWidth;
    d.frame_h = frameHeight;
    d.storage = FS_STORAGE_FP32;
    d.device = 0;
    if (fs_create(&d, &g_scene) != FS_OK) {
        std::snprintf(g_lastError, sizeof(g_lastError), "initialize: %s", fs_last_global_error());
        g_scene = nullptr;
        return false;
    }
    g_frameIndex = 0;
    g_haveApplied = false;
    std::snprintf(g_lastError, sizeof(g_lastError), "No error.");
    // A failed initialization frees everything it created, as the reference's
    // does (:2981): later calls then report "not initialized".
    if (!apply(FireSettings{})) {
        fireCudaShutdown();
        return false;
    }
    // The fastest SOR tile for this grid (every tile gives the same result):
    // one timed pressure pass per tile on the freshly reset fields, which the
    // autotuner leaves as they were.
    fs_autotune_report report{};
    if (fs_autotune(g_scene, 1, &report) != FS_OK) {
        fail_fs("SOR autotune");
        fireCudaShutdown();
        return false;
    }
    return true;
}

bool fireCudaStepAndRender(std::uint32_t* bgraPixels, const FireSettings& settings) {
    return step_and_render(bgraPixels, settings, nullptr, true, true, false);
}

bool fireCudaStepAndRenderMeasured(std::uint32_t* bgraPixels, const FireSettings& settings, FireCudaFrameMetrics* metrics) {
    if (metrics == nullptr) return fail("Metrics output pointer was null.");
    return step_and_render(bgraPixels, settings, metrics, true, true, false);
}

bool fireCudaSelectDeviceForD3D11(void*) { return fail("D3D11 interop is not available in this library."); }
bool fireCudaRegisterD3D11Texture(void*) { return fail("D3D11 interop is not available in this library."); }
bool fireCudaRegisterD3D11TextureSlot(int, void*) { return fail("D3D11 interop is not available in this library."); }
bool fireCudaSetD3D11TextureSlot(int) { return fail("D3D11 interop is not available in this library."); }

// Step only, no render (src/fire_cuda.cu:3624-3626).
bool fireCudaStepD3D11(const FireSettings& settings) { return step_and_render(nullptr, settings, nullptr, true, false, false); }

bool fireCudaStepAndRenderD3D11(const FireSettings& settings) { return step_and_render(nullptr, settings, nullptr, true, true, true); }

bool fireCudaStepAndRenderD3D11Measured(const FireSettings& settings, FireCudaFrameMetrics* metrics) {
    if (metrics == nullptr) return fail("Metrics output pointer was null.");
    return step_and_render(nullptr, settings, metrics, true, true, true);
}

bool fireCudaRenderD3D11(const FireSettings& settings) { return step_and_render(nullptr, settings, nullptr, false, true, true); }

void fireCudaUnregisterD3D11Texture() {}

bool fireCudaSynchronize() {
    if (!g_scene) return fail("Renderer is not initialized.");
    return fs_synchronize(g_scene) == FS_OK ? true : fail_fs("synchronize");
}

bool fireCudaReset() {
    if (!g_scene) return fail("Renderer is not initialized.");
    if (fs_reset(g_scene) != FS_OK) return fail_fs("reset");
    g_frameIndex = 0;
    return true;
}

bool fireCudaGetDiagnostics(FireCudaDiagnostics* diagnostics) {
    if (diagnostics == nullptr) return fail("Diagnostics output pointer was null.");
    FireCudaDiagnostics out{};
    if (hipDriverGetVersion(&out.driverVersion) != hipSuccess || hipRuntimeGetVersion(&out.runtimeVersion) != hipSuccess ||
        hipGetDeviceCount(&out.deviceCount) != hipSuccess)
        return fail("HIP runtime query failed.");
    if (out.deviceCount <= 0) return fail("No HIP device was found.");
    int device = 0;
    hipDeviceProp_t props{};
    if (hipGetDevice(&device) != hipSuccess || hipGetDeviceProperties(&props, device) != hipSuccess)
        return fail("HIP device query failed.");
    out.activeDevice = device;
    out.computeMajor = props.major;
    out.computeMinor = props.minor;
    out.totalGlobalMem = static_cast<std::uint64_t>(props.totalGlobalMem);
    std::snprintf(out.deviceName, sizeof(out.deviceName), "%s", props.name);
    *diagnostics = out;
    return true;
}

void fireCudaShutdown() {
    if (g_scene) fs_destroy(g_scene);
    g_scene = nullptr;
    g_haveApplied = false;
}

const char* fireCudaLastError() { return g_lastError; }

fs_scene* fireCudaScene() { return g_scene; }

void fireApplyCanonicalSettings(FireSettings* s) {
    if (!s) return;
    s->sceneId = std::max(0, std::min(2, s->sceneId));
    s->cinematicMode = 1;
    s->raymarchSteps = 104;
    s->emberCount = 176;
    s->exposure = 0.84f;
    s->reflectionGain = 0.54f;
    s->smokeDarkness = 1.26f;
    s->intensity = 1.42f;
    s->smoke = 0.96f;
    s->turbulence = 1.34f;
    s->renderDebugMode = 0;
}

bool fireApplySceneEmitter(FireSettings* s, const char* assetsDir) {
    if (!s) return fail("Settings pointer was null.");
    fs_params p;
    fs_default_params(&p);
    if (fs_scene_emitter_params(std::max(0, std::min(2, s->sceneId)), assetsDir, &p) != FS_OK) return fail("Unknown scene id.");
    s->emitterCenterX = p.emitter_center_x;
    s->emitterCenterZ = p.emitter_center_z;
    s->emitterHeightNorm = p.emitter_height_norm;
    s->emitterHeightBandNorm = p.emitter_height_band_norm;
    s->emitterRadius = p.emitter_radius;
    s->burnerCenterCount = p.burner_center_count;
    for (int i = 0; i < 4; ++i) {
        s->burnerCenterX[i] = p.burner_center_x[i];
        s->burnerCenterY[i] = p.burner_center_y[i];
        s->burnerCenterZ[i] = p.burner_center_z[i];
    }
    return true;
}
