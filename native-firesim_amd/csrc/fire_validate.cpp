// fire_validate.cpp - the reference's validation run (runValidation
// src/main.cpp:4220-4583) on this library, through fire_api.h.
//
// Runs the canonical scene for --validation-frames frames (default 96) with
// the reference's per-frame settings, writes the same three artefacts with the
// same columns and keys (validation-metrics.csv, validation-report.json,
// validation-frame.bmp; validation-app-frame.bmp holds the same frame, as the
// display overlay is not part of this library), applies the reference's
// stability checks and any --targets= envelope CSV ("metric,min,max" rows),
// and exits 0 when validation passes, 4 when it fails, 2/3 on setup errors.
//
// Options as the reference: --targets= --output-dir= --validation-frames=
// --scene= --pool-fire-calibration --dataset-id= --manifest=, plus
// --assets= (scene contract directory, default assets/fire-scenes). The
// real-burn calibration comparison (--calibration=, --geometry=) is not part
// of this library: providing either marks the run as failed, as the reference
// does for an unreadable calibration file.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <sys/stat.h>
#include <vector>

#include "../../include/firesim.h"
#include "../fire_api.h"

namespace {

std::string arg_value(const std::string& args, const char* key) {
    const size_t at = args.find(key);
    if (at == std::string::npos) return {};
    const size_t start = at + std::string(key).size();
    const size_t end = args.find(' ', start);
    return args.substr(start, end == std::string::npos ? std::string::npos : end - start);
}

int arg_int(const std::string& args, const char* key, int fallback, int lo, int hi) {
    const std::string v = arg_value(args, key);
    if (v.empty()) return fallback;
    return std::max(lo, std::min(hi, std::atoi(v.c_str())));
}

std::string trim(const std::string& v) {
    const size_t b = v.find_first_not_of(" \t\r\n");
    if (b == std::string::npos) return {};
    return v.substr(b, v.find_last_not_of(" \t\r\n") - b + 1);
}

std::string json_escape(const std::string& v) {
    std::string o;
    for (char c : v) {
        if (c == '"' || c == '\\') o += '\\';
        o += c;
    }
    return o;
}

struct Target {
    std::string metric;
    float minimum = 0.0f, maximum = 0.0f, observed = 0.0f;
    bool matched = false, passed = false;
};

// "metric,min,max" rows; '#' comments and a "metric" header are skipped
// (loadTargetEnvelopes src/main.cpp:3679-3710).
std::vector<Target> load_targets(const std::string& path) {
    std::vector<Target> out;
    if (path.empty()) return out;
    std::ifstream in(path);
    std::string line;
    while (std::getline(in, line)) {
        line = trim(line);
        if (line.empty() || line[0] == '#') continue;
        const size_t c1 = line.find(','), c2 = c1 == std::string::npos ? c1 : line.find(',', c1 + 1);
        if (c2 == std::string::npos) continue;
        Target t;
        t.metric = trim(line.substr(0, c1));
        if (t.metric == "metric") continue;
        t.minimum = std::strtof(trim(line.substr(c1 + 1, c2 - c1 - 1)).c_str(), nullptr);
        t.maximum = std::strtof(trim(line.substr(c2 + 1)).c_str(), nullptr);
        out.push_back(t);
    }
    return out;
}

bool make_dirs(const std::string& path) {
    std::string cur;
    std::stringstream ss(path);
    std::string part;
    if (!path.empty() && path[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        mkdir(cur.c_str(), 0755);
    }
    struct stat st;
    return stat(path.c_str(), &st) == 0 && S_ISDIR(st.st_mode);
}

}  // namespace

int main(int argc, char** argv) {
    std::string args;
    for (int i = 1; i < argc; ++i) args += std::string(argv[i]) + " ";
    const std::string targetPath = arg_value(args, "--targets=");
    const std::string calibrationPath = arg_value(args, "--calibration=");
    const std::string geometryPath = arg_value(args, "--geometry=");
    const std::string manifestPath = arg_value(args, "--manifest=");
    const std::string datasetId = arg_value(args, "--dataset-id=");
    const std::string outArg = arg_value(args, "--output-dir=");
    const std::string assets = arg_value(args, "--assets=");
    const std::string outputDir = outArg.empty() ? "out" : outArg;
    const int frames = arg_int(args, "--validation-frames=", 96, 8, 240);
    const int scene = arg_int(args, "--scene=", 0, 0, 2);
    const bool poolFire = args.find("--pool-fire-calibration") != std::string::npos;
    constexpr int kW = 960, kH = 540, kGridW = 384, kGridH = 240;
    if (!make_dirs(outputDir)) return 3;
    const std::string metricsPath = outputDir + "/validation-metrics.csv";
    const std::string reportPath = outputDir + "/validation-report.json";
    const std::string rawPath = outputDir + "/validation-frame.bmp";
    const std::string appPath = outputDir + "/validation-app-frame.bmp";
    std::vector<uint32_t> frame(static_cast<size_t>(kW) * kH, 0xff000000u);
    if (!fireCudaInitialize(kW, kH, kGridW, kGridH)) {
        std::fprintf(stderr, "initialize: %s\n", fireCudaLastError());
        return 2;
    }
    std::ofstream csv(metricsPath, std::ios::binary);
    if (!csv) {
        fireCudaShutdown();
        return 3;
    }
    csv << "frame,timeSeconds,gridX,gridY,gridZ,pressureIterations,gpuSolveMs,gpuRenderMs,"
           "gpuVelocityMs,gpuReactionMs,gpuProjectionMs,gpuLightingMs,gpuRaymarchMs,gpuPackMs,"
           "heatSum,fuelSum,oxygenSum,sootSum,charSum,ashSum,pyrolysisSum,progressSum,"
           "turbulenceEnergySum,sootOpticalDepthSum,maxHeat,maxFuel,maxSoot,maxPyrolysis,"
           "maxProgress,maxTurbulenceEnergy,flameHeightMeters,"
           "meanOpticalDepth,meanSceneLight,meanSceneShadow,flameMassProxy,smokeMassProxy,flameSmokeOverlapProxy,"
           "meanVolumetricShadow,meanRoomIrradiance,heatReleaseProxy,divergenceBeforeL2,divergenceAfterL2,"
           "divergenceBeforeMax,divergenceAfterMax,divergenceReduction,invalidCells\n";
    csv << std::fixed << std::setprecision(6);

    bool stable = true;
    double tSolve = 0, tRender = 0, tVel = 0, tReact = 0, tProj = 0, tLight = 0, tRay = 0, tPack = 0, tReduction = 0;
    float worstL2 = 0.0f, worstMax = 0.0f, minReduction = 1.0f, maxHeat = 0.0f, maxFlame = 0.0f;
    FireCudaFrameMetrics last{};
    for (int i = 0; i < frames; ++i) {
        // per-frame settings of the reference run (:4290-4318)
        FireSettings s;
        s.width = kW;
        s.height = kH;
        s.dt = 1.0f / 60.0f;
        s.mouseX = poolFire ? 0.50f : 0.50f + 0.07f * std::sin(static_cast<float>(i) * 0.071f);
        s.mouseY = poolFire ? 0.12f : 0.12f + 0.035f * std::sin(static_cast<float>(i) * 0.053f);
        s.leftDown = 0;
        s.rightDown = 0;
        s.showGizmos = 0;
        s.activeGizmo = 1;
        s.sceneId = scene;
        s.reset = i == 0 ? 1 : 0;
        s.wind = poolFire ? 0.0f : 0.10f + 0.05f * std::sin(static_cast<float>(i) * 0.037f);
        s.detail = 0.98f;
        fireApplyCanonicalSettings(&s);
        fireApplySceneEmitter(&s, assets.empty() ? nullptr : assets.c_str());
        s.turbulence = std::max(s.turbulence, 1.08f);
        s.smoke = std::max(s.smoke, 1.02f);
        s.intensity = std::max(s.intensity, 1.18f);
        s.cameraYaw = 0.18f;
        s.cameraPitch = 0.08f;
        s.cameraDistance = 2.62f;
        FireCudaFrameMetrics m{};
        if (!fireCudaStepAndRenderMeasured(frame.data(), s, &m)) {
            std::fprintf(stderr, "frame %d: %s\n", i, fireCudaLastError());
            stable = false;
            break;
        }
        csv << m.frameIndex << "," << m.timeSeconds << "," << m.gridX << "," << m.gridY << "," << m.gridZ << "," << m.pressureIterations << ","
            << m.gpuSolveMs << "," << m.gpuRenderMs << "," << m.gpuVelocityMs << "," << m.gpuReactionMs << "," << m.gpuProjectionMs << ","
            << m.gpuLightingMs << "," << m.gpuRaymarchMs << "," << m.gpuPackMs << "," << m.heatSum << "," << m.fuelSum << "," << m.oxygenSum
            << "," << m.sootSum << "," << m.charSum << "," << m.ashSum << "," << m.pyrolysisSum << "," << m.progressSum << ","
            << m.turbulenceEnergySum << "," << m.sootOpticalDepthSum << "," << m.maxHeat << "," << m.maxFuel << "," << m.maxSoot << ","
            << m.maxPyrolysis << "," << m.maxProgress << "," << m.maxTurbulenceEnergy << "," << m.flameHeightMeters << ","
            << m.meanOpticalDepth << "," << m.meanSceneLight << "," << m.meanSceneShadow << "," << m.flameMassProxy << ","
            << m.smokeMassProxy << "," << m.flameSmokeOverlapProxy << "," << m.meanVolumetricShadow << "," << m.meanRoomIrradiance << ","
            << m.heatReleaseProxy << "," << m.divergenceBeforeL2 << "," << m.divergenceAfterL2 << "," << m.divergenceBeforeMax << ","
            << m.divergenceAfterMax << "," << m.divergenceReduction << "," << m.invalidCells << "\n";
        // per-frame stability checks (:4334-4336)
        stable = stable && m.invalidCells == 0 && m.divergenceAfterL2 < 1.25f && m.divergenceAfterMax < 8.0f;
        tSolve += m.gpuSolveMs;
        tRender += m.gpuRenderMs;
        tVel += m.gpuVelocityMs;
        tReact += m.gpuReactionMs;
        tProj += m.gpuProjectionMs;
        tLight += m.gpuLightingMs;
        tRay += m.gpuRaymarchMs;
        tPack += m.gpuPackMs;
        tReduction += m.divergenceReduction;
        worstL2 = std::max(worstL2, m.divergenceAfterL2);
        worstMax = std::max(worstMax, m.divergenceAfterMax);
        minReduction = std::min(minReduction, m.divergenceReduction);
        maxHeat = std::max(maxHeat, m.maxHeat);
        maxFlame = std::max(maxFlame, m.flameHeightMeters);
        last = m;
    }
    fs_image_stats img{};
    fs_image_stats_of(frame.data(), static_cast<int64_t>(frame.size()), &img);
    const double n = std::max(1, last.frameIndex + 1);
    const double aSolve = tSolve / n, aRender = tRender / n, aVel = tVel / n, aReact = tReact / n, aProj = tProj / n;
    const double aLight = tLight / n, aRay = tRay / n, aPack = tPack / n, aReduction = tReduction / n;
    std::vector<Target> targets = load_targets(targetPath);
    stable = stable && (targetPath.empty() || !targets.empty());
    stable = stable && calibrationPath.empty() && geometryPath.empty();
    for (Target& t : targets) {
        const std::string& k = t.metric;
        t.matched = true;
        if (k == "averageDivergenceReduction") t.observed = static_cast<float>(aReduction);
        else if (k == "worstDivergenceAfterL2") t.observed = worstL2;
        else if (k == "worstDivergenceAfterMax") t.observed = worstMax;
        else if (k == "maxFlameHeightMeters") t.observed = maxFlame;
        else if (k == "finalMeanOpticalDepth") t.observed = last.meanOpticalDepth;
        else if (k == "finalMeanSceneLight") t.observed = last.meanSceneLight;
        else if (k == "finalMeanSceneShadow") t.observed = last.meanSceneShadow;
        else if (k == "finalHeatReleaseProxy") t.observed = last.heatReleaseProxy;
        else if (k == "finalCharSum") t.observed = last.charSum;
        else if (k == "finalAshSum") t.observed = last.ashSum;
        else if (k == "finalPyrolysisSum") t.observed = last.pyrolysisSum;
        else if (k == "finalProgressSum") t.observed = last.progressSum;
        else if (k == "finalTurbulenceEnergySum") t.observed = last.turbulenceEnergySum;
        else if (k == "finalSootOpticalDepthSum") t.observed = last.sootOpticalDepthSum;
        else if (k == "maxHeat") t.observed = maxHeat;
        else if (k == "maxPyrolysis") t.observed = last.maxPyrolysis;
        else if (k == "maxProgress") t.observed = last.maxProgress;
        else if (k == "maxTurbulenceEnergy") t.observed = last.maxTurbulenceEnergy;
        else if (k == "imageMeanLuma") t.observed = img.mean_luma;
        else if (k == "imageMaxLuma") t.observed = img.max_luma;
        else if (k == "imageMeanSaturation") t.observed = img.mean_saturation;
        else if (k == "imageWarmFireFraction") t.observed = img.warm_fire_fraction;
        else if (k == "imageTanSmokeFraction") t.observed = img.tan_smoke_fraction;
        else if (k == "imageBlackSmokeFraction") t.observed = img.black_smoke_fraction;
        else if (k == "imageWhiteCoreFraction") t.observed = img.white_core_fraction;
        else if (k == "imageBrightPixels") t.observed = static_cast<float>(img.bright_pixels);
        else if (k == "imageBrightFraction") t.observed = img.bright_fraction;
        else if (k == "averageGpuSolveMs") t.observed = static_cast<float>(aSolve);
        else if (k == "averageGpuRenderMs") t.observed = static_cast<float>(aRender);
        else if (k == "averageGpuLightingMs") t.observed = static_cast<float>(aLight);
        else t.matched = false;  // includes the calibration* metrics
        t.passed = t.matched && t.observed >= t.minimum && t.observed <= t.maximum;
        stable = stable && t.passed;
    }
    // end-of-run checks (:4462-4467)
    stable = stable && last.frameIndex >= frames - 1 && maxHeat > 0.50f && maxFlame > 0.20f && img.max_luma > 0.12f &&
             img.bright_pixels > 32 && (aReduction > 0.02 || worstL2 < 0.02f);
    const bool wroteRaw = fs_write_bmp(rawPath.c_str(), frame.data(), kW, kH) == FS_OK;
    const bool wroteApp = fs_write_bmp(appPath.c_str(), frame.data(), kW, kH) == FS_OK;
    std::ofstream json(reportPath, std::ios::binary);
    if (!json) {
        fireCudaShutdown();
        return 3;
    }
    json << std::fixed << std::setprecision(6) << "{\n";
    json << "  \"validationOk\": " << (stable && wroteRaw && wroteApp ? "true" : "false") << ",\n";
    json << "  \"model\": {\n"
         << "    \"runtime\": \"HIP (MI355X, gfx950)\",\n"
         << "    \"pressureSolver\": \"weighted red-black SOR\",\n"
         << "    \"scalarTransport\": \"clamped MacCormack/BFECC correction over semi-Lagrangian backtraces\",\n"
         << "    \"combustion\": \"fuel-bed char/ash pyrolysis plus oxygen-limited Arrhenius progress variable\",\n"
         << "    \"turbulence\": \"LES-style scalar turbulence-energy closure\",\n"
         << "    \"soot\": \"soot optical depth with oxidation feedback and particle-size-derived absorption/scattering\",\n"
         << "    \"renderer\": \"linear HDR blackbody Beer-Lambert participating media with scene radiance volume GI, volume shadowing, emitter scattering, and ACES display tonemapping\"\n"
         << "  },\n";
    json << "  \"runtimeConfig\": \"" << (poolFire ? "nist-pool-fire-calibration" : "canonical") << "\",\n";
    json << "  \"datasetId\": \"" << json_escape(datasetId) << "\",\n";
    json << "  \"manifestPath\": \"" << json_escape(manifestPath) << "\",\n";
    json << "  \"outputDir\": \"" << json_escape(outputDir) << "\",\n";
    json << "  \"requestedGrid\": [" << kGridW << ", " << kGridH << "],\n";
    json << "  \"raymarchSteps\": 104,\n  \"emberCount\": 176,\n";
    json << "  \"frames\": " << frames << ",\n";
    json << "  \"grid\": [" << last.gridX << ", " << last.gridY << ", " << last.gridZ << "],\n";
    json << "  \"pressureIterations\": " << last.pressureIterations << ",\n";
    json << "  \"targetEnvelopePath\": \"" << json_escape(targetPath) << "\",\n";
    json << "  \"calibrationPath\": \"" << json_escape(calibrationPath) << "\",\n";
    json << "  \"geometryPath\": \"" << json_escape(geometryPath) << "\",\n";
    json << "  \"calibration\": {\"provided\": " << (calibrationPath.empty() ? "false" : "true")
         << ", \"readable\": false, \"geometryProvided\": " << (geometryPath.empty() ? "false" : "true")
         << ", \"geometryReadable\": false, \"sampleCount\": 0},\n";
    const std::pair<const char*, double> scalars[] = {
        {"averageGpuSolveMs", aSolve}, {"averageGpuRenderMs", aRender}, {"averageGpuVelocityMs", aVel},
        {"averageGpuReactionMs", aReact}, {"averageGpuProjectionMs", aProj}, {"averageGpuLightingMs", aLight},
        {"averageGpuRaymarchMs", aRay}, {"averageGpuPackMs", aPack}, {"averageDivergenceReduction", aReduction},
        {"minimumDivergenceReduction", minReduction}, {"worstDivergenceAfterL2", worstL2}, {"worstDivergenceAfterMax", worstMax},
        {"finalMeanOpticalDepth", last.meanOpticalDepth}, {"finalMeanSceneLight", last.meanSceneLight},
        {"finalMeanSceneShadow", last.meanSceneShadow}, {"finalFlameMassProxy", last.flameMassProxy},
        {"finalSmokeMassProxy", last.smokeMassProxy}, {"finalFlameSmokeOverlapProxy", last.flameSmokeOverlapProxy},
        {"finalMeanVolumetricShadow", last.meanVolumetricShadow}, {"finalMeanRoomIrradiance", last.meanRoomIrradiance},
        {"finalHeatReleaseProxy", last.heatReleaseProxy}, {"finalCharSum", last.charSum}, {"finalAshSum", last.ashSum},
        {"finalPyrolysisSum", last.pyrolysisSum}, {"finalProgressSum", last.progressSum},
        {"finalTurbulenceEnergySum", last.turbulenceEnergySum}, {"finalSootOpticalDepthSum", last.sootOpticalDepthSum},
        {"maxHeat", maxHeat}, {"maxPyrolysis", last.maxPyrolysis}, {"maxProgress", last.maxProgress},
        {"maxTurbulenceEnergy", last.maxTurbulenceEnergy}, {"maxFlameHeightMeters", maxFlame},
        {"imageMeanLuma", img.mean_luma}, {"imageMaxLuma", img.max_luma}, {"imageMeanSaturation", img.mean_saturation},
        {"imageWarmFireFraction", img.warm_fire_fraction}, {"imageTanSmokeFraction", img.tan_smoke_fraction},
        {"imageBlackSmokeFraction", img.black_smoke_fraction}, {"imageWhiteCoreFraction", img.white_core_fraction}};
    for (const auto& kv : scalars) json << "  \"" << kv.first << "\": " << kv.second << ",\n";
    json << "  \"imageBrightPixels\": " << img.bright_pixels << ",\n";
    json << "  \"imageBrightFraction\": " << img.bright_fraction << ",\n";
    json << "  \"targetEnvelopes\": [\n";
    for (size_t i = 0; i < targets.size(); ++i) {
        const Target& t = targets[i];
        json << "    {\"metric\": \"" << json_escape(t.metric) << "\", \"min\": " << t.minimum << ", \"max\": " << t.maximum
             << ", \"observed\": " << t.observed << ", \"matched\": " << (t.matched ? "true" : "false")
             << ", \"passed\": " << (t.passed ? "true" : "false") << "}" << (i + 1 == targets.size() ? "\n" : ",\n");
    }
    json << "  ],\n";
    json << "  \"outputs\": [\"" << json_escape(metricsPath) << "\", \"" << json_escape(reportPath) << "\", \"" << json_escape(rawPath)
         << "\", \"" << json_escape(appPath) << "\"]\n}\n";
    const bool ok = stable && wroteRaw && wroteApp && json.good() && csv.good();
    fireCudaShutdown();
    return ok ? 0 : 4;
}
