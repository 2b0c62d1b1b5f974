// fire_smoke.cpp - the reference's smoke-test sequence (runCudaSmokeTest,
// src/main.cpp:4170-4218) driven through fire_api.h: initialise for a
// 960x540 frame and the canonical 384x240 grid request (176x208x128 cells),
// run 64 step+render frames with the same per-frame settings, write the last
// frame as a BMP. Usage: fire_smoke <output.bmp> [assets dir]. Exit codes as
// the reference: 0 ok, 2 initialisation failed, 3 a frame or the BMP failed.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/firesim.h"
#include "../fire_api.h"

int main(int argc, char** argv) {
    const char* out = argc > 1 ? argv[1] : "fire-smoke-test-frame.bmp";
    const char* assets = argc > 2 ? argv[2] : nullptr;
    constexpr int kFrameWidth = 960, kFrameHeight = 540;
    std::vector<std::uint32_t> frame(static_cast<size_t>(kFrameWidth) * kFrameHeight, 0xff000000u);
    if (!fireCudaInitialize(kFrameWidth, kFrameHeight, 384, 240)) {
        std::fprintf(stderr, "initialize failed: %s\n", fireCudaLastError());
        return 2;
    }
    bool ok = true;
    FireCudaFrameMetrics metrics{};
    for (int i = 0; i < 64 && ok; ++i) {
        FireSettings s;
        s.width = kFrameWidth;
        s.height = kFrameHeight;
        s.dt = 1.0f / 60.0f;
        s.mouseX = 0.50f + 0.08f * std::sin(static_cast<float>(i) * 0.08f);
        s.mouseY = 0.13f + 0.04f * std::sin(static_cast<float>(i) * 0.045f);
        s.wind = 0.08f;
        s.detail = 0.92f;
        fireApplyCanonicalSettings(&s);
        fireApplySceneEmitter(&s, assets);
        s.turbulence = std::max(s.turbulence, 1.08f);
        s.smoke = std::max(s.smoke, 1.05f);
        s.intensity = std::max(s.intensity, 1.05f);
        s.cameraYaw = 0.18f;
        s.cameraPitch = 0.08f;
        s.cameraDistance = 2.62f;
        s.showGizmos = 0;  // the frame call gets simSettings.showGizmos = 0 (src/main.cpp:4204-4206)
        ok = (i == 63) ? fireCudaStepAndRenderMeasured(frame.data(), s, &metrics) : fireCudaStepAndRender(frame.data(), s);
        if (!ok) std::fprintf(stderr, "frame %d failed: %s\n", i, fireCudaLastError());
    }
    if (ok) ok = fs_write_bmp(out, frame.data(), kFrameWidth, kFrameHeight) == FS_OK;
    if (ok) {
        std::printf("{\"grid\": [%d, %d, %d], \"frame\": %d, \"invalidCells\": %d, \"maxHeat\": %.6f, \"heatSum\": %.6f, "
                    "\"flameHeightMeters\": %.6f, \"divergenceAfterL2\": %.6g, \"divergenceAfterMax\": %.6g, "
                    "\"gpuSolveMs\": %.4f, \"gpuProjectionMs\": %.4f, \"renderMs\": %.4f}\n",
                    metrics.gridX, metrics.gridY, metrics.gridZ, metrics.frameIndex, metrics.invalidCells, metrics.maxHeat,
                    metrics.heatSum, metrics.flameHeightMeters, metrics.divergenceAfterL2, metrics.divergenceAfterMax,
                    metrics.gpuSolveMs, metrics.gpuProjectionMs, metrics.gpuRenderMs);
    }
    fireCudaShutdown();
    return ok ? 0 : 3;
}
