// fs_launch.hpp - host-side view of one scene's device buffers and the kernel
// launchers that the orchestration in scene.cpp calls.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "fs_device.hpp"

namespace fs {

// Number of doubles each metrics block writes (22 sums, 9 maxima, invalid count).
constexpr int kMetricSlots = 32;
constexpr int kMetricBlocks = 1024;

struct DeviceScene {
    int nx = 0, ny = 0, nz = 0, lnx = 0, lny = 0, lnz = 0, fw = 0, fh = 0;
    bool bf16 = false;
    void* s[NSCALAR] = {};      // current scalars
    void* sNext[NSCALAR] = {};  // advect-react output, swapped after the pass
    float *u = nullptr, *v = nullptr, *w = nullptr;
    float *uN = nullptr, *vN = nullptr, *wN = nullptr;
    float *p = nullptr, *d = nullptr;
    float* p2 = nullptr;  // second pressure buffer: tile SOR launches read one, write the other
    float *cu = nullptr, *cv = nullptr, *cw = nullptr;  // per-cell force contributions
    float4 *light = nullptr, *lightNext = nullptr;
    float* shadow = nullptr;
    float4* hdr = nullptr;
    uint32_t* bgra = nullptr;
    double* partials = nullptr;  // per-block partial sums
    size_t partialCapacity = 0;  // doubles
    double* acc = nullptr;       // [0] heat release of the last step, [1] since reset
    double* metricsOut = nullptr;  // kMetricSlots doubles
    float* clock = nullptr;        // [simulation time, render time]
    int sorKernel = 2;           // enum fs_sor_kernel
};

// Grid and block shape of the cell-parallel passes (64 lanes along x).
dim3 cell_block();
dim3 cell_grid(int nx, int ny, int nz);

hipError_t launch_reset(const DeviceScene& s, const DevParams& p, hipStream_t st);
hipError_t launch_advance_clock(DeviceScene& s, const DevParams& p, hipStream_t st);
hipError_t launch_advect_velocity(DeviceScene& s, const DevParams& p, hipStream_t st);
hipError_t launch_advect_react(DeviceScene& s, const DevParams& p, hipStream_t st);
hipError_t launch_force(DeviceScene& s, const DevParams& p, hipStream_t st);
hipError_t launch_pressure(DeviceScene& s, const DevParams& p, hipStream_t st);
hipError_t launch_project(DeviceScene& s, const DevParams& p, hipStream_t st);
hipError_t launch_light(DeviceScene& s, const DevParams& p, hipStream_t st);
hipError_t launch_render(DeviceScene& s, const DevParams& p, hipStream_t st);
hipError_t launch_tonemap(DeviceScene& s, const DevParams& p, hipStream_t st);
hipError_t launch_metrics(DeviceScene& s, const DevParams& p, hipStream_t st);
// Converts between caller FP32 data and a scalar field's storage type.
hipError_t launch_pack_scalar(const DeviceScene& s, int field, const float* src, hipStream_t st);
hipError_t launch_unpack_scalar(const DeviceScene& s, int field, float* dst, hipStream_t st);

size_t react_partial_count(int nx, int ny, int nz);
// Kernel launches one pressure pass issues for `iterations` SOR iterations.
int pressure_launch_count(const DeviceScene& s, int iterations);
// Half-sweeps one temporally blocked SOR launch performs (1 for the plain kernel).
int sor_depth(int sorKernel);

}  // namespace fs
