// fs_launch.hpp - host-side view of one scene's device buffers and the kernel
// launchers that the orchestration in scene.cpp calls.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "firesim.h"
#include "fs_device.hpp"

namespace fs {

// Number of doubles each metrics block writes (22 sums, 9 maxima, invalid count).
constexpr int kMetricSlots = 32;
constexpr int kMetricBlocks = 1024;

// One spark of the ember pass, as the splat kernel needs it (screen-space
// centre, core radius, streak radius and direction, life fade, colour and
// the pixel bounding box; active == 0 for sparks that do not draw).
struct Ember {
    float sx, sy, radius, streak, tdx, tdy, fade, r, g, b;
    int x0, x1, y0, y1, active, pad;
};
static_assert(sizeof(Ember) == 64, "one spark is 64 bytes");

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
    float* psi[3] = {};  // curl-noise potential on grid edges, nx * (ny + 1) * nz each
    float4 *light = nullptr, *lightNext = nullptr;
    float* shadow = nullptr;
    float* dirT = nullptr;  // transmittance toward the directional light, one per light cell
    float4* hdr = nullptr;
    uint32_t* bgra = nullptr;
    uint16_t* rgba16 = nullptr;    // FP16 RGBA frame, 4 halves per pixel
    double* partials = nullptr;  // per-block partial sums
    size_t partialCapacity = 0;  // doubles
    double* acc = nullptr;       // [0] heat release of the last step, [1] since reset
    double* metricsOut = nullptr;  // kMetricSlots doubles
    float* clock = nullptr;        // [simulation time, render time]
    Ember* embers = nullptr;       // kMaxEmbers sparks of the last render
    int sorKernel = FS_SOR_TB_T24_Z10_K4;  // enum fs_sor_kernel
};

// ------------------------------------------------------------ batches ----
// A batch steps many scenes of different sizes with one launch per pass. The
// launch is persistent: a fixed grid of workgroups walks the concatenated
// block lists of all scenes (per-scene tile grids identical to the one-scene
// launches, so every cell sees the same arithmetic and results equal the
// one-scene path bit for bit). Scenes are ordered by SOR half-sweep count,
// largest first, so the scenes still active in pressure launch j are a prefix.
constexpr int kMaxBatch = 512;
constexpr int kMaxSorLaunches = 4096;

// One scene as the passes of one step see it (the buffer roles the one-scene
// launchers would assign before the step).
struct BatchScene {
    DevParams p;
    void* sIn[NSCALAR];   // scalars at step start
    void* sOut[NSCALAR];  // advect-react output = scalars after the step
    float *velA[3];       // u, v, w at step start; force output; projected
    float *velB[3];       // advected velocity; force input
    float *p0, *p1;       // pressure buffers in launch order (p0 zeroed by the force pass)
    float* pFinal;        // buffer holding the solution after the pressure pass
    float *d, *cu, *cv, *cw;
    float* psi[3];
    float4 *light, *lightNext;
    float* shadow;
    float* dirT;
    double* partials;
    double* acc;
    float* clock;
    int halfSweeps;                 // 2 * SOR iterations
    int cellG[3], faceG[3], sorG[3], lightG[3];  // per-pass block grids
};

struct BatchTable {
    int count;
    int sorDepth;  // half-sweeps per pressure launch (1: plain half-sweep kernel)
    uint32_t cellStart[kMaxBatch + 1];  // prefix sums of per-scene block counts
    uint32_t faceStart[kMaxBatch + 1];
    uint32_t sorStart[kMaxBatch + 1];
    uint32_t lightStart[kMaxBatch + 1];
    uint16_t sorActive[kMaxSorLaunches];  // scenes taking part in pressure launch j
    BatchScene scene[1];  // `count` entries follow
};

inline size_t batch_table_bytes(int count) { return sizeof(BatchTable) + sizeof(BatchScene) * (count > 0 ? count - 1 : 0); }

// Persistent walk over the concatenated block lists of the first `count`
// scenes (start[] = prefix sums of their block counts). The hardware hands
// workgroup w to XCD w % 8; XCD i takes the i-th eighth of the logical blocks
// and its workgroups stride through it, so the tiles an XCD works on are
// neighbours that share its L2. f(scene, local block) runs with
// workgroup-uniform arguments, so it may contain barriers.
template <class F>
__device__ __forceinline__ void batch_blocks(const uint32_t* __restrict__ start, int count, F&& f) {
    if (count <= 0) return;
    const uint32_t total = start[count];
    const unsigned per = gridDim.x / kXcdCount;
    const unsigned xcd = blockIdx.x % kXcdCount, slot = blockIdx.x / kXcdCount;
    if (slot >= per) return;
    const uint32_t c0 = static_cast<uint32_t>((static_cast<uint64_t>(total) * xcd) / kXcdCount);
    const uint32_t c1 = static_cast<uint32_t>((static_cast<uint64_t>(total) * (xcd + 1)) / kXcdCount);
    uint32_t b = c0 + slot;
    if (b >= c1) return;
    int lo = 0, hi = count - 1;  // scene of block b: last s with start[s] <= b
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (start[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    int s = lo;
    for (; b < c1; b += per) {
        while (start[s + 1] <= b) ++s;
        f(s, b - start[s]);
    }
}

// Local block number -> 3D block coordinate of a per-scene grid g.
__device__ __forceinline__ uint3 block_of(uint32_t local, const int* g) {
    const uint32_t gx = static_cast<uint32_t>(g[0]), gy = static_cast<uint32_t>(g[1]);
    return make_uint3(local % gx, (local / gx) % gy, local / (gx * gy));
}

hipError_t launch_batch_light(const BatchTable* table, bool bf16, int grid, hipStream_t st);

// Launches one batched step: clock, the six passes. `grid` is the persistent
// workgroup count (a multiple of 8); `sorLaunches` the pressure launch count.
hipError_t launch_batch_step(const BatchTable* table, bool bf16, int sorKernel, int sorLaunches, int grid, hipStream_t st);
// The pressure launches of a batched step alone (autotuning). zeroGuess (tile
// kernels only): start from p = 0 without reading p.
hipError_t launch_batch_pressure(const BatchTable* table, int sorKernel, int sorLaunches, int grid, hipStream_t st, bool zeroGuess = false);
// Persistent grid size for this device (workgroups resident at once, a multiple of 16).
int batch_grid(int device);

// Grid and block shape of the cell-parallel passes (64 lanes along x).
dim3 cell_block();
dim3 cell_grid(int nx, int ny, int nz);
// z planes of a force-contribution workgroup (64 x 4 x kForcePlanes cells,
// one curl-noise lattice build each), and that kernel's grid.
constexpr int kForcePlanes = 4;
dim3 force_grid(int nx, int ny, int nz);
dim3 light_block();
dim3 light_grid(int lnx, int lny, int lnz);
dim3 sor_grid(int sorKernel, int nx, int ny, int nz);

hipError_t launch_reset(const DeviceScene& s, const DevParams& p, hipStream_t st);
hipError_t launch_advance_clock(DeviceScene& s, const DevParams& p, hipStream_t st);
// advanceClock: also advance the device clock by one step (launch_advance_clock).
hipError_t launch_advect_velocity(DeviceScene& s, const DevParams& p, hipStream_t st, bool advanceClock = false);
// finalize = false leaves the heat-release sum of the react partials to the
// next launch_force with finalizeRelease = true (one launch fewer per step).
hipError_t launch_advect_react(DeviceScene& s, const DevParams& p, hipStream_t st, bool finalize = true);
// writeP = false skips the force pass's p = 0 store, for a step whose pressure
// pass then runs with zeroGuess.
hipError_t launch_force(DeviceScene& s, const DevParams& p, hipStream_t st, bool finalizeRelease = false, bool writeP = true);
hipError_t launch_pressure(DeviceScene& s, const DevParams& p, hipStream_t st, bool zeroGuess = false);
// The pressure pass on bare buffers: p holds the initial guess, p2 is the
// second buffer the tile kernels ping-pong with (unused by the half-sweep
// kernel). *resultInP2 tells where the solution ends (an odd launch count).
// zeroGuess (tile kernels only; the half-sweep kernel is rejected): the
// initial guess is p = 0 and p is not read.
hipError_t launch_pressure_buffers(float* p, float* p2, const float* d, int nx, int ny, int nz, float omega, int iterations, int sorKernel,
                                   hipStream_t st, bool* resultInP2, bool zeroGuess = false);
// n floats from a to b (n < 2^30), a plain kernel (no memset or copy node).
hipError_t launch_copy_f32(const float* a, float* b, size_t n, hipStream_t st);
hipError_t launch_project(DeviceScene& s, const DevParams& p, hipStream_t st);
// bytes / 16 float4 elements from a to b (bytes a multiple of 16).
hipError_t launch_copy4(const void* a, void* b, size_t bytes, hipStream_t st);
hipError_t launch_light(DeviceScene& s, const DevParams& p, hipStream_t st);
hipError_t launch_render(DeviceScene& s, const DevParams& p, hipStream_t st);
hipError_t launch_tonemap(DeviceScene& s, const DevParams& p, hipStream_t st);
hipError_t launch_pack_fp16(DeviceScene& s, const DevParams& p, hipStream_t st);
hipError_t launch_metrics(DeviceScene& s, const DevParams& p, hipStream_t st);
// Converts between caller FP32 data and a scalar field's storage type.
hipError_t launch_pack_scalar(const DeviceScene& s, int field, const float* src, hipStream_t st);
hipError_t launch_unpack_scalar(const DeviceScene& s, int field, float* dst, hipStream_t st);

size_t react_partial_count(int nx, int ny, int nz);
// Kernel launches one pressure pass issues for `iterations` SOR iterations.
int pressure_launch_count(const DeviceScene& s, int iterations);
// Half-sweeps one temporally blocked SOR launch performs (1 for the plain kernel).
int sor_depth(int sorKernel);
int sor_tile_edge(int sorKernel);
int sor_tile_depth(int sorKernel);

}  // namespace fs
