// batch.cpp - many scenes stepped together (fs_batch_* in include/firesim.h).
//
// The reference runs one simulation per process (src/fire_cuda.cu:116-195).
// A batch keeps any number of resident scenes and advances all of them with
// one persistent launch per pass: the launch walks the concatenated tile lists
// of every member (BatchTable, fs_launch.hpp). Per step the host writes the
// table of buffer roles into pinned memory, copies it to the device on the
// batch stream, and replays the captured step graph for the step's pressure
// launch count (or launches directly). Each member's host bookkeeping (buffer
// swaps, clocks, frame index) then advances exactly as fs_step would.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>
#include <string>
#include <vector>

#include "scene_impl.hpp"

using fs::BatchScene;
using fs::BatchTable;

struct fs_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<fs_scene*> members;  // in insertion order
    bool bf16 = false;
    int sorKernel = FS_SOR_TB_T24_Z10_K4;
    bool useGraph = true;
    int grid = 0;
    BatchTable* devTable = nullptr;
    BatchTable* hostTable[2] = {nullptr, nullptr};  // pinned ring
    hipEvent_t hostFree[2] = {nullptr, nullptr};    // recorded after the copy out of hostTable[i]
    int ring = 0;
    // Captured steps. A graph's kernels are fixed at capture by the SOR kernel,
    // the storage mode (the FP32 or BF16 passes) and the pressure launch count,
    // so all three key it; the table it reads is rewritten before every launch.
    std::map<std::tuple<int, bool, int>, hipGraphExec_t> graphs;
    GraphGraveyard graveyard;  // dropped graphs whose last launch may still run
    hipEvent_t joinEvent = nullptr;
    std::string lastError = "no error";

    int fail(int status, const std::string& msg) {
        lastError = msg;
        return status;
    }
    int hip_fail(const char* what, hipError_t e) { return fail(FS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e)); }
};

namespace {

int half_sweeps(const fs_scene* s) { return 2 * s->prm.sor_iterations; }

int launches_of(int halfSweeps, int sorKernel) {
    const int k = fs::sor_depth(sorKernel);
    return (halfSweeps + k - 1) / k;
}

uint32_t blocks(dim3 g) { return g.x * g.y * g.z; }

void set3(int* dst, dim3 g) {
    dst[0] = static_cast<int>(g.x);
    dst[1] = static_cast<int>(g.y);
    dst[2] = static_cast<int>(g.z);
}

// Members ordered by SOR half-sweeps, most first (stable).
std::vector<fs_scene*> step_order(const fs_batch* b) {
    std::vector<fs_scene*> order = b->members;
    std::stable_sort(order.begin(), order.end(), [](const fs_scene* a, const fs_scene* c) { return half_sweeps(a) > half_sweeps(c); });
    return order;
}

// Fills t for one step from the members' current buffer assignment and
// returns the pressure launch count.
int build_table(const fs_batch* b, int sorKernel, BatchTable* t) {
    const std::vector<fs_scene*> order = step_order(b);
    const int n = static_cast<int>(order.size());
    const bool tiled = fs::sor_tile_edge(sorKernel) != 0;
    t->count = n;
    t->sorDepth = fs::sor_depth(sorKernel);
    t->cellStart[0] = t->faceStart[0] = t->sorStart[0] = t->lightStart[0] = 0;
    int maxLaunches = 0;
    for (int i = 0; i < n; ++i) {
        const fs_scene* s = order[i];
        const DeviceScene& d = s->ds;
        BatchScene& e = t->scene[i];
        e.p = fsi::dev_params(*s);
        for (int k = 0; k < fs::NSCALAR; ++k) {
            e.sIn[k] = d.s[k];
            e.sOut[k] = d.sNext[k];
        }
        e.velA[0] = d.u; e.velA[1] = d.v; e.velA[2] = d.w;
        e.velB[0] = d.uN; e.velB[1] = d.vN; e.velB[2] = d.wN;
        e.p0 = d.p;
        e.p1 = d.p2;
        e.halfSweeps = half_sweeps(s);
        const int launches = launches_of(e.halfSweeps, sorKernel);
        e.pFinal = (tiled && (launches & 1)) ? d.p2 : d.p;
        e.d = d.d; e.cu = d.cu; e.cv = d.cv; e.cw = d.cw;
        for (int k = 0; k < 3; ++k) e.psi[k] = d.psi[k];
        e.light = d.light; e.lightNext = d.lightNext; e.shadow = d.shadow; e.dirT = d.dirT;
        e.partials = d.partials; e.acc = d.acc; e.clock = d.clock;
        const dim3 cg = fs::cell_grid(d.nx, d.ny, d.nz), fg = fs::cell_grid(d.nx, d.ny + 1, d.nz + 1);  // face passes: x = nx - 1 also covers x = nx
        const dim3 sg = fs::sor_grid(sorKernel, d.nx, d.ny, d.nz), lg = fs::light_grid(d.lnx, d.lny, d.lnz);
        set3(e.cellG, cg);
        set3(e.faceG, fg);
        set3(e.sorG, sg);
        set3(e.lightG, lg);
        t->cellStart[i + 1] = t->cellStart[i] + blocks(cg);
        t->faceStart[i + 1] = t->faceStart[i] + blocks(fg);
        t->sorStart[i + 1] = t->sorStart[i] + blocks(sg);
        t->lightStart[i + 1] = t->lightStart[i] + blocks(lg);
        maxLaunches = std::max(maxLaunches, launches);
    }
    // scenes are sorted by half-sweeps, so those in launch j are a prefix
    int active = n;
    for (int j = 0; j < maxLaunches; ++j) {
        while (active > 0 && launches_of(t->scene[active - 1].halfSweeps, sorKernel) <= j) --active;
        t->sorActive[j] = static_cast<uint16_t>(active);
    }
    return maxLaunches;
}

size_t table_bytes(int count) { return fs::batch_table_bytes(std::max(count, 1)); }

// Graphs may have launches queued or running on the batch stream: they are
// destroyed once the stream has passed them.
void drop_batch_graphs(fs_batch* b) {
    for (auto& kv : b->graphs) b->graveyard.retire(kv.second, b->stream);
    b->graphs.clear();
    b->graveyard.reap(false);
}

// Host bookkeeping of one completed step, the same as fs_step's.
void advance_members(fs_batch* b) {
    const bool tiled = fs::sor_tile_edge(b->sorKernel) != 0;
    for (fs_scene* s : b->members) {
        for (int k = 0; k < fs::NSCALAR; ++k) std::swap(s->ds.s[k], s->ds.sNext[k]);
        if (tiled && (launches_of(half_sweeps(s), b->sorKernel) & 1)) std::swap(s->ds.p, s->ds.p2);
        s->renderTime += s->prm.dt;
        s->time += s->prm.dt;
        ++s->frameIndex;
        // the scene's own step graphs assume its own SOR kernel's buffer order
        if (s->graphsValid) fsi::drop_graphs(s);
    }
}

int graph_for(fs_batch* b, int launches, hipGraphExec_t* out) {
    const auto key = std::make_tuple(b->sorKernel, b->bf16, launches);
    auto it = b->graphs.find(key);
    if (it != b->graphs.end()) {
        *out = it->second;
        return FS_OK;
    }
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return b->hip_fail("hipStreamBeginCapture", e);
    const hipError_t le = fs::launch_batch_step(b->devTable, b->bf16, b->sorKernel, launches, b->grid, b->stream);
    e = hipStreamEndCapture(b->stream, &graph);
    if (le != hipSuccess || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return b->hip_fail("batch step capture", le != hipSuccess ? le : e);
    }
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return b->hip_fail("hipGraphInstantiate", e);
    b->graphs[key] = exec;
    *out = exec;
    return FS_OK;
}

// Writes the step table, copies it to the device and enqueues the step.
int enqueue_batch_step(fs_batch* b) {
    const int slot = b->ring;
    b->ring ^= 1;
    hipError_t e = hipEventSynchronize(b->hostFree[slot]);
    if (e != hipSuccess) return b->hip_fail("batch table event", e);
    BatchTable* t = b->hostTable[slot];
    const int launches = build_table(b, b->sorKernel, t);
    e = hipMemcpyAsync(b->devTable, t, table_bytes(t->count), hipMemcpyHostToDevice, b->stream);
    if (e == hipSuccess) e = hipEventRecord(b->hostFree[slot], b->stream);
    if (e != hipSuccess) return b->hip_fail("batch table copy", e);
    if (b->useGraph) {
        hipGraphExec_t exec = nullptr;
        if (int st = graph_for(b, launches, &exec)) return st;
        e = hipGraphLaunch(exec, b->stream);
        if (e != hipSuccess) return b->hip_fail("hipGraphLaunch", e);
    } else {
        e = fs::launch_batch_step(b->devTable, b->bf16, b->sorKernel, launches, b->grid, b->stream);
        if (e != hipSuccess) return b->hip_fail("batch step", e);
    }
    advance_members(b);
    return FS_OK;
}

int set_device(fs_batch* b) {
    const hipError_t e = hipSetDevice(b->device);
    return e == hipSuccess ? FS_OK : b->hip_fail("hipSetDevice", e);
}

void release(fs_batch* b) {
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    drop_batch_graphs(b);
    b->graveyard.reap(true);
    for (int i = 0; i < 2; ++i) {
        if (b->hostFree[i]) (void)hipEventDestroy(b->hostFree[i]);
        if (b->hostTable[i]) (void)hipHostFree(b->hostTable[i]);
    }
    if (b->joinEvent) (void)hipEventDestroy(b->joinEvent);
    if (b->devTable) (void)hipFree(b->devTable);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // namespace

extern "C" {

int fs_batch_create(int32_t device, fs_batch** out) {
    if (!out) return fsi::global_fail(FS_ERR_INVALID_ARGUMENT, "null output");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) return fsi::global_fail(FS_ERR_HIP, "no HIP device");
    if (device < 0 || device >= count) return fsi::global_fail(FS_ERR_INVALID_ARGUMENT, "device ordinal out of range");
    e = hipSetDevice(device);
    if (e != hipSuccess) return fsi::global_fail(FS_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    fs_batch* b = new (std::nothrow) fs_batch;
    if (!b) return fsi::global_fail(FS_ERR_OUT_OF_MEMORY, "batch handle");
    b->device = device;
    b->grid = fs::batch_grid(device);
    const size_t bytes = table_bytes(fs::kMaxBatch);
    e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&b->devTable, bytes);
    for (int i = 0; e == hipSuccess && i < 2; ++i) {
        e = hipHostMalloc(&b->hostTable[i], bytes, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->hostFree[i], hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventRecord(b->hostFree[i], b->stream);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->joinEvent, hipEventDisableTiming);
    if (e != hipSuccess) {
        const std::string msg = std::string("batch create: ") + hipGetErrorString(e);
        (void)hipGetLastError();
        release(b);
        return fsi::global_fail(e == hipErrorOutOfMemory ? FS_ERR_OUT_OF_MEMORY : FS_ERR_HIP, msg);
    }
    *out = b;
    return FS_OK;
}

int fs_batch_destroy(fs_batch* b) {
    if (!b) return FS_OK;
    (void)hipSetDevice(b->device);
    while (!b->members.empty()) (void)fs_batch_remove(b, b->members.back());
    (void)hipStreamSynchronize(b->stream);
    release(b);
    return FS_OK;
}

int fs_batch_add(fs_batch* b, fs_scene* s) {
    if (!b) return fsi::global_fail(FS_ERR_NOT_INITIALIZED, "null batch");
    if (!s) return b->fail(FS_ERR_INVALID_ARGUMENT, "null scene");
    if (s->batch) return b->fail(FS_ERR_INVALID_ARGUMENT, s->batch == b ? "scene is already in this batch" : "scene belongs to another batch");
    if (s->device != b->device) return b->fail(FS_ERR_INVALID_ARGUMENT, "scene lives on another device");
    if (static_cast<int>(b->members.size()) >= fs::kMaxBatch) return b->fail(FS_ERR_INVALID_ARGUMENT, "batch is full");
    if (!b->members.empty() && s->dims.bf16 != b->bf16) return b->fail(FS_ERR_INVALID_ARGUMENT, "a batch holds scenes of one storage mode");
    if (int st = set_device(b)) return st;
    // the scene's pending work precedes everything the batch issues for it
    hipError_t e = hipEventRecord(b->joinEvent, s->ownStream);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->joinEvent, 0);
    if (e != hipSuccess) return b->hip_fail("join scene stream", e);
    if (b->members.empty()) b->bf16 = s->dims.bf16;
    b->members.push_back(s);
    s->batch = b;
    s->stream = b->stream;
    return FS_OK;
}

int fs_batch_remove(fs_batch* b, fs_scene* s) {
    if (!b) return fsi::global_fail(FS_ERR_NOT_INITIALIZED, "null batch");
    auto it = std::find(b->members.begin(), b->members.end(), s);
    if (!s || it == b->members.end()) return b->fail(FS_ERR_INVALID_ARGUMENT, "scene is not in this batch");
    if (int st = set_device(b)) return st;
    // later work on the scene's own stream waits for the batch's
    hipError_t e = hipEventRecord(b->joinEvent, b->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(s->ownStream, b->joinEvent, 0);
    b->members.erase(it);
    s->batch = nullptr;
    s->stream = s->ownStream;
    return e == hipSuccess ? FS_OK : b->hip_fail("leave batch stream", e);
}

int fs_batch_count(const fs_batch* b, int32_t* count) {
    if (!b) return fsi::global_fail(FS_ERR_NOT_INITIALIZED, "null batch");
    if (!count) return fsi::global_fail(FS_ERR_INVALID_ARGUMENT, "null output");
    *count = static_cast<int32_t>(b->members.size());
    return FS_OK;
}

int fs_batch_set_tuning(fs_batch* b, const fs_tuning* t) {
    if (!b) return fsi::global_fail(FS_ERR_NOT_INITIALIZED, "null batch");
    if (!t) return b->fail(FS_ERR_INVALID_ARGUMENT, "null tuning");
    if (t->sor_kernel < 0 || t->sor_kernel >= FS_SOR_KERNEL_COUNT) return b->fail(FS_ERR_INVALID_ARGUMENT, "unknown SOR kernel");
    // Members whose last batched step used the other kernel's buffer order are
    // unaffected: the table is rebuilt from their state every step. Graphs are
    // keyed by kernel, so those of the other kernel stay valid.
    b->sorKernel = t->sor_kernel;
    b->useGraph = t->use_graph != 0;
    return FS_OK;
}

int fs_batch_get_tuning(const fs_batch* b, fs_tuning* out) {
    if (!b) return fsi::global_fail(FS_ERR_NOT_INITIALIZED, "null batch");
    if (!out) return fsi::global_fail(FS_ERR_INVALID_ARGUMENT, "null output");
    out->sor_kernel = b->sorKernel;
    out->use_graph = b->useGraph ? 1 : 0;
    return FS_OK;
}

int fs_batch_autotune(fs_batch* b, int32_t reps, fs_autotune_report* out) {
    if (!b) return fsi::global_fail(FS_ERR_NOT_INITIALIZED, "null batch");
    if (reps < 1 || reps > 100) return b->fail(FS_ERR_INVALID_ARGUMENT, "reps must be in [1, 100]");
    if (b->members.empty()) return b->fail(FS_ERR_INVALID_ARGUMENT, "batch is empty");
    if (int st = set_device(b)) return st;
    // keep every member's pressure; cu is per-step scratch
    hipError_t e = hipSuccess;
    for (fs_scene* s : b->members)
        if (e == hipSuccess) e = hipMemcpyAsync(s->ds.cu, s->ds.p, s->dims.cells() * 4, hipMemcpyDeviceToDevice, b->stream);
    hipEvent_t t0 = nullptr, t1 = nullptr;
    if (e == hipSuccess) e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    fs_autotune_report rep{};
    int best = b->sorKernel;
    double bestMs = 1e300;
    for (int k = 0; e == hipSuccess && k < FS_SOR_KERNEL_COUNT; ++k) {
        e = hipEventSynchronize(b->hostFree[b->ring]);
        if (e != hipSuccess) break;
        BatchTable* t = b->hostTable[b->ring];
        const int launches = build_table(b, k, t);
        e = hipMemcpyAsync(b->devTable, t, table_bytes(t->count), hipMemcpyHostToDevice, b->stream);
        if (e == hipSuccess) e = hipEventRecord(b->hostFree[b->ring], b->stream);
        if (e == hipSuccess) e = fs::launch_batch_pressure(b->devTable, k, launches, b->grid, b->stream);  // warm
        if (e == hipSuccess) e = hipEventRecord(t0, b->stream);
        for (int r = 0; e == hipSuccess && r < reps; ++r) e = fs::launch_batch_pressure(b->devTable, k, launches, b->grid, b->stream);
        if (e == hipSuccess) e = hipEventRecord(t1, b->stream);
        if (e == hipSuccess) e = hipEventSynchronize(t1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
        rep.sor_ms[k] = static_cast<double>(ms) / reps;
        if (e == hipSuccess && rep.sor_ms[k] < bestMs) {
            bestMs = rep.sor_ms[k];
            best = k;
        }
    }
    for (fs_scene* s : b->members)
        if (e == hipSuccess) e = hipMemcpyAsync(s->ds.p, s->ds.cu, s->dims.cells() * 4, hipMemcpyDeviceToDevice, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (e != hipSuccess) return b->hip_fail("batch autotune", e);
    b->sorKernel = best;
    rep.chosen = best;
    if (out) *out = rep;
    return FS_OK;
}

int fs_batch_step(fs_batch* b, int32_t nsteps) {
    if (!b) return fsi::global_fail(FS_ERR_NOT_INITIALIZED, "null batch");
    if (nsteps < 0) return b->fail(FS_ERR_INVALID_ARGUMENT, "nsteps must be >= 0");
    if (b->members.empty() || nsteps == 0) return FS_OK;
    if (int st = set_device(b)) return st;
    for (int i = 0; i < nsteps; ++i)
        if (int st = enqueue_batch_step(b)) return st;
    return FS_OK;
}

int fs_batch_step_latency(fs_batch* b, int32_t nsteps, double* stepMs) {
    if (!b) return fsi::global_fail(FS_ERR_NOT_INITIALIZED, "null batch");
    if (nsteps < 1 || nsteps > 100000) return b->fail(FS_ERR_INVALID_ARGUMENT, "nsteps must be in [1, 100000]");
    if (b->members.empty()) return b->fail(FS_ERR_INVALID_ARGUMENT, "batch is empty");
    if (int st = set_device(b)) return st;
    std::vector<hipEvent_t> ev(nsteps + 1, nullptr);
    hipError_t e = hipSuccess;
    for (auto& x : ev)
        if (e == hipSuccess) e = hipEventCreate(&x);
    int st = FS_OK;
    for (int i = 0; e == hipSuccess && st == FS_OK && i < nsteps; ++i) {
        e = hipEventRecord(ev[i], b->stream);
        if (e == hipSuccess) st = enqueue_batch_step(b);
    }
    if (e == hipSuccess && st == FS_OK) e = hipEventRecord(ev[nsteps], b->stream);
    if (e == hipSuccess && st == FS_OK) e = hipStreamSynchronize(b->stream);
    for (int i = 0; e == hipSuccess && st == FS_OK && i < nsteps; ++i) {
        float ms = 0.0f;
        e = hipEventElapsedTime(&ms, ev[i], ev[i + 1]);
        if (stepMs) stepMs[i] = ms;
    }
    for (auto& x : ev)
        if (x) (void)hipEventDestroy(x);
    if (st != FS_OK) return st;
    return e == hipSuccess ? FS_OK : b->hip_fail("batch step latency", e);
}

int fs_batch_synchronize(fs_batch* b) {
    if (!b) return fsi::global_fail(FS_ERR_NOT_INITIALIZED, "null batch");
    if (int st = set_device(b)) return st;
    const hipError_t e = hipStreamSynchronize(b->stream);
    return e == hipSuccess ? FS_OK : b->hip_fail("hipStreamSynchronize", e);
}

const char* fs_batch_last_error(const fs_batch* b) { return b ? b->lastError.c_str() : "null batch"; }

void* fs_batch_stream(fs_batch* b) { return b ? static_cast<void*>(b->stream) : nullptr; }

}  // extern "C"
