// scene_impl.hpp - the scene handle behind include/firesim.h, shared by the
// per-scene ABI (scene.cpp) and batches (batch.cpp). Not part of the ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "../../include/firesim.h"
#include "fs_launch.hpp"

using fs::DeviceScene;
using fs::DevParams;

struct fs_batch;

// Executable graphs that may still be running. A graph dropped while a launch
// of it is queued or running on a stream is not destroyed at once: retire()
// records an event on that stream behind the launch, and the graph is
// destroyed by a later reap() once the event has completed (reap(true) waits
// for it). So no hipGraphExecDestroy ever races a launch in flight.
class GraphGraveyard {
public:
    void retire(hipGraphExec_t g, hipStream_t st) {
        if (!g) return;
        hipEvent_t ev = nullptr;
        if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess || hipEventRecord(ev, st) != hipSuccess) {
            // no event: fall back to waiting for the stream before destroying
            (void)hipGetLastError();
            if (ev) (void)hipEventDestroy(ev);
            (void)hipStreamSynchronize(st);
            (void)hipGraphExecDestroy(g);
            return;
        }
        dead_.push_back({g, ev});
    }
    void reap(bool wait) {
        size_t keep = 0;
        for (size_t i = 0; i < dead_.size(); ++i) {
            Entry& e = dead_[i];
            const hipError_t q = wait ? hipEventSynchronize(e.done) : hipEventQuery(e.done);
            if (q == hipErrorNotReady) {
                dead_[keep++] = e;
                continue;
            }
            (void)hipGraphExecDestroy(e.exec);
            (void)hipEventDestroy(e.done);
        }
        dead_.resize(keep);
    }
    size_t pending() const { return dead_.size(); }

private:
    struct Entry {
        hipGraphExec_t exec;
        hipEvent_t done;
    };
    std::vector<Entry> dead_;
};

struct Dims {
    int nx, ny, nz, lnx, lny, lnz, fw, fh;
    bool bf16;
    size_t cells() const { return static_cast<size_t>(nx) * ny * nz; }
    size_t lcells() const { return static_cast<size_t>(lnx) * lny * lnz; }
    size_t pixels() const { return static_cast<size_t>(fw) * fh; }
    size_t partials() const {
        return std::max(fs::react_partial_count(nx, ny, nz), static_cast<size_t>(fs::kMetricBlocks) * fs::kMetricSlots);
    }
};

struct fs_scene {
    int device = 0;
    Dims dims{};
    hipStream_t stream = nullptr;     // stream all of the scene's work is issued on
    hipStream_t ownStream = nullptr;  // the scene's own stream (stream == ownStream outside a batch)
    fs_batch* batch = nullptr;        // batch the scene belongs to, if any
    void* pool = nullptr;
    size_t poolBytes = 0;
    bool ownsPool = true;  // false: fs_create_in memory, left to the caller
    DeviceScene ds;
    fs_params prm{};
    std::vector<fs_emitter> emitters;
    float time = 0.0f, renderTime = 0.0f;
    int frameIndex = 0;
    std::string lastError = "no error";
    // Step graphs. Scalar fields and (for an odd number of SOR launches) the
    // pressure buffers change places every step, so two one-step graphs are
    // captured, from the buffer assignment before an even and an odd step;
    // state[g] is the assignment graph g starts from.
    bool useGraph = true;
    bool graphsValid = false;
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    DeviceScene state[2];
    int parity = 0;  // graph to run next
    GraphGraveyard graveyard;  // dropped step graphs whose last launch may still run

    int fail(int status, const std::string& msg) {
        lastError = msg;
        return status;
    }
    int hip_fail(const char* what, hipError_t e) { return fail(FS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e)); }
};

namespace fsi {
// fs_create / fs_create_in (memory may be null) without the initial reset:
// the fields hold garbage until reset_async or fs_reset.
int create_unreset(const fs_scene_desc* desc, void* memory, uint64_t bytes, fs_scene** out);
// fs_reset without the final wait: enqueued on the scene's current stream.
int reset_async(fs_scene* s);
// Kernel parameter block of a scene's current parameters and emitters.
DevParams dev_params(const fs_scene& sc);
// Releases the scene's captured step graphs (they are recaptured on demand).
void drop_graphs(fs_scene* s);
// Sets the thread's handle-less error message; returns status.
int global_fail(int status, const std::string& msg);
}  // namespace fsi
