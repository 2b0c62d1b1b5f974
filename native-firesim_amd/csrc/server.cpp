// server.cpp - continuous batching of simulation requests (fs_server_* in
// include/firesim.h).
//
// The reference's worker (runCudaWorker src/main.cpp:3426-3677) keeps one
// simulation resident and steps/renders it on a fixed cadence. The server keeps
// many resident at once: requests are admitted by HBM budget into one fs_batch,
// every admitted request advances one step per scheduling round (one batched
// launch per pass), frames are rendered asynchronously into pinned host
// buffers, and finished requests leave the batch while others keep running.
// Device memory of retired requests is recycled for later ones (hipFree
// synchronises the device, so it is only called when the cache must shrink).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "scene_impl.hpp"

namespace {

using Clock = std::chrono::steady_clock;

constexpr int kFrameBuffers = 3;  // at most two rounds in flight, so one finished frame is always stable

struct Request {
    uint64_t id = 0;
    fs_request r{};
    int state = FS_REQ_PENDING;
    fs_scene* scene = nullptr;
    void* memory = nullptr;  // recycled block (null when the scene owns its memory)
    uint64_t memoryBytes = 0;
    uint64_t sceneBytes = 0;
    uint64_t cells = 0;
    int enqueued = 0;   // steps enqueued
    int completed = 0;  // steps observed complete
    int framesEnqueued = 0;
    int frames = 0;
    uint32_t* frame[kFrameBuffers] = {nullptr, nullptr, nullptr};
    int frameStep = -1;  // step index of the newest complete frame
    int frameBuf = -1;
    std::vector<uint32_t> finalFrame;  // newest frame, kept after the pinned buffers are recycled
    double arrival = 0.0, admitted = -1.0, finished = -1.0;
    double readyAt = 0.0;
    std::vector<double> lat;
    std::string error;
};

struct Round {
    hipEvent_t done = nullptr;
    std::vector<uint64_t> stepped;                  // requests advanced in this round
    std::vector<std::pair<uint64_t, int>> renders;  // (request, frame buffer) rendered after the step
    std::vector<int> renderSteps;
};

double pct(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * static_cast<double>(v.size() - 1);
    const size_t i = static_cast<size_t>(pos);
    const double f = pos - static_cast<double>(i);
    return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

}  // namespace

struct fs_server {
    fs_server_desc desc{};
    fs_batch* batch = nullptr;
    // Every entry point holds mu while it reads or changes server state, so
    // requests may be submitted (and status read) from other threads while one
    // thread runs the scheduling loop. fs_server_run releases it while it sleeps.
    std::mutex mu;
    std::map<uint64_t, Request> reqs;
    std::multimap<double, uint64_t> pending;  // by arrival
    std::deque<uint64_t> waiting;             // arrived, FIFO
    std::vector<uint64_t> active;             // admitted, in the batch
    std::vector<uint64_t> draining;           // all steps enqueued, not yet observed complete
    std::deque<Round> inflight;
    std::vector<hipEvent_t> eventPool;
    std::multimap<uint64_t, void*> cache;     // recycled device blocks by size
    std::multimap<size_t, uint32_t*> framePool;  // recycled pinned frame buffers by size
    uint64_t cachedBytes = 0;
    uint64_t usedBytes = 0;
    uint64_t peakBytes = 0;
    uint64_t nextId = 1;
    bool started = false;
    Clock::time_point epoch;
    bool tuned = false;
    int64_t roundsStepped = 0;
    int64_t steps = 0, voxelSteps = 0, frames = 0;
    int64_t done = 0, failed = 0, submitted = 0;
    int maxActiveSeen = 0;
    double firstArrival = -1.0, lastCompletion = 0.0;
    std::vector<double> latencies;
    std::vector<double> queueMs;
    std::string lastError = "no error";

    int fail(int status, const std::string& msg) {
        lastError = msg;
        return status;
    }
    double now() {
        if (!started) {
            started = true;
            epoch = Clock::now();
        }
        return std::chrono::duration<double>(Clock::now() - epoch).count();
    }
};

namespace {

hipEvent_t take_event(fs_server* sv) {
    if (!sv->eventPool.empty()) {
        hipEvent_t e = sv->eventPool.back();
        sv->eventPool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    return e;
}

// Device memory for a scene: a recycled block of 1x..1.5x the size, else a
// new allocation (shrinking the cache first if the budget requires it).
void* get_block(fs_server* sv, uint64_t bytes, uint64_t* got) {
    auto it = sv->cache.lower_bound(bytes);
    if (it != sv->cache.end() && it->first <= bytes + bytes / 2) {
        void* p = it->second;
        *got = it->first;
        sv->cachedBytes -= it->first;
        sv->cache.erase(it);
        return p;
    }
    while (!sv->cache.empty() && sv->usedBytes + sv->cachedBytes + bytes > sv->desc.hbm_budget_bytes) {
        auto big = std::prev(sv->cache.end());
        (void)hipFree(big->second);
        sv->cachedBytes -= big->first;
        sv->cache.erase(big);
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    *got = bytes;
    return p;
}

void put_block(fs_server* sv, void* p, uint64_t bytes) {
    if (!p) return;
    sv->cache.emplace(bytes, p);
    sv->cachedBytes += bytes;
}

size_t frame_bytes(const Request& q) { return static_cast<size_t>(q.r.desc.frame_w) * q.r.desc.frame_h * 4; }

// Pinned frame buffers come from a pool: hipHostMalloc is slow and would
// stall the scheduling loop at every admission.
bool get_frames(fs_server* sv, Request& q) {
    const size_t bytes = frame_bytes(q);
    for (auto& f : q.frame) {
        auto it = sv->framePool.find(bytes);
        if (it != sv->framePool.end()) {
            f = it->second;
            sv->framePool.erase(it);
        } else if (hipHostMalloc(reinterpret_cast<void**>(&f), bytes, hipHostMallocDefault) != hipSuccess) {
            f = nullptr;
            return false;
        }
    }
    return true;
}

void put_frames(fs_server* sv, Request& q) {
    for (auto& f : q.frame) {
        if (f) sv->framePool.emplace(frame_bytes(q), f);
        f = nullptr;
    }
}

// Creates the scene of an arrived request; false if it does not fit now.
bool admit(fs_server* sv, Request& q, double t) {
    fs_scene_desc d = q.r.desc;
    d.device = sv->desc.device;
    if (static_cast<int>(sv->active.size() + sv->draining.size()) >= sv->desc.max_active) return false;
    if (sv->usedBytes + q.sceneBytes > sv->desc.hbm_budget_bytes) return false;
    if (sv->desc.max_round_voxels > 0 && !sv->active.empty()) {
        uint64_t cells = q.cells;
        for (uint64_t id : sv->active) cells += sv->reqs[id].cells;
        if (cells > sv->desc.max_round_voxels) return false;
    }
    // Nothing here waits for the device: the scene is created without its
    // initial reset, joins the batch, and is reset on the batch stream, in
    // order before its first step.
    int st = FS_OK;
    if (q.r.keep_scene) {
        st = fsi::create_unreset(&d, nullptr, 0, &q.scene);
    } else {
        q.memory = get_block(sv, q.sceneBytes, &q.memoryBytes);
        st = q.memory ? fsi::create_unreset(&d, q.memory, q.memoryBytes, &q.scene) : FS_ERR_OUT_OF_MEMORY;
    }
    if (st == FS_OK) st = fs_set_params(q.scene, &q.r.params);
    if (st == FS_OK) st = fs_set_emitters(q.scene, q.r.emitters, q.r.emitter_count);
    if (st == FS_OK && q.r.render_every > 0 && !get_frames(sv, q)) st = FS_ERR_OUT_OF_MEMORY;
    if (st == FS_OK) st = fs_batch_add(sv->batch, q.scene);
    if (st == FS_OK) st = fsi::reset_async(q.scene);
    if (st != FS_OK) {
        q.error = q.scene ? fs_last_error(q.scene) : fs_last_global_error();
        if (q.scene) fs_destroy(q.scene);
        q.scene = nullptr;
        put_block(sv, q.memory, q.memoryBytes);
        q.memory = nullptr;
        put_frames(sv, q);
        q.state = FS_REQ_FAILED;
        q.finished = t;
        ++sv->failed;
        return true;  // leaves the queue either way
    }
    q.state = FS_REQ_RUNNING;
    q.admitted = t;
    q.readyAt = t;  // step latency counts from admission; the wait before it is the queue time
    sv->usedBytes += q.sceneBytes;
    sv->peakBytes = std::max(sv->peakBytes, sv->usedBytes);
    sv->queueMs.push_back((t - q.arrival) * 1e3);
    sv->active.push_back(q.id);
    return true;
}

void retire(fs_server* sv, Request& q, double t) {
    q.state = FS_REQ_DONE;
    q.finished = t;
    ++sv->done;
    if (q.frameBuf >= 0) {
        const uint32_t* f = q.frame[q.frameBuf];
        q.finalFrame.assign(f, f + frame_bytes(q) / 4);
    }
    put_frames(sv, q);
    if (!q.r.keep_scene) {
        fs_destroy(q.scene);  // its work completed: no wait
        q.scene = nullptr;
        put_block(sv, q.memory, q.memoryBytes);
        q.memory = nullptr;
        sv->usedBytes -= q.sceneBytes;
    }
}

// Books a completed round: per-request step latencies and frames.
void complete_round(fs_server* sv, Round& rd, double t) {
    for (uint64_t id : rd.stepped) {
        Request& q = sv->reqs[id];
        const double ms = (t - q.readyAt) * 1e3;
        q.lat.push_back(ms);
        sv->latencies.push_back(ms);
        q.readyAt = t;
        ++q.completed;
        ++sv->steps;
        sv->voxelSteps += static_cast<int64_t>(q.cells);
    }
    for (size_t i = 0; i < rd.renders.size(); ++i) {
        Request& q = sv->reqs[rd.renders[i].first];
        q.frameBuf = rd.renders[i].second;
        q.frameStep = rd.renderSteps[i];
        ++q.frames;
        ++sv->frames;
    }
    sv->lastCompletion = t;
    // requests whose last step has now completed
    for (size_t i = 0; i < sv->draining.size();) {
        Request& q = sv->reqs[sv->draining[i]];
        if (q.completed == q.r.steps) {
            retire(sv, q, t);
            sv->draining.erase(sv->draining.begin() + static_cast<long>(i));
        } else {
            ++i;
        }
    }
    sv->eventPool.push_back(rd.done);
    rd.done = nullptr;
}

// Observes finished rounds; blocks on the oldest while `keep` or more are in flight.
int drain(fs_server* sv, size_t keep) {
    while (!sv->inflight.empty()) {
        Round& rd = sv->inflight.front();
        hipError_t e = hipEventQuery(rd.done);
        if (e == hipErrorNotReady && sv->inflight.size() >= keep && keep > 0) e = hipEventSynchronize(rd.done);
        if (e == hipErrorNotReady) break;
        if (e != hipSuccess) return sv->fail(FS_ERR_HIP, std::string("round completion: ") + hipGetErrorString(e));
        complete_round(sv, rd, sv->now());
        sv->inflight.pop_front();
    }
    return FS_OK;
}

bool all_done(const fs_server* sv) { return sv->done + sv->failed == sv->submitted; }

}  // namespace

extern "C" {

int fs_server_create(const fs_server_desc* desc, fs_server** out) {
    if (!out) return fsi::global_fail(FS_ERR_INVALID_ARGUMENT, "null output");
    *out = nullptr;
    if (!desc) return fsi::global_fail(FS_ERR_INVALID_ARGUMENT, "null server descriptor");
    if (desc->max_active < 1 || desc->max_active > FS_MAX_BATCH) return fsi::global_fail(FS_ERR_INVALID_ARGUMENT, "max_active must be in [1, FS_MAX_BATCH]");
    if (desc->hbm_budget_bytes == 0) return fsi::global_fail(FS_ERR_INVALID_ARGUMENT, "hbm_budget_bytes must be positive");
    if (desc->sor_kernel < -1 || desc->sor_kernel >= FS_SOR_KERNEL_COUNT) return fsi::global_fail(FS_ERR_INVALID_ARGUMENT, "unknown SOR kernel");
    fs_server* sv = new (std::nothrow) fs_server;
    if (!sv) return fsi::global_fail(FS_ERR_OUT_OF_MEMORY, "server handle");
    sv->desc = *desc;
    if (int st = fs_batch_create(desc->device, &sv->batch)) {
        delete sv;
        return st;
    }
    fs_tuning t{desc->sor_kernel < 0 ? FS_SOR_TB_T24_Z10_K4 : desc->sor_kernel, desc->use_graph};
    fs_batch_set_tuning(sv->batch, &t);
    sv->tuned = desc->sor_kernel >= 0;
    *out = sv;
    return FS_OK;
}

int fs_server_destroy(fs_server* sv) {
    if (!sv) return FS_OK;
    (void)hipSetDevice(sv->desc.device);
    (void)fs_batch_synchronize(sv->batch);
    for (auto& kv : sv->reqs) {
        Request& q = kv.second;
        if (q.scene) fs_destroy(q.scene);
        q.scene = nullptr;
        if (q.memory) (void)hipFree(q.memory);
        for (auto& f : q.frame)
            if (f) (void)hipHostFree(f);
    }
    for (auto& kv : sv->framePool) (void)hipHostFree(kv.second);
    for (auto& rd : sv->inflight)
        if (rd.done) (void)hipEventDestroy(rd.done);
    for (hipEvent_t e : sv->eventPool) (void)hipEventDestroy(e);
    for (auto& kv : sv->cache) (void)hipFree(kv.second);
    fs_batch_destroy(sv->batch);
    delete sv;
    return FS_OK;
}

int fs_server_submit(fs_server* sv, const fs_request* r, uint64_t* id) {
    if (!sv) return fsi::global_fail(FS_ERR_NOT_INITIALIZED, "null server");
    if (!r) return sv->fail(FS_ERR_INVALID_ARGUMENT, "null request");
    if (r->steps < 1) return sv->fail(FS_ERR_INVALID_ARGUMENT, "steps must be >= 1");
    if (r->render_every < 0) return sv->fail(FS_ERR_INVALID_ARGUMENT, "render_every must be >= 0");
    if (r->emitter_count < 0 || r->emitter_count > FS_MAX_EMITTERS) return sv->fail(FS_ERR_INVALID_ARGUMENT, "emitter_count out of range");
    fs_scene_desc d = r->desc;
    d.device = sv->desc.device;
    uint64_t bytes = 0;
    if (fs_scene_bytes(&d, &bytes) != FS_OK) return sv->fail(FS_ERR_INVALID_ARGUMENT, fs_last_global_error());
    if (bytes > sv->desc.hbm_budget_bytes) return sv->fail(FS_ERR_INVALID_ARGUMENT, "request exceeds the whole HBM budget");
    std::lock_guard<std::mutex> lock(sv->mu);
    Request q;
    q.id = sv->nextId++;
    q.r = *r;
    q.r.desc.device = sv->desc.device;
    q.sceneBytes = bytes;
    q.cells = static_cast<uint64_t>(d.nx) * d.ny * d.nz;
    q.arrival = r->arrival_s < 0.0 ? sv->now() : r->arrival_s;
    sv->pending.emplace(q.arrival, q.id);
    sv->reqs.emplace(q.id, std::move(q));
    ++sv->submitted;
    if (id) *id = sv->nextId - 1;
    return FS_OK;
}

}  // extern "C"

namespace {

// One scheduling round; the caller holds sv->mu.
int tick_locked(fs_server* sv, int32_t* activeOut) {
    if (activeOut) *activeOut = 0;
    if (hipSetDevice(sv->desc.device) != hipSuccess) return sv->fail(FS_ERR_HIP, "hipSetDevice");
    if (int st = drain(sv, 2)) return st;
    const double t = sv->now();
    while (!sv->pending.empty() && sv->pending.begin()->first <= t) {
        const uint64_t id = sv->pending.begin()->second;
        sv->pending.erase(sv->pending.begin());
        sv->reqs[id].state = FS_REQ_WAITING;
        if (sv->firstArrival < 0.0) sv->firstArrival = sv->reqs[id].arrival;
        sv->waiting.push_back(id);
    }
    while (!sv->waiting.empty() && admit(sv, sv->reqs[sv->waiting.front()], t)) sv->waiting.pop_front();
    if (sv->active.empty()) return FS_OK;
    uint64_t activeCells = 0;
    for (uint64_t id : sv->active) activeCells += sv->reqs[id].cells;
    // tune on a representative mix: once 16M cells are resident (or after 64 rounds)
    if (!sv->tuned && (activeCells >= (16ull << 20) || sv->roundsStepped >= 64)) {
        fs_autotune_report rep;
        if (int st = fs_batch_autotune(sv->batch, 1, &rep)) return sv->fail(st, fs_batch_last_error(sv->batch));
        sv->tuned = true;
    }
    Round rd;
    rd.done = take_event(sv);
    if (!rd.done) return sv->fail(FS_ERR_HIP, "event create");
    if (int st = fs_batch_step(sv->batch, 1)) {
        sv->eventPool.push_back(rd.done);
        return sv->fail(st, fs_batch_last_error(sv->batch));
    }
    rd.stepped = sv->active;
    for (size_t i = 0; i < sv->active.size();) {
        Request& q = sv->reqs[sv->active[i]];
        ++q.enqueued;
        if (q.r.render_every > 0 && (q.enqueued % q.r.render_every == 0 || q.enqueued == q.r.steps)) {
            const int buf = q.framesEnqueued % kFrameBuffers;
            if (int st = fs_render_bgra8_async(q.scene, q.frame[buf])) return sv->fail(st, fs_last_error(q.scene));
            ++q.framesEnqueued;
            rd.renders.emplace_back(q.id, buf);
            rd.renderSteps.push_back(q.enqueued);
        }
        if (q.enqueued == q.r.steps) {
            // out of the batch now; its memory is released once this round completes
            if (int st = fs_batch_remove(sv->batch, q.scene)) return sv->fail(st, fs_batch_last_error(sv->batch));
            sv->draining.push_back(q.id);
            sv->active.erase(sv->active.begin() + static_cast<long>(i));
        } else {
            ++i;
        }
    }
    const hipError_t e = hipEventRecord(rd.done, static_cast<hipStream_t>(fs_batch_stream(sv->batch)));
    if (e != hipSuccess) return sv->fail(FS_ERR_HIP, std::string("round event: ") + hipGetErrorString(e));
    sv->maxActiveSeen = std::max(sv->maxActiveSeen, static_cast<int>(rd.stepped.size()));
    ++sv->roundsStepped;
    if (activeOut) *activeOut = static_cast<int32_t>(rd.stepped.size());
    sv->inflight.push_back(std::move(rd));
    return FS_OK;
}

}  // namespace

extern "C" {

int fs_server_tick(fs_server* sv, int32_t* activeOut) {
    if (!sv) return fsi::global_fail(FS_ERR_NOT_INITIALIZED, "null server");
    std::lock_guard<std::mutex> lock(sv->mu);
    return tick_locked(sv, activeOut);
}

int fs_server_run(fs_server* sv) {
    if (!sv) return fsi::global_fail(FS_ERR_NOT_INITIALIZED, "null server");
    for (;;) {
        double wait = 0.0;
        {
            std::lock_guard<std::mutex> lock(sv->mu);
            int32_t n = 0;
            if (int st = tick_locked(sv, &n)) return st;
            if (n > 0) continue;
            if (!sv->inflight.empty()) {
                if (int st = drain(sv, 1)) return st;  // wait for the oldest round
                continue;
            }
            if (all_done(sv)) return FS_OK;
            if (sv->pending.empty()) {
                if (!sv->waiting.empty()) return sv->fail(FS_ERR_OUT_OF_MEMORY, "a waiting request cannot be admitted");
                return FS_OK;
            }
            wait = sv->pending.begin()->first - sv->now();
        }
        // idle until the next arrival, without the lock so submissions get in
        if (wait > 0.0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.05)));
    }
}

int fs_server_status(fs_server* sv, uint64_t id, fs_request_status* out) {
    if (!sv) return fsi::global_fail(FS_ERR_NOT_INITIALIZED, "null server");
    std::lock_guard<std::mutex> lock(sv->mu);
    if (!out) return sv->fail(FS_ERR_INVALID_ARGUMENT, "null output");
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return sv->fail(FS_ERR_INVALID_ARGUMENT, "unknown request id");
    const Request& q = it->second;
    fs_request_status s{};
    s.state = q.state;
    s.steps_done = q.completed;
    s.frames = q.frames;
    s.frame_w = q.r.desc.frame_w;
    s.frame_h = q.r.desc.frame_h;
    s.arrival_s = q.arrival;
    s.admitted_s = q.admitted;
    s.finished_s = q.finished;
    s.step_ms_p50 = pct(q.lat, 0.50);
    s.step_ms_p99 = pct(q.lat, 0.99);
    s.step_ms_max = q.lat.empty() ? 0.0 : *std::max_element(q.lat.begin(), q.lat.end());
    *out = s;
    if (q.state == FS_REQ_FAILED) sv->lastError = q.error;
    return FS_OK;
}

int fs_server_frame(fs_server* sv, uint64_t id, uint32_t* bgra, int32_t* frameStep) {
    if (!sv) return fsi::global_fail(FS_ERR_NOT_INITIALIZED, "null server");
    std::lock_guard<std::mutex> lock(sv->mu);
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return sv->fail(FS_ERR_INVALID_ARGUMENT, "unknown request id");
    const Request& q = it->second;
    if (q.frameBuf < 0) return sv->fail(FS_ERR_INVALID_ARGUMENT, "no frame rendered yet");
    const uint32_t* src = q.frame[q.frameBuf] ? q.frame[q.frameBuf] : q.finalFrame.data();
    if (bgra) std::memcpy(bgra, src, frame_bytes(q));
    if (frameStep) *frameStep = q.frameStep;
    return FS_OK;
}

int fs_server_take_scene(fs_server* sv, uint64_t id, fs_scene** out) {
    if (!sv) return fsi::global_fail(FS_ERR_NOT_INITIALIZED, "null server");
    std::lock_guard<std::mutex> lock(sv->mu);
    if (!out) return sv->fail(FS_ERR_INVALID_ARGUMENT, "null output");
    auto it = sv->reqs.find(id);
    if (it == sv->reqs.end()) return sv->fail(FS_ERR_INVALID_ARGUMENT, "unknown request id");
    Request& q = it->second;
    if (q.state != FS_REQ_DONE || !q.scene) return sv->fail(FS_ERR_INVALID_ARGUMENT, "request is not done or its scene was not kept");
    *out = q.scene;
    q.scene = nullptr;
    sv->usedBytes -= q.sceneBytes;
    return FS_OK;
}

int fs_server_stats_get(fs_server* sv, fs_server_stats* out) {
    if (!sv) return fsi::global_fail(FS_ERR_NOT_INITIALIZED, "null server");
    std::lock_guard<std::mutex> lock(sv->mu);
    if (!out) return sv->fail(FS_ERR_INVALID_ARGUMENT, "null output");
    fs_server_stats s{};
    s.requests_submitted = sv->submitted;
    s.requests_done = sv->done;
    s.requests_failed = sv->failed;
    s.steps = sv->steps;
    s.voxel_steps = sv->voxelSteps;
    s.rounds = sv->roundsStepped;
    s.frames = sv->frames;
    s.max_active_seen = sv->maxActiveSeen;
    fs_tuning t{};
    fs_batch_get_tuning(sv->batch, &t);
    s.sor_kernel = t.sor_kernel;
    s.peak_resident_bytes = sv->peakBytes;
    s.busy_s = sv->firstArrival >= 0.0 ? sv->lastCompletion - sv->firstArrival : 0.0;
    s.step_ms_p50 = pct(sv->latencies, 0.50);
    s.step_ms_p99 = pct(sv->latencies, 0.99);
    double sum = 0.0, mx = 0.0;
    for (double v : sv->latencies) {
        sum += v;
        mx = std::max(mx, v);
    }
    s.step_ms_mean = sv->latencies.empty() ? 0.0 : sum / static_cast<double>(sv->latencies.size());
    s.step_ms_max = mx;
    s.queue_ms_p50 = pct(sv->queueMs, 0.50);
    s.queue_ms_p99 = pct(sv->queueMs, 0.99);
    *out = s;
    return FS_OK;
}

int fs_server_step_latencies(fs_server* sv, double* outMs, int64_t capacity, int64_t* count) {
    if (!sv) return fsi::global_fail(FS_ERR_NOT_INITIALIZED, "null server");
    std::lock_guard<std::mutex> lock(sv->mu);
    const int64_t n = static_cast<int64_t>(sv->latencies.size());
    if (count) *count = n;
    if (outMs) std::copy(sv->latencies.begin(), sv->latencies.begin() + std::min(n, std::max<int64_t>(capacity, 0)), outMs);
    return FS_OK;
}

double fs_server_clock(fs_server* sv) {
    if (!sv) return 0.0;
    std::lock_guard<std::mutex> lock(sv->mu);
    return sv->now();
}

const char* fs_server_last_error(const fs_server* sv) { return sv ? sv->lastError.c_str() : "null server"; }

}  // extern "C"
