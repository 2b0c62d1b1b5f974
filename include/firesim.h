/*
 * firesim.h - C ABI of the MI355X (gfx950) fire/smoke step.
 *
 * The reference exposes the step as C++ free functions over one process-wide
 * singleton (src/fire_cuda.h:103-119). This ABI exposes the same operations
 * over explicit per-scene handles so that many scenes can live on one GPU:
 *
 *   reference (src/fire_cuda.h)          this ABI
 *   fireCudaInitialize      :103   ->    fs_create (+ fs_set_params)
 *   fireCudaReset           :116   ->    fs_reset
 *   fireCudaStepD3D11       :110   ->    fs_step            (step only, no render)
 *   fireCudaStepAndRender   :104   ->    fs_step + fs_render_bgra8
 *   fireCudaStepAndRenderMeasured :105 -> fs_step + fs_render_bgra8 + fs_metrics
 *   fireCudaRenderD3D11     :113   ->    fs_render_hdr      (host float4 buffer, no D3D11)
 *   fireCudaSynchronize     :115   ->    fs_synchronize
 *   fireCudaShutdown        :118   ->    fs_destroy
 *   fireCudaLastError       :119   ->    fs_last_error
 *   FireSettings       :65-101     ->    fs_params (+ fs_emitter list)
 *   FireCudaFrameMetrics :16-63    ->    fs_metrics
 *
 * Conventions: every entry point returns an int status (FS_OK == 0) and never
 * throws; host buffers belong to the caller; device memory belongs to the
 * handle. Different handles may be driven from different threads; one handle
 * must be driven by one thread at a time. All work of a handle runs in order
 * on that handle's own HIP stream.
 */
#ifndef FIRESIM_H
#define FIRESIM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FS_ABI_VERSION 2

enum fs_status {
    FS_OK = 0,
    FS_ERR_INVALID_ARGUMENT = 1,
    FS_ERR_NOT_INITIALIZED = 2,
    FS_ERR_HIP = 3,
    FS_ERR_OUT_OF_MEMORY = 4,
    FS_ERR_UNSUPPORTED = 5,
    FS_ERR_IO = 6
};

/* Scalar, face and light fields (reference globals src/fire_cuda.cu:116-151).
 * Scalars are nx*ny*nz, x fastest: idx = (z*ny + y)*nx + x (src/fire_cuda.cu:449).
 * U is (nx+1)*ny*nz, V is nx*(ny+1)*nz, W is nx*ny*(nz+1) (src/fire_cuda.cu:456-487).
 * LIGHT is 4 floats (rgb radiance, extinction) per light cell, SHADOW 1 float. */
enum fs_field_id {
    FS_FIELD_HEAT = 0,
    FS_FIELD_FUEL = 1,
    FS_FIELD_OXYGEN = 2,
    FS_FIELD_SOOT = 3,
    FS_FIELD_CHAR = 4,
    FS_FIELD_ASH = 5,
    FS_FIELD_PYROLYSIS = 6,
    FS_FIELD_PROGRESS = 7,
    FS_FIELD_TURBULENCE = 8,
    FS_FIELD_SOOT_OPTICS = 9,
    FS_FIELD_U = 10,
    FS_FIELD_V = 11,
    FS_FIELD_W = 12,
    FS_FIELD_PRESSURE = 13,
    FS_FIELD_DIVERGENCE = 14,
    FS_FIELD_LIGHT = 15,
    FS_FIELD_SHADOW = 16,
    FS_FIELD_DIR_TRANSMITTANCE = 17, /* light grid; written by the light stage while dir_light is on */
    FS_FIELD_COUNT = 18
};

/* Passes of one step, in the order fs_step runs them (src/fire_cuda.cu:3196-3318). */
enum fs_stage {
    FS_STAGE_ADVECT_VELOCITY = 0, /* K5-K7  advectU/V/WKernel :880-932 */
    FS_STAGE_ADVECT_REACT = 1,    /* K8     advectReactKernel :934-1126 */
    FS_STAGE_FORCE = 2,           /* K9-K11 force (gather form), boundary clamp, divergence :1128-1199,:1261-1291 */
    FS_STAGE_PRESSURE = 3,        /* K12    red-black SOR, sor_iterations sweeps :1201-1227 */
    FS_STAGE_PROJECT = 4,         /* K13-K14 gradient subtract + wall zeroing :1229-1245,:1293-1317 */
    FS_STAGE_LIGHT = 5,           /* K17-K18 light build + 2 propagations :1445-1565 */
    FS_STAGE_COUNT = 6
};

/* Storage precision of the advected scalar fields. */
enum fs_storage {
    FS_STORAGE_FP32 = 0, /* every field FP32 (reference) */
    FS_STORAGE_BF16 = 1  /* heat, fuel, soot stored BF16, all arithmetic FP32 */
};

/* Scene ids. 0..2 are the reference's procedural fuel beds
 * (fuelBedMaterial src/fire_cuda.cu:746-793); 3 has no bed, so only
 * fs_emitter spheres and the mouse sources inject fuel and heat. */
enum fs_scene_kind {
    FS_SCENE_ROOM_TRAY = 0,
    FS_SCENE_CAMPFIRE = 1,
    FS_SCENE_GAS_BURNER = 2,
    FS_SCENE_EMITTERS_ONLY = 3
};

#define FS_MAX_EMITTERS 16

typedef struct fs_scene fs_scene;

typedef struct fs_scene_desc {
    int32_t nx, ny, nz;        /* grid cells; each >= 4, (nx+1)(ny+1)(nz+1) <= 2^30 */
    int32_t frame_w, frame_h;  /* render target; each >= 1 */
    int32_t storage;           /* enum fs_storage */
    int32_t device;            /* HIP device ordinal */
} fs_scene_desc;

/* Per-step knobs. Field meaning follows FireSettings (src/fire_cuda.h:65-101);
 * fs_clamp_params applies the reference clamps (src/fire_cuda.cu:3135-3179)
 * with the extensions noted on each field. All fields are 4 bytes wide. */
typedef struct fs_params {
    float dt;                    /* [0.001, 1/30] */
    float mouse_x, mouse_y;      /* [0, 1] */
    int32_t left_down, right_down;
    int32_t scene_id;            /* enum fs_scene_kind, [0, 3] */
    float wind_x;                /* [-1, 1] (reference "wind", x only) */
    float wind_z;                /* [-1, 1] extension: same profile on w */
    float turbulence;            /* [0, 1.8] (reference floor 0.05 lifted to 0) */
    float detail;                /* [0.1, 1.5] */
    float smoke;                 /* [0, 1.6] */
    float intensity;             /* [0.2, 2.4] */
    float camera_yaw;
    float camera_pitch;          /* [-0.75, 0.75] */
    float camera_distance;       /* [1.5, 6] */
    int32_t cinematic_mode;      /* 0/1 */
    int32_t raymarch_steps;      /* [24, 256] (reference cap 120 lifted) */
    int32_t render_debug_mode;   /* [0, 6] */
    float exposure;              /* [0.45, 2.4] */
    float reflection_gain;       /* [0, 2.2] */
    float smoke_darkness;        /* [0.35, 2.4] */
    float emitter_center_x;      /* [-1.05, 1.05] */
    float emitter_center_z;      /* [-0.82, 0.82] */
    float emitter_height_norm;   /* [0, 0.96] */
    float emitter_height_band_norm; /* [0.006, 0.2] */
    float emitter_radius;        /* [0.035, 0.8] */
    int32_t burner_center_count; /* [0, 4] */
    float burner_center_x[4];
    float burner_center_y[4];
    float burner_center_z[4];
    float sor_omega;             /* [0.05, 1.95]; reference constant 1.52 (:105) */
    int32_t sor_iterations;      /* [1, 1000]; reference constant 40 (:104) */
    uint32_t noise_seed;         /* seed of the integer-hash noise */
    int32_t ember_count;         /* [0, 192] sparks splatted into the HDR frame (:2323-2452) */
    /* Extensions without a reference counterpart unless noted; 0 keeps the
     * reference behaviour. */
    int32_t closed_top;          /* 0/1: 1 makes the top v face a wall like the floor (the reference's top is open, :905-908, :1275-1281) */
    int32_t noise_forcing;       /* 0: the reference's fbm noise forcing (:1164-1165); 1: divergence-free curl noise of
                                  * amplitude curl_noise in its place */
    float curl_noise;            /* [0, 4]: with noise_forcing 1, faces accelerate by curl_noise * (curl of the potential)
                                  * per second; 0 gives no noise forcing at all */
    int32_t noise_mode;          /* 0: integer-hash noise (default, portable); 1: the reference's sin hash (:381-443),
                                  * a visual mode whose values depend on the platform's sinf */
    int32_t show_gizmos;         /* 0/1: draw the source ring and axis gizmos on the BGRA frame (drawGizmos :1931-1951) */
    int32_t active_gizmo;        /* any value; 1 highlights the source ring (FireSettings::activeGizmo) */
    int32_t dir_light;           /* 0/1: a directional key light scattered by the smoke and lighting the room */
    float dir_light_x, dir_light_y, dir_light_z; /* direction the light travels (normalized; zero gives straight down) */
    float dir_light_r, dir_light_g, dir_light_b; /* [0, 20] radiance */
} fs_params;

/* Spherical fuel + heat source in voxel units (cell i has centre i + 0.5).
 * Inside the sphere, each step adds fuel_rate*dt to fuel and raises heat to
 * (temperature_k - 293)/335 (the reference's heat <-> kelvin map, :1033). */
typedef struct fs_emitter {
    float cx, cy, cz, radius;
    float fuel_rate;
    float temperature_k;
} fs_emitter;

/* Frame metrics (FireCudaFrameMetrics src/fire_cuda.h:16-63), FP64 sums. */
typedef struct fs_metrics {
    int32_t frame_index;
    int32_t grid_x, grid_y, grid_z;
    int32_t pressure_iterations;
    int32_t invalid_cells;
    double time_seconds;
    double heat_sum, fuel_sum, oxygen_sum, soot_sum, char_sum, ash_sum;
    double pyrolysis_sum, progress_sum, turbulence_energy_sum, soot_optical_depth_sum;
    double max_heat, max_fuel, max_soot, max_pyrolysis, max_progress, max_turbulence_energy;
    double flame_height_meters;
    double mean_optical_depth, mean_scene_light, mean_scene_shadow;
    double flame_mass_proxy, smoke_mass_proxy, flame_smoke_overlap_proxy;
    double mean_volumetric_shadow, mean_room_irradiance, heat_release_proxy;
    double divergence_before_l2, divergence_before_max;
    double divergence_after_l2, divergence_after_max, divergence_reduction;
    /* Not in the reference: reaction heat release 7*burnMass summed over cells
     * (last step and since reset), and the FP64 SOR residual L2. */
    double heat_release_step;
    double heat_release_total;
    double sor_residual_l2;
} fs_metrics;

/* Per-pass GPU times of fs_step_timed, from HIP events recorded on the scene
 * stream between passes (the reference's measured path records the same kind
 * of per-pass cudaEvent intervals, src/fire_cuda.cu:3188-3491). */
typedef struct fs_stage_times {
    double stage_ms[FS_STAGE_COUNT]; /* summed over the timed steps */
    double total_ms;                 /* first pass start to last pass end */
    int32_t steps;
    int32_t pressure_launches;       /* kernel launches in one pressure pass */
} fs_stage_times;

/* Pressure-solve kernels. All give bit-identical results (same per-cell FP32
 * expression in the same red/black order); they differ in how many
 * half-sweeps one launch performs from an LDS-resident tile of p (d read from
 * global memory) and in the tile shape: x/y edge T, z depth TZ, K half-sweeps
 * per launch on a (T+2K)^2 x (TZ+2K) box. fs_autotune picks per grid. */
enum fs_sor_kernel {
    FS_SOR_HALF_SWEEP = 0,      /* one launch per half-sweep, as the reference (:3280-3284) */
    FS_SOR_TB_T16_K4 = 1,       /* 16^3 tiles, 4 half-sweeps per launch */
    FS_SOR_TB_T24_K4 = 2,       /* 24^3 tiles, 4 half-sweeps per launch */
    FS_SOR_TB_T24_Z8_K4 = 3,    /* 24x24x8 tiles */
    FS_SOR_TB_T24_Z16_K4 = 4,   /* 24x24x16 tiles */
    FS_SOR_TB_T24_Z28_K4 = 5,   /* 24x24x28 tiles */
    FS_SOR_TB_T24_Z10_K4 = 6,   /* 24x24x10 tiles, two workgroups per CU (default) */
    FS_SOR_TB_T24_Z12_K4 = 7,   /* 24x24x12 tiles, two workgroups per CU */
    FS_SOR_TB_T24_Z6_K4 = 8,    /* 24x24x6 tiles, for small grids */
    FS_SOR_TB_T26_Z13_K3 = 9,   /* 26x26x13 tiles, 3 half-sweeps per launch (5 tiles span 128 cells) */
    FS_SOR_TB_T26_Z9_K3 = 10,   /* 26x26x9 tiles, 3 half-sweeps per launch, two workgroups per CU */
    FS_SOR_TB_T22_Z8_K5 = 11,   /* 22x22x8 tiles, 5 half-sweeps per launch: fewest launches, for small grids */
    FS_SOR_KERNEL_COUNT = 12
};

/* Launch-shape choices of a scene (picked by the autotuner). */
typedef struct fs_tuning {
    int32_t sor_kernel; /* enum fs_sor_kernel */
    int32_t use_graph;  /* 1: fs_step replays a captured HIP graph of the step (default); 0: direct launches */
} fs_tuning;

/* Result of fs_autotune: time of one pressure pass per SOR kernel. */
typedef struct fs_autotune_report {
    double sor_ms[FS_SOR_KERNEL_COUNT];
    int32_t chosen; /* enum fs_sor_kernel now set on the scene */
} fs_autotune_report;

/* ---- host-only helpers (no device needed) ---- */
int fs_abi_version(void);
void fs_default_params(fs_params* out);
/* Applies the clamps documented on fs_params; in and out may alias. */
int fs_clamp_params(const fs_params* in, fs_params* out);
/* Light grid of a scene: clamp((n+3)/4, lo, hi) (src/fire_cuda.cu:2920-2922). */
int fs_light_dims(int32_t nx, int32_t ny, int32_t nz, int32_t* lnx, int32_t* lny, int32_t* lnz);
/* Device bytes fs_create will allocate for desc (admission by HBM budget). */
int fs_scene_bytes(const fs_scene_desc* desc, uint64_t* bytes);
/* 32-bit top-down BGRA BMP (reference writeBmp src/main.cpp:2511-2536). */
int fs_write_bmp(const char* path, const uint32_t* bgra, int32_t width, int32_t height);
/* Thread-local message of the last failed call that had no handle. */
const char* fs_last_global_error(void);

/* ---- device measurement ---- */
/* Achievable device-memory bandwidth of `device`: `reps` streaming float4
 * copies of `bytes` (a multiple of 16, 1 MiB .. 64 GiB) between two buffers
 * the call allocates and frees; *gbps = (read + write bytes) per second / 1e9,
 * from HIP events around the copies. No counterpart in the reference: the
 * achievable ceiling reported beside the nominal HBM peak. */
int fs_copy_bandwidth(int32_t device, uint64_t bytes, int32_t reps, double* gbps);

/* ---- pressure pass without allocation ----
 * No counterpart in the reference. For callers that cannot allocate device
 * memory: code under HIP graph capture, serving loops and harnesses that
 * forbid hipMalloc. The pressure pass of fs_run_stage(FS_STAGE_PRESSURE)
 * (red-black SOR, :1201-1227, the same per-cell FP32 update and red/black
 * order, so bit-identical results) on caller device buffers: p (nx*ny*nz
 * floats, x fastest) holds the initial guess and receives the solution, d
 * (nx*ny*nz floats) is the right-hand side. Enqueued on `stream` (a
 * hipStream_t; null is the null stream) of `device`, which becomes the
 * calling thread's current device; returns without waiting. The tile kernels
 * ping-pong with a library-owned static device array of FS_NOALLOC_MAX_CELLS
 * floats (64 MiB, held by every GPU the library runs on until the process
 * ends) instead of a second buffer; grids of more cells are rejected with
 * FS_ERR_INVALID_ARGUMENT. The call makes no hipMalloc, hipFree, memset or
 * other allocating call. Every call on one GPU shares that array: calls that
 * overlap on one GPU (on different streams, or in different threads without
 * ordering) race and give wrong results. */
#define FS_NOALLOC_MAX_CELLS (1u << 24)
int fs_pressure_solve_noalloc(float* p, const float* d, int32_t nx, int32_t ny, int32_t nz, float omega, int32_t iterations,
                              int32_t sor_kernel, int32_t device, void* stream);

/* ---- scene lifecycle ---- */
int fs_create(const fs_scene_desc* desc, fs_scene** out);
/* fs_create in caller-owned device memory (256-byte aligned, at least
 * fs_scene_bytes(desc) bytes), for callers that recycle memory across scenes;
 * fs_destroy leaves that memory to the caller. */
int fs_create_in(const fs_scene_desc* desc, void* device_memory, uint64_t bytes, fs_scene** out);
int fs_destroy(fs_scene* scene);
int fs_reset(fs_scene* scene);
int fs_set_params(fs_scene* scene, const fs_params* params);
int fs_get_params(const fs_scene* scene, fs_params* out);
/* The descriptor the scene was created with. */
int fs_get_desc(const fs_scene* scene, fs_scene_desc* out);
int fs_set_emitters(fs_scene* scene, const fs_emitter* emitters, int32_t count);
int fs_set_tuning(fs_scene* scene, const fs_tuning* tuning);
int fs_get_tuning(const fs_scene* scene, fs_tuning* out);
/* Times every SOR kernel on this scene's grid and current divergence (HIP
 * events, `reps` pressure passes each), selects the fastest and restores the
 * pressure field. The choice depends only on grid size and iteration count;
 * all kernels give identical results. */
int fs_autotune(fs_scene* scene, int32_t reps, fs_autotune_report* out);
const char* fs_last_error(const fs_scene* scene);

/* ---- stepping (asynchronous on the scene stream) ---- */
int fs_step(fs_scene* scene, int32_t nsteps);
/* Runs one pass with the current time, without advancing time. */
int fs_run_stage(fs_scene* scene, int32_t stage);
/* fs_step with per-pass event timing (at most 4096 steps per call); blocks
 * until the steps finish. step_ms, if not null, receives nsteps per-step GPU
 * times. Replaces fireCudaStepAndRenderMeasured's timing half (:105). */
int fs_step_timed(fs_scene* scene, int32_t nsteps, fs_stage_times* out, double* step_ms);
/* fs_step that also returns each step's GPU time (HIP events between steps,
 * production path, graph replay when enabled); blocks until done. */
int fs_step_latency(fs_scene* scene, int32_t nsteps, double* step_ms);
int fs_synchronize(fs_scene* scene);
int fs_set_time(fs_scene* scene, float time, float render_time);
int fs_get_time(const fs_scene* scene, float* time, float* render_time, int32_t* frame_index);

/* ---- rendering (synchronous: returns when the host buffer is written) ---- */
int fs_render_hdr(fs_scene* scene, float* rgba_host);        /* frame_w*frame_h*4 floats */
int fs_render_bgra8(fs_scene* scene, uint32_t* bgra_host);   /* frame_w*frame_h pixels */
/* fs_render_bgra8 without the wait: render, tonemap and the copy into
 * bgra_host are enqueued on the scene stream; the frame is complete once
 * fs_synchronize (or an event after this call) has passed. bgra_host should be
 * pinned (hipHostMalloc) for the copy to overlap other work. */
int fs_render_bgra8_async(fs_scene* scene, uint32_t* bgra_host);
/* The HDR frame as FP16 RGBA (4 halves per pixel: max(0, r|g|b) rounded to
 * nearest even, alpha 1.0), the surface format the reference writes instead
 * of the tonemap when it renders into a D3D11 texture (packFp16SurfaceKernel
 * src/fire_cuda.cu:2299-2321, dispatched at :3384-3403). */
int fs_render_rgba16f(fs_scene* scene, uint16_t* rgba16_host);   /* frame_w*frame_h*4 halves */
/* Tonemaps a caller HDR frame with the scene's exposure (K21 alone). */
int fs_tonemap(fs_scene* scene, const float* rgba_host, uint32_t* bgra_host);

/* ---- field access (synchronous) ---- */
int fs_field_count(const fs_scene* scene, int32_t field, uint64_t* count);
int fs_read_field(fs_scene* scene, int32_t field, float* host);
int fs_write_field(fs_scene* scene, int32_t field, const float* host);
int fs_metrics_collect(fs_scene* scene, fs_metrics* out);

/* ---- measurement support ---- */
/* Device stream of the scene (hipStream_t) for event timing by the caller. */
void* fs_stream(fs_scene* scene);

/* ---- reference scenes and frame statistics (host only) ---- */
/* Emitter and burner placement of reference scene scene_id (0 room, 1
 * campfire, 2 gas burner), read from <assets_dir>/<scene>/scene.json and
 * emitter-mask.json as the reference does (loadSceneEmitterParams
 * src/main.cpp:1397-1476), written into the emitter_* and burner_* fields of
 * *params (applySceneEmitterParams :433-447). Missing files give the
 * reference's built-in values. assets_dir null: "assets/fire-scenes". */
int fs_scene_emitter_params(int32_t scene_id, const char* assets_dir, fs_params* params);

/* Colour statistics of a BGRA frame (computeImageStats src/main.cpp:2635-2690). */
typedef struct fs_image_stats {
    float mean_luma, max_luma, mean_saturation;
    float warm_fire_fraction, tan_smoke_fraction, black_smoke_fraction, white_core_fraction;
    float bright_fraction;
    int32_t bright_pixels;
} fs_image_stats;
int fs_image_stats_of(const uint32_t* bgra, int64_t pixels, fs_image_stats* out);

/* ---- batches: many scenes, one launch per pass ----
 * No counterpart in the reference, which has one simulation per process
 * (src/fire_cuda.cu:116-195). A batch steps every member scene once per
 * fs_batch_step with one persistent launch per pass over all members (grids
 * of any size, any SOR iteration count, one storage mode per batch). Results
 * equal stepping each scene alone, bit for bit. While a scene is a member,
 * all of its work (including per-scene calls) is issued on the batch stream.
 * Membership and parameters may change between steps. */
typedef struct fs_batch fs_batch;

#define FS_MAX_BATCH 512

int fs_batch_create(int32_t device, fs_batch** out);
/* Removes every member (they keep their state) and frees the batch. */
int fs_batch_destroy(fs_batch* batch);
int fs_batch_add(fs_batch* batch, fs_scene* scene);
int fs_batch_remove(fs_batch* batch, fs_scene* scene);
int fs_batch_count(const fs_batch* batch, int32_t* count);
/* sor_kernel applies to every member; use_graph replays captured step graphs. */
int fs_batch_set_tuning(fs_batch* batch, const fs_tuning* tuning);
int fs_batch_get_tuning(const fs_batch* batch, fs_tuning* out);
/* Times each SOR kernel on the current members (HIP events, `reps` pressure
 * passes each), keeps the fastest and restores every member's pressure. */
int fs_batch_autotune(fs_batch* batch, int32_t reps, fs_autotune_report* out);
/* Advances every member by nsteps (asynchronous on the batch stream). */
int fs_batch_step(fs_batch* batch, int32_t nsteps);
/* fs_batch_step that blocks and returns each step's GPU time (HIP events). */
int fs_batch_step_latency(fs_batch* batch, int32_t nsteps, double* step_ms);
int fs_batch_synchronize(fs_batch* batch);
const char* fs_batch_last_error(const fs_batch* batch);
void* fs_batch_stream(fs_batch* batch);

/* ---- serving: continuous batching of simulation requests ----
 * Extends the reference's worker loop (runCudaWorker src/main.cpp:3426-3677:
 * resident state, step/render cadence, frame publish) from one simulation to
 * a stream of requests. Each scheduling round (fs_server_tick):
 *   1. takes note of finished GPU work and retires requests that are done;
 *   2. admits arrived requests in FIFO order while their device memory fits
 *      the HBM budget and a batch slot is free (state stays resident);
 *   3. advances every admitted request by one step in one batched launch per
 *      pass, and enqueues a frame render for requests at their render cadence.
 * Requests arriving while others run join at the next round, so their work
 * overlaps on the device. A request's results are the same as running it alone
 * (fs_batch guarantees bit-identical steps). The host keeps at most two rounds
 * in flight. Step latency of a request's k-th step is the time from when it
 * became due (arrival for the first step, completion of the previous step
 * otherwise) to the host observing its completion. */
typedef struct fs_server fs_server;

typedef struct fs_server_desc {
    int32_t device;
    int32_t max_active;         /* concurrently admitted requests, [1, FS_MAX_BATCH] */
    uint64_t hbm_budget_bytes;  /* device memory for resident requests (fs_scene_bytes each) */
    int32_t sor_kernel;         /* enum fs_sor_kernel; -1: autotune on the first rounds' members */
    int32_t use_graph;          /* replay captured batch step graphs */
    uint64_t max_round_voxels;  /* 0: no cap; else a request is admitted only while the admitted
                                 * requests' cells stay within it (a round's time, hence every
                                 * step's latency, scales with them); a larger request runs alone */
} fs_server_desc;

typedef struct fs_request {
    fs_scene_desc desc;         /* grid, frame size, storage (device is the server's) */
    fs_params params;
    fs_emitter emitters[FS_MAX_EMITTERS];
    int32_t emitter_count;
    int32_t steps;              /* steps to run, >= 1 */
    int32_t render_every;       /* render a frame every k steps (0: never) */
    int32_t keep_scene;         /* 1: keep the scene after the last step for fs_server_take_scene */
    double arrival_s;           /* arrival on the server clock (seconds since the first tick); < 0: now */
} fs_request;

enum fs_request_state { FS_REQ_PENDING = 0, FS_REQ_WAITING = 1, FS_REQ_RUNNING = 2, FS_REQ_DONE = 3, FS_REQ_FAILED = 4 };

typedef struct fs_request_status {
    int32_t state;              /* enum fs_request_state */
    int32_t steps_done;         /* steps whose completion has been observed */
    int32_t frames;             /* frames rendered */
    int32_t frame_w, frame_h;   /* frame size of the request (fs_server_frame writes frame_w*frame_h pixels) */
    double arrival_s, admitted_s, finished_s;
    double step_ms_p50, step_ms_p99, step_ms_max;
} fs_request_status;

typedef struct fs_server_stats {
    int64_t requests_submitted, requests_done, requests_failed;
    int64_t steps;              /* scene-steps completed */
    int64_t voxel_steps;        /* sum of cells over completed scene-steps */
    int64_t rounds;             /* scheduling rounds that stepped at least one request */
    int64_t frames;
    int32_t max_active_seen;
    int32_t sor_kernel;         /* SOR kernel in use */
    uint64_t peak_resident_bytes;
    double busy_s;              /* first arrival to last completion */
    double step_ms_p50, step_ms_p99, step_ms_mean, step_ms_max;  /* a step's latency: from the request's
                                 * previous step completion (its admission, for the first) to this one */
    double queue_ms_p50, queue_ms_p99;   /* arrival to admission */
} fs_server_stats;

int fs_server_create(const fs_server_desc* desc, fs_server** out);
/* Destroys every request's scene and frees the server. */
int fs_server_destroy(fs_server* server);
int fs_server_submit(fs_server* server, const fs_request* request, uint64_t* id);
/* Every fs_server_* call holds the server's lock while it touches server
 * state: requests may be submitted and status read from other threads while
 * one thread runs fs_server_run or fs_server_tick. fs_server_destroy must not
 * overlap any other call on the server. */
/* One scheduling round; *active receives the number of requests stepped. */
int fs_server_tick(fs_server* server, int32_t* active);
/* Rounds until every submitted request is done (sleeps until the next arrival when idle). */
int fs_server_run(fs_server* server);
int fs_server_status(fs_server* server, uint64_t id, fs_request_status* out);
/* Latest frame of a request: frame_w*frame_h BGRA pixels of the request's
 * descriptor (reported by fs_server_status), so bgra must hold that many;
 * FS_ERR_INVALID_ARGUMENT if none yet. */
int fs_server_frame(fs_server* server, uint64_t id, uint32_t* bgra, int32_t* frame_step);
/* Hands a finished request's scene (kept with keep_scene) to the caller, who then owns it. */
int fs_server_take_scene(fs_server* server, uint64_t id, fs_scene** out);
int fs_server_stats_get(fs_server* server, fs_server_stats* out);
/* Every observed step latency in ms, in completion order; *count receives the total. */
int fs_server_step_latencies(fs_server* server, double* out_ms, int64_t capacity, int64_t* count);
double fs_server_clock(fs_server* server);
const char* fs_server_last_error(const fs_server* server);

#ifdef __cplusplus
}
#endif

#endif /* FIRESIM_H */
