// fs_device.hpp - device-side types and helpers shared by the gfx950 kernels.
//
// Everything here is FP32 arithmetic on CDNA4 (64-lane wavefronts). The math
// follows the reference's device library (src/fire_cuda.cu:219-733); the lattice
// noise uses the PCG integer hash (Jarzynski & Olano, JCGT 2020) in place of the
// reference's sin-hash so that results do not depend on a sinf implementation.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace fs {

enum ScalarId { HEAT = 0, FUEL, OXYGEN, SOOT, CHARF, ASH, PYRO, PROGRESS, TURB, SOPT, NSCALAR };

constexpr int kMaxEmitters = 16;

// Seed of one noise stream, and which lattice hash it uses: the PCG integer
// hash (default) or the reference's sin hash (a visual mode).
struct NoiseKey {
    uint32_t salt;
    int sinHash;
};
constexpr int kMaxEmbers = 192;  // kMaxEmberCount, src/fire_cuda.cu:110
constexpr float kAdvectionScale = 0.62f;  // src/fire_cuda.cu:107

// Per-pass parameter block (SimParams src/fire_cuda.cu:16-68), passed by value.
struct DevParams {
    int nx, ny, nz;
    float invN[3];  // 1/nx, 1/ny, 1/nz: normalized coordinates are index * invN (within one ulp of index / n)
    int lnx, lny, lnz;
    int fw, fh;
    float dt;
    float mouseX, mouseY;
    int leftDown, rightDown, sceneId;
    float windX, windZ, turbulence, detail, smokeGain, intensity;
    int cinematic, raymarchSteps, debugMode;
    float exposure, reflectionGain, smokeDarkness;
    float emitterCenterX, emitterCenterZ, emitterHeightNorm, emitterHeightBandNorm, emitterRadius;
    int burnerCount;
    float burnerX[4], burnerY[4], burnerZ[4];
    float omega;
    int iterations;
    NoiseKey saltNoise2, saltNoise3, saltJitter, saltDither, saltGrain, saltEmber;
    NoiseKey saltCurl[3];  // the three components of the curl-noise potential
    int emberCount;
    int closedTop;        // 1: the top v face is a wall
    int curlOn;           // 1: divergence-free curl noise replaces the fbm noise terms
    float curlNoise;      // its amplitude
    int showGizmos, activeGizmo;
    int dirLight;
    float dirLightDir[3];    // unit direction the light travels
    float dirLightColor[3];
    float camEye[3], camForward[3], camRight[3], camUp[3];
    float tanHalfFov, aspect;
    int emitterCount;
    float emitter[kMaxEmitters][6];  // cx, cy, cz, r (voxels), fuel rate, heat level
    // Device clock [simulation time, render time]. Kernels read the time here
    // rather than from this (by-value) block, so that a captured step graph
    // stays valid while a one-thread kernel advances the clock every step.
    const float* clock;
};

// Element index into a device array. Scene creation rejects grids with more
// than 2^30 values in any field ((nx+1)(ny+1)(nz+1) <= 2^30), so an element's
// byte offset fits 32 bits. ld_at/st_at address an element as uniform base +
// 32-bit byte offset in the global address space: the loads then take the
// offset directly (no 64-bit address math), and pointers read from a batch
// table (which the compiler cannot place) still give global, not flat,
// accesses. Every array they touch is device memory of a scene.
using Idx = uint32_t;
constexpr double kMaxFieldValues = 1073741824.0;  // 2^30

// (Vector types such as float4 cannot bind address-space-qualified
// references; they take the generic path.)
#define FS_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ T ld_at(const T* __restrict__ base, Idx i) {
    if constexpr (std::is_arithmetic_v<T>) {
        const FS_GLOBAL char* b = (const FS_GLOBAL char*)base;
        return *(const FS_GLOBAL T*)(b + i * static_cast<Idx>(sizeof(T)));
    } else {
        return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + i * static_cast<Idx>(sizeof(T)));
    }
}
template <class T>
__device__ __forceinline__ void st_at(T* __restrict__ base, Idx i, T v) {
    if constexpr (std::is_arithmetic_v<T>) {
        FS_GLOBAL char* b = (FS_GLOBAL char*)base;
        *(FS_GLOBAL T*)(b + i * static_cast<Idx>(sizeof(T))) = v;
    } else {
        *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + i * static_cast<Idx>(sizeof(T))) = v;
    }
}

// st_at as a nontemporal store, for a value that only a later launch reads.
template <class T>
__device__ __forceinline__ void st_stream(T* __restrict__ base, Idx i, T v) {
    static_assert(std::is_arithmetic_v<T>, "scalar element types only");
    FS_GLOBAL char* b = (FS_GLOBAL char*)base;
    __builtin_nontemporal_store(v, (FS_GLOBAL T*)(b + i * static_cast<Idx>(sizeof(T))));
}

// Scalar field set. When BF16 is true, heat, fuel and soot are stored as
// bfloat16 (round to nearest even) and widened to FP32 on load.
template <bool BF16>
struct Scalars {
    void* f[NSCALAR];

    static constexpr __host__ __device__ bool narrow(int k) { return BF16 && (k == HEAT || k == FUEL || k == SOOT); }

    template <int K>
    __device__ __forceinline__ float ld(Idx i) const {
        if constexpr (narrow(K)) {
            const uint32_t b = static_cast<uint32_t>(ld_at(static_cast<const uint16_t*>(f[K]), i)) << 16;
            return __uint_as_float(b);
        } else {
            return ld_at(static_cast<const float*>(f[K]), i);
        }
    }
    template <int K>
    __device__ __forceinline__ void st(Idx i, float v) const {
        if constexpr (narrow(K)) {
            uint32_t b = __float_as_uint(v);
            b += 0x7FFFu + ((b >> 16) & 1u);
            st_at(static_cast<uint16_t*>(f[K]), i, static_cast<uint16_t>(b >> 16));
        } else {
            st_at(static_cast<float*>(f[K]), i, v);
        }
    }
};

struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 f3(float x, float y, float z) { return F3{x, y, z}; }
__device__ __forceinline__ F3 add3(F3 a, F3 b) { return f3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ F3 sub3(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ F3 mul3(F3 a, float s) { return f3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ F3 mul3v(F3 a, F3 b) { return f3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

__device__ __forceinline__ float sat(float v) { return fminf(1.0f, fmaxf(0.0f, v)); }
__device__ __forceinline__ float lerpf(float a, float b, float t) { return a + (b - a) * t; }
__device__ __forceinline__ F3 lerp3(F3 a, F3 b, float t) { return add3(a, mul3(sub3(b, a), t)); }
__device__ __forceinline__ float lum3(F3 c) { return c.x * 0.2126f + c.y * 0.7152f + c.z * 0.0722f; }
__device__ __forceinline__ F3 normalize3(F3 v) {
    const float len = sqrtf(fmaxf(0.000001f, dot3(v, v)));
    return f3(v.x / len, v.y / len, v.z / len);
}

// Smoothstep with inverted edges (src/fire_cuda.cu:227-234). The quotient is
// formed with the reciprocal of the edge span: with constant edges (nearly
// every call) that is one multiply by a folded constant instead of a full
// FP32 division, at most one ulp apart.
__device__ __forceinline__ float sstep(float e0, float e1, float x) {
    const bool inv = e1 < e0;
    const float lo = inv ? e1 : e0;
    const float hi = inv ? e0 : e1;
    const float t = sat((x - lo) * (1.0f / fmaxf(0.000001f, hi - lo)));
    const float v = t * t * (3.0f - 2.0f * t);
    return inv ? 1.0f - v : v;
}

// ------------------------------------------------------------- noise ----
__device__ __forceinline__ uint32_t pcg(uint32_t v) {
    const uint32_t state = v * 747796405u + 2891336453u;
    const uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
__device__ __forceinline__ float unit24(uint32_t h) { return static_cast<float>(h >> 8) * (1.0f / 16777216.0f); }

// The reference's lattice hashes hash21/hash31 (src/fire_cuda.cu:381-388):
// fract(sin(.) * 43758.5453) in FP32. Their values depend on the platform's
// sinf, so they serve only the visual noise mode (NoiseKey::sinHash).
__device__ __forceinline__ float fract(float v) { return v - floorf(v); }
__device__ __forceinline__ float sin_hash21(float x, float y) {
    const float a = x * 127.1f + y * 311.7f, b = x * 269.5f + y * 183.3f;
    return fract(sinf(a + b) * 43758.5453123f);
}
__device__ __forceinline__ float sin_hash31(float x, float y, float z) { return fract(sinf(x * 127.1f + y * 311.7f + z * 74.7f) * 43758.5453123f); }

__device__ __forceinline__ float hash2(int x, int y, NoiseKey k) {
    if (k.sinHash) return sin_hash21(static_cast<float>(x), static_cast<float>(y));
    return unit24(pcg(pcg(k.salt + static_cast<uint32_t>(y)) + static_cast<uint32_t>(x)));
}
// A hash of integer inputs (xi, yi) that the reference takes at the float
// inputs (xf, yf) (pixel jitter, dither, grain, spark seeds).
__device__ __forceinline__ float hash2f(float xf, float yf, int xi, int yi, NoiseKey k) {
    return k.sinHash ? sin_hash21(xf, yf) : hash2(xi, yi, k);
}

// Value noise, reference interpolation src/fire_cuda.cu:390-399: the four
// lattice values around (px, py), blended with smoothstep weights.
__device__ __forceinline__ float lattice2(uint32_t rowHash, int x) { return unit24(pcg(rowHash + static_cast<uint32_t>(x))); }
__device__ __forceinline__ float vblend2(float a, float b, float c, float d, float fx, float fy) {
    const float ux = fx * fx * (3.0f - 2.0f * fx);
    const float uy = fy * fy * (3.0f - 2.0f * fy);
    return lerpf(lerpf(a, b, ux), lerpf(c, d, ux), uy);
}
// The reference's sin-hash value noise (:390-399), for the visual mode.
__device__ __forceinline__ float vnoise2_sin(float px, float py) {
    const float fx0 = floorf(px), fy0 = floorf(py);
    return vblend2(sin_hash21(fx0, fy0), sin_hash21(fx0 + 1.0f, fy0), sin_hash21(fx0, fy0 + 1.0f), sin_hash21(fx0 + 1.0f, fy0 + 1.0f), px - fx0,
                   py - fy0);
}
__device__ inline float vnoise2(float px, float py, NoiseKey k) {
    if (k.sinHash) return vnoise2_sin(px, py);
    const float fx0 = floorf(px), fy0 = floorf(py);
    const int ix = static_cast<int>(fx0), iy = static_cast<int>(fy0);
    // one row hash per lattice y, shared by both x corners
    const uint32_t r0 = pcg(k.salt + static_cast<uint32_t>(iy));
    const uint32_t r1 = pcg(k.salt + static_cast<uint32_t>(iy + 1));
    return vblend2(lattice2(r0, ix), lattice2(r0, ix + 1), lattice2(r1, ix), lattice2(r1, ix + 1), px - fx0, py - fy0);
}

// vnoise2 evaluated by a whole wavefront at once, for lanes that are
// neighbouring cells (all 64 lanes active). Each lattice value is then needed
// by several lanes: lane L hashes one lattice point, and every lane takes its
// four corners from the lanes that hashed them (ds_bpermute). The values are
// those of vnoise2, bit for bit; when the lanes do not have the assumed shape
// (all 64 active, a lattice row or column shared by all lanes, the others
// within a span of 64) the per-lane hashes of vnoise2 are used instead.
__device__ __forceinline__ bool full_wave() { return __builtin_amdgcn_read_exec() == ~0ull; }
__device__ __forceinline__ float lane_value(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(lane << 2, __float_as_int(v)));
}
// Lanes on one lattice row (same floor(py)): the two row hashes are
// wave-uniform (scalar unit), lane L hashes columns base + L of both rows.
__device__ __forceinline__ float vnoise2_row_wave(float px, float py, NoiseKey k) {
    if (k.sinHash) return vnoise2_sin(px, py);
    const uint32_t salt = k.salt;
    const float fx0 = floorf(px), fy0 = floorf(py);
    const int ix = static_cast<int>(fx0), iy = static_cast<int>(fy0);
    const int iyw = __builtin_amdgcn_readfirstlane(iy);
    if (!full_wave() || !__all(iy == iyw)) return vnoise2(px, py, k);
    const uint32_t r0 = pcg(salt + static_cast<uint32_t>(iyw));
    const uint32_t r1 = pcg(salt + static_cast<uint32_t>(iyw + 1));
    const int base = __builtin_amdgcn_readfirstlane(ix);
    const int off = ix - base;
    if (__all(off >= 0 && off < 63)) {
        const float h0 = lattice2(r0, base + static_cast<int>(__lane_id())), h1 = lattice2(r1, base + static_cast<int>(__lane_id()));
        return vblend2(lane_value(h0, off), lane_value(h0, off + 1), lane_value(h1, off), lane_value(h1, off + 1), px - fx0, py - fy0);
    }
    return vblend2(lattice2(r0, ix), lattice2(r0, ix + 1), lattice2(r1, ix), lattice2(r1, ix + 1), px - fx0, py - fy0);
}
// Lanes on one lattice column (same floor(px)): lane L hashes row base + L
// and its two column values.
__device__ __forceinline__ float vnoise2_col_wave(float px, float py, NoiseKey k) {
    if (k.sinHash) return vnoise2_sin(px, py);
    const uint32_t salt = k.salt;
    const float fx0 = floorf(px), fy0 = floorf(py);
    const int ix = static_cast<int>(fx0), iy = static_cast<int>(fy0);
    const int ixw = __builtin_amdgcn_readfirstlane(ix);
    const int base = __builtin_amdgcn_readfirstlane(iy);
    const int off = iy - base;
    if (full_wave() && __all(ix == ixw && off >= 0 && off < 63)) {
        const uint32_t r = pcg(salt + static_cast<uint32_t>(base + static_cast<int>(__lane_id())));
        const float g0 = lattice2(r, ixw), g1 = lattice2(r, ixw + 1);
        return vblend2(lane_value(g0, off), lane_value(g1, off), lane_value(g0, off + 1), lane_value(g1, off + 1), px - fx0, py - fy0);
    }
    return vnoise2(px, py, k);
}

// src/fire_cuda.cu:401-421
__device__ inline float vnoise3(float px, float py, float pz, NoiseKey k) {
    const float fx0 = floorf(px), fy0 = floorf(py), fz0 = floorf(pz);
    const int ix = static_cast<int>(fx0), iy = static_cast<int>(fy0), iz = static_cast<int>(fz0);
    const float fx = px - fx0, fy = py - fy0, fz = pz - fz0;
    const float ux = fx * fx * (3.0f - 2.0f * fx);
    const float uy = fy * fy * (3.0f - 2.0f * fy);
    const float uz = fz * fz * (3.0f - 2.0f * fz);
    if (k.sinHash) {
        const float c00 = lerpf(sin_hash31(fx0, fy0, fz0), sin_hash31(fx0 + 1.0f, fy0, fz0), ux);
        const float c10 = lerpf(sin_hash31(fx0, fy0 + 1.0f, fz0), sin_hash31(fx0 + 1.0f, fy0 + 1.0f, fz0), ux);
        const float c01 = lerpf(sin_hash31(fx0, fy0, fz0 + 1.0f), sin_hash31(fx0 + 1.0f, fy0, fz0 + 1.0f), ux);
        const float c11 = lerpf(sin_hash31(fx0, fy0 + 1.0f, fz0 + 1.0f), sin_hash31(fx0 + 1.0f, fy0 + 1.0f, fz0 + 1.0f), ux);
        return lerpf(lerpf(c00, c10, uy), lerpf(c01, c11, uy), uz);
    }
    const uint32_t salt = k.salt;
    const uint32_t z0 = pcg(salt + static_cast<uint32_t>(iz));
    const uint32_t z1 = pcg(salt + static_cast<uint32_t>(iz + 1));
    const uint32_t r00 = pcg(z0 + static_cast<uint32_t>(iy));
    const uint32_t r10 = pcg(z0 + static_cast<uint32_t>(iy + 1));
    const uint32_t r01 = pcg(z1 + static_cast<uint32_t>(iy));
    const uint32_t r11 = pcg(z1 + static_cast<uint32_t>(iy + 1));
    const uint32_t xa = static_cast<uint32_t>(ix), xb = static_cast<uint32_t>(ix + 1);
    const float c000 = unit24(pcg(r00 + xa)), c100 = unit24(pcg(r00 + xb));
    const float c010 = unit24(pcg(r10 + xa)), c110 = unit24(pcg(r10 + xb));
    const float c001 = unit24(pcg(r01 + xa)), c101 = unit24(pcg(r01 + xb));
    const float c011 = unit24(pcg(r11 + xa)), c111 = unit24(pcg(r11 + xb));
    const float c00 = lerpf(c000, c100, ux);
    const float c10 = lerpf(c010, c110, ux);
    const float c01 = lerpf(c001, c101, ux);
    const float c11 = lerpf(c011, c111, ux);
    return lerpf(lerpf(c00, c10, uy), lerpf(c01, c11, uy), uz);
}

// 5-octave fbm, src/fire_cuda.cu:423-432. WAVE selects the octave noise:
// 0 per lane (vnoise2), 1 lanes on one lattice row, 2 lanes on one lattice
// column (the wavefront forms above; same values).
template <int WAVE>
__device__ __forceinline__ float fbm2_t(float px, float py, NoiseKey salt) {
    float sum = 0.0f, amp = 0.5f;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        float n;
        if constexpr (WAVE == 1) n = vnoise2_row_wave(px, py, salt);
        else if constexpr (WAVE == 2) n = vnoise2_col_wave(px, py, salt);
        else n = vnoise2(px, py, salt);
        sum += amp * n;
        px = px * 2.03f + 17.7f;
        py = py * 2.01f - 9.2f;
        amp *= 0.5f;
    }
    return sum;
}
__device__ inline float fbm2(float px, float py, NoiseKey salt) { return fbm2_t<0>(px, py, salt); }

// 3-octave fbm, saturated, src/fire_cuda.cu:434-443
__device__ inline float fbm3(float px, float py, float pz, NoiseKey salt) {
    float sum = 0.0f, amp = 0.5f;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        sum += amp * vnoise3(px, py, pz, salt);
        px = px * 2.03f + 17.7f;
        py = py * 2.01f - 9.2f;
        pz = pz * 2.07f + 5.1f;
        amp *= 0.5f;
    }
    return sat(sum * 1.142857f);
}

// ----------------------------------------------------------- samplers ----
// Clamped trilinear sample of an x-fastest array of size dx*dy*dz with
// coordinate multipliers sx, sy, sz (src/fire_cuda.cu:489-600).
struct Tri {
    Idx i000, i100, i010, i110, i001, i101, i011, i111;
    float tx, ty, tz;
};

// Coordinates may be float or double. The step's backtraces use double: the
// sample position f = u * (n - 1) reaches ~n cells, where one float ulp is up
// to 1.5e-5 of a cell, and that position error times the field gradient would
// dominate the FP32 error of the sampled value. The blend weights are float.
__device__ __forceinline__ float sat_c(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }
__device__ __forceinline__ double sat_c(double x) { return fmin(fmax(x, 0.0), 1.0); }

// One axis of a sample: the clamped coordinate u in [0, 1] times the
// multiplier s, its floor cell i0 (and i1 = i0 + 1, clamped to d - 1) and the
// fraction. A trilinear setup is three of these, so a caller sampling several
// arrays at one position (or one array at positions sharing a coordinate) can
// form each axis once.
struct Axis {
    int i0, i1;
    float t;
};
template <class T>
__device__ __forceinline__ Axis axis_setup(int d, T s, T u) {
    u = sat_c(u);
    const T f = u * s;
    const T fl = floor(f);
    const int i0 = static_cast<int>(fl);
    return Axis{i0, min(i0 + 1, d - 1), static_cast<float>(f - fl)};
}
// Corner offsets of an x-fastest array with rows of dx and planes of dx*dy.
__device__ __forceinline__ Tri tri_of(const Axis& ax, const Axis& ay, const Axis& az, int dx, int dy) {
    Tri t;
    t.tx = ax.t;
    t.ty = ay.t;
    t.tz = az.t;
    // corner (x0, y0, z0) and the steps to the +1 corners (0 where clamped)
    const Idx row = static_cast<Idx>(dx), plane = row * static_cast<Idx>(dy);
    const Idx ox = static_cast<Idx>(ax.i1 - ax.i0);
    const Idx oy = ay.i1 != ay.i0 ? row : 0u;
    const Idx oz = az.i1 != az.i0 ? plane : 0u;
    t.i000 = static_cast<Idx>(az.i0) * plane + static_cast<Idx>(ay.i0) * row + static_cast<Idx>(ax.i0);
    t.i100 = t.i000 + ox;
    t.i010 = t.i000 + oy;
    t.i110 = t.i010 + ox;
    t.i001 = t.i000 + oz;
    t.i101 = t.i001 + ox;
    t.i011 = t.i001 + oy;
    t.i111 = t.i011 + ox;
    return t;
}
template <class T>
__device__ __forceinline__ Tri tri_setup(int dx, int dy, int dz, T sx, T sy, T sz, T u, T v, T w) {
    return tri_of(axis_setup<T>(dx, sx, u), axis_setup<T>(dy, sy, v), axis_setup<T>(dz, sz, w), dx, dy);
}

__device__ __forceinline__ float tri_blend(const Tri& t, float c000, float c100, float c010, float c110, float c001, float c101,
                                           float c011, float c111) {
    const float c00 = lerpf(c000, c100, t.tx);
    const float c10 = lerpf(c010, c110, t.tx);
    const float c01 = lerpf(c001, c101, t.tx);
    const float c11 = lerpf(c011, c111, t.tx);
    return lerpf(lerpf(c00, c10, t.ty), lerpf(c01, c11, t.ty), t.tz);
}

__device__ __forceinline__ float tri_plain(const float* __restrict__ f, const Tri& t) {
    return tri_blend(t, ld_at(f, t.i000), ld_at(f, t.i100), ld_at(f, t.i010), ld_at(f, t.i110), ld_at(f, t.i001), ld_at(f, t.i101),
                     ld_at(f, t.i011), ld_at(f, t.i111));
}

template <int K, bool B>
__device__ __forceinline__ float tri_scalar(const Scalars<B>& S, const Tri& t) {
    return tri_blend(t, S.template ld<K>(t.i000), S.template ld<K>(t.i100), S.template ld<K>(t.i010), S.template ld<K>(t.i110),
                     S.template ld<K>(t.i001), S.template ld<K>(t.i101), S.template ld<K>(t.i011), S.template ld<K>(t.i111));
}

template <class T>
__device__ __forceinline__ Tri tri_scalar_setup(const DevParams& p, T u, T v, T w) {
    return tri_setup<T>(p.nx, p.ny, p.nz, static_cast<T>(p.nx - 1), static_cast<T>(p.ny - 1), static_cast<T>(p.nz - 1), u, v, w);
}

template <int K, bool B>
__device__ __forceinline__ float sample_scalar(const Scalars<B>& S, const DevParams& p, float u, float v, float w) {
    return tri_scalar<K, B>(S, tri_scalar_setup<float>(p, u, v, w));
}

template <class T>
__device__ __forceinline__ float sample_u(const float* __restrict__ f, const DevParams& p, T u, T v, T w) {
    return tri_plain(f, tri_setup<T>(p.nx + 1, p.ny, p.nz, static_cast<T>(p.nx), static_cast<T>(p.ny - 1), static_cast<T>(p.nz - 1), u, v, w));
}
template <class T>
__device__ __forceinline__ float sample_v(const float* __restrict__ f, const DevParams& p, T u, T v, T w) {
    return tri_plain(f, tri_setup<T>(p.nx, p.ny + 1, p.nz, static_cast<T>(p.nx - 1), static_cast<T>(p.ny), static_cast<T>(p.nz - 1), u, v, w));
}
template <class T>
__device__ __forceinline__ float sample_w(const float* __restrict__ f, const DevParams& p, T u, T v, T w) {
    return tri_plain(f, tri_setup<T>(p.nx, p.ny, p.nz + 1, static_cast<T>(p.nx - 1), static_cast<T>(p.ny - 1), static_cast<T>(p.nz), u, v, w));
}
template <class T>
__device__ __forceinline__ F3 sample_vel(const float* __restrict__ U, const float* __restrict__ V, const float* __restrict__ W,
                                         const DevParams& p, T u, T v, T w) {
    return f3(sample_u<T>(U, p, u, v, w), sample_v<T>(V, p, u, v, w), sample_w<T>(W, p, u, v, w));
}

// Light grid: float4 trilinear and nearest shadow (src/fire_cuda.cu:618-664).
__device__ __forceinline__ float4 sample_light(const float4* __restrict__ L, const DevParams& p, float u, float v, float w) {
    const Tri t = tri_setup(p.lnx, p.lny, p.lnz, static_cast<float>(p.lnx - 1), static_cast<float>(p.lny - 1),
                            static_cast<float>(p.lnz - 1), u, v, w);
    const float4 a = ld_at(L, t.i000), b = ld_at(L, t.i100), c = ld_at(L, t.i010), d = ld_at(L, t.i110);
    const float4 e = ld_at(L, t.i001), f = ld_at(L, t.i101), g = ld_at(L, t.i011), h = ld_at(L, t.i111);
    float4 o;
    o.x = tri_blend(t, a.x, b.x, c.x, d.x, e.x, f.x, g.x, h.x);
    o.y = tri_blend(t, a.y, b.y, c.y, d.y, e.y, f.y, g.y, h.y);
    o.z = tri_blend(t, a.z, b.z, c.z, d.z, e.z, f.z, g.z, h.z);
    o.w = tri_blend(t, a.w, b.w, c.w, d.w, e.w, f.w, g.w, h.w);
    return o;
}
__device__ __forceinline__ float sample_shadow(const float* __restrict__ S, const DevParams& p, float u, float v, float w) {
    u = sat(u);
    v = sat(v);
    w = sat(w);
    const int x = min(max(static_cast<int>(u * static_cast<float>(p.lnx - 1) + 0.5f), 0), p.lnx - 1);
    const int y = min(max(static_cast<int>(v * static_cast<float>(p.lny - 1) + 0.5f), 0), p.lny - 1);
    const int z = min(max(static_cast<int>(w * static_cast<float>(p.lnz - 1) + 0.5f), 0), p.lnz - 1);
    return ld_at(S, (static_cast<Idx>(z) * static_cast<Idx>(p.lny) + static_cast<Idx>(y)) * static_cast<Idx>(p.lnx) + static_cast<Idx>(x));
}

// --------------------------------------------------------- fuel beds ----
// src/fire_cuda.cu:734-744
__device__ inline float capsule_mask(float x, float z, float cx, float cz, float angle, float halfLength, float radius) {
    float sa, ca;
    sincosf(angle, &sa, &ca);
    const float px = x - cx, pz = z - cz;
    const float qx = px * ca + pz * sa;
    const float qz = -px * sa + pz * ca;
    const float over = fmaxf(fabsf(qx) - halfLength, 0.0f);
    return sstep(radius, radius * 0.22f, sqrtf(over * over + qz * qz));
}

// Procedural bed material (src/fire_cuda.cu:746-793); scene 3 has none.
__device__ inline float bed_material(float x, float z, const DevParams& p) {
    if (p.sceneId == 3) return 0.0f;
    const NoiseKey s2 = p.saltNoise2;
    if (p.sceneId == 1) {
        const float logA = capsule_mask(x, z, -0.06f, 0.01f, 0.38f, 0.34f, 0.060f);
        const float logB = capsule_mask(x, z, 0.07f, -0.02f, -0.58f, 0.33f, 0.058f);
        const float logC = capsule_mask(x, z, 0.01f, 0.08f, 1.34f, 0.24f, 0.050f);
        const float cAB = fminf(logA, logB) * 1.25f;
        const float cBC = fminf(logB, logC) * 1.10f;
        const float cCA = fminf(logC, logA) * 0.92f;
        const float coalNoise = fbm2(x * 9.4f + z * 1.7f, z * 11.2f - x * 3.1f, s2);
        const float coalBed = sstep(0.38f, 0.08f, sqrtf(x * x * 1.35f + z * z * 2.40f)) * sstep(0.34f, 0.88f, coalNoise);
        const float cracks = sstep(0.65f, 0.93f, fbm2(x * 21.0f - z * 6.0f, z * 24.0f + x * 4.0f, s2));
        const float contact = sat(cAB + cBC + cCA);
        return sat((logA * 0.72f + logB * 0.70f + logC * 0.54f + coalBed * 0.84f + contact * 1.18f) * (1.0f - cracks * 0.24f));
    }
    if (p.sceneId == 2) {
        const float R = fmaxf(0.04f, p.emitterRadius);
        const float cx = p.burnerCount > 0 ? p.burnerX[0] : p.emitterCenterX;
        const float cz = p.burnerCount > 0 ? p.burnerZ[0] : p.emitterCenterZ;
        const float px = x - cx, pz = z - cz;
        const float radius = sqrtf(px * px + pz * pz);
        const float angle = atan2f(pz, px);
        const float ring = sstep(R * 0.072f, R * 0.014f, fabsf(radius - R));
        const float inner = sstep(R * 0.052f, R * 0.012f, fabsf(radius - R * 0.52f));
        const float ports = 0.5f + 0.5f * cosf(angle * 32.0f);
        const float portMask = sstep(0.80f, 0.992f, ports);
        const float pilot = sstep(R * 0.22f, 0.000f, radius) * 0.045f;
        const float source = fmaxf(0.0f, (ring * (0.035f + portMask * 1.72f) + inner * 0.055f + pilot) * sstep(R * 1.28f, R * 1.05f, radius));
        return sat(source);
    }
    const float tray = sstep(1.08f, 0.030f, fabsf(x)) * sstep(0.56f, 0.026f, fabsf(z));
    const float emberMat = sstep(1.00f, 0.018f, fabsf(x)) * sstep(0.50f, 0.020f, fabsf(z));
    const float mound = expf(-(x * x * 0.70f + z * z * 2.15f));
    const float logMass = fbm2(x * 3.7f + z * 0.9f, z * 5.4f - x * 0.6f, s2);
    const float chips = fbm2(x * 15.0f + z * 4.0f, z * 18.0f - x * 5.0f, s2);
    const float strand = 0.5f + 0.5f * sinf((x * 19.0f + z * 7.0f) + fbm2(z * 4.0f, x * 3.0f, s2) * 4.2f);
    const float cracks = sstep(0.44f, 0.74f, chips) * sstep(0.20f, 0.84f, logMass);
    const float voids = sstep(0.70f, 0.93f, chips * 0.72f + strand * 0.34f);
    const float clumps = sstep(0.22f, 0.86f, logMass * 0.82f + chips * 0.36f + strand * 0.18f);
    const float porous = (0.24f + clumps * 0.94f + emberMat * 0.20f) * (0.64f + mound * 0.46f);
    return sat(tray * porous * (1.0f - cracks * 0.46f) * (1.0f - voids * 0.34f));
}

// src/fire_cuda.cu:795-803, given the material already evaluated.
__device__ inline float bed_source(float material, float x, float z, float h, const DevParams& p, float time) {
    if (material == 0.0f) return 0.0f;
    const float height = p.sceneId == 2
        ? sstep(fmaxf(0.005f, p.emitterHeightBandNorm * 0.34f), fmaxf(0.0015f, p.emitterHeightBandNorm * 0.055f), fabsf(h - p.emitterHeightNorm))
        : sstep(p.sceneId == 1 ? 0.245f : 0.145f, 0.00f, h);
    const float breathing = 0.84f + 0.16f * fbm2(x * 9.0f + time * 0.035f, z * 11.0f - time * 0.025f, p.saltNoise2);
    return sat(height * sat(material * breathing));
}

// src/fire_cuda.cu:1420-1443
__device__ inline F3 blackbody(float tempK) {
    const float t = fmaxf(1000.0f, fminf(12000.0f, tempK)) / 100.0f;
    float r, g, b;
    if (t <= 66.0f) {
        r = 1.0f;
        g = sat((99.4708025861f * logf(t) - 161.1195681661f) / 255.0f);
    } else {
        r = sat(329.698727446f * powf(t - 60.0f, -0.1332047592f) / 255.0f);
        g = sat(288.1221695283f * powf(t - 60.0f, -0.0755148492f) / 255.0f);
    }
    if (t >= 66.0f) {
        b = 1.0f;
    } else if (t <= 19.0f) {
        b = 0.0f;
    } else {
        b = sat((138.5177312231f * logf(t - 10.0f) - 305.0447927307f) / 255.0f);
    }
    return f3(r, g, b);
}

// src/fire_cuda.cu:359-368
__device__ __forceinline__ float domain_fade(float u, float v, float w) {
    const float side = sstep(0.000f, 0.095f, u) * sstep(1.000f, 0.905f, u) * sstep(0.000f, 0.100f, w) * sstep(1.000f, 0.900f, w);
    return sat(side * sstep(0.000f, 0.018f, v) * sstep(1.000f, 0.520f, v));
}

// ------------------------------------------------------ XCD mapping ----
// MI355X dispatches the workgroups of a launch to its 8 XCDs round-robin by
// linear workgroup id, and each XCD has its own L2. xcd_block() renumbers the
// workgroups so that every XCD receives one contiguous range of logical blocks
// (z-major, i.e. a slab of the domain): neighbouring tiles then share an L2 and
// each XCD's working set is 1/8 of the field instead of all of it.
constexpr unsigned kXcdCount = 8;
__device__ __forceinline__ uint3 xcd_block() {
    const unsigned gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
    const unsigned n = gx * gy * gz;
    const unsigned hw = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const unsigned xcd = hw % kXcdCount, slot = hw / kXcdCount;
    const unsigned logical = xcd * (n / kXcdCount) + min(xcd, n % kXcdCount) + slot;
    return make_uint3(logical % gx, (logical / gx) % gy, logical / (gx * gy));
}

// ------------------------------------------------------------ reduce ----
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, 64));
    return v;
}

}  // namespace fs
