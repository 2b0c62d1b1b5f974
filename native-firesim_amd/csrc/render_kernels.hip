// render_kernels.hip - light-transmittance volume, HDR raymarch and tonemap.
//
//   k_light_build      K17 buildSceneLightKernel      src/fire_cuda.cu:1445-1525
//   k_light_propagate  K18 propagateSceneLightKernel  src/fire_cuda.cu:1527-1565
//   k_render           K19 renderKernel               src/fire_cuda.cu:1953-2244
//   k_tonemap          K21 tonemapKernel              src/fire_cuda.cu:2246-2297
#include "fs_launch.hpp"

namespace fs {

// ---------------------------------------------------------------- K17 ----
template <bool B>
__device__ __forceinline__ void light_build_at(float4* __restrict__ L, float* __restrict__ SH, const Scalars<B>& S, const DevParams& p, int lx, int ly,
                                               int lz) {
    if (lx >= p.lnx || ly >= p.lny || lz >= p.lnz) return;
    const float u = (static_cast<float>(lx) + 0.5f) / static_cast<float>(p.lnx);
    const float v = (static_cast<float>(ly) + 0.5f) / static_cast<float>(p.lny);
    const float w = (static_cast<float>(lz) + 0.5f) / static_cast<float>(p.lnz);
    const float fade = domain_fade(u, v, w);
    const Tri tc = tri_scalar_setup(p, u, v, w);
    const float heat = tri_scalar<HEAT, B>(S, tc);
    const float fuel = tri_scalar<FUEL, B>(S, tc);
    const float oxygen = tri_scalar<OXYGEN, B>(S, tc);
    const float soot = tri_scalar<SOOT, B>(S, tc);
    const float charMass = tri_scalar<CHARF, B>(S, tc);
    const float ash = tri_scalar<ASH, B>(S, tc);
    const float pyro = tri_scalar<PYRO, B>(S, tc);
    const float progress = tri_scalar<PROGRESS, B>(S, tc);
    const float sootOptics = tri_scalar<SOPT, B>(S, tc);
    const float cellX = 1.0f / static_cast<float>(p.nx);
    const float cellY = 1.0f / static_cast<float>(p.ny);
    const float cellZ = 1.0f / static_cast<float>(p.nz);
    const float gx = sample_scalar<HEAT, B>(S, p, u + cellX, v, w) - sample_scalar<HEAT, B>(S, p, u - cellX, v, w);
    const float gy = sample_scalar<HEAT, B>(S, p, u, v + cellY, w) - sample_scalar<HEAT, B>(S, p, u, v - cellY, w);
    const float gz = sample_scalar<HEAT, B>(S, p, u, v, w + cellZ) - sample_scalar<HEAT, B>(S, p, u, v, w - cellZ);
    const float front = sqrtf(gx * gx + gy * gy + gz * gz) + fuel * oxygen * 0.035f + progress * 0.16f;
    const float t = ld_at(p.clock, 0u);
    const float fine = fbm3(u * 34.0f + t * 0.45f, v * 42.0f - t * 0.82f, w * 34.0f + t * 0.22f, p.saltNoise3);
    const float shear = fbm3(u * 12.0f - t * 0.38f, v * 11.0f - t * 0.62f, w * 13.0f + t * 0.20f, p.saltNoise3);
    const float reactionFront = sstep(0.035f, 0.54f, front + pyro * 0.10f);
    const float heightFade = sstep(0.70f, 0.070f, v);
    const float lowerShadow = sstep(0.62f, 0.06f, v);
    const float plumeShadow = sstep(0.12f, 0.88f, v);
    const float sheetGate = sstep(0.56f, 0.94f, fine * 0.48f + shear * 0.42f + reactionFront * 0.26f + heat * 0.018f) *
                            sstep(0.04f, 0.86f, reactionFront) * heightFade;
    const float sheetCore = sat(sheetGate * (0.42f + progress * 0.22f + pyro * 0.34f));
    const float tempK = 293.0f + heat * 360.0f + fuel * 44.0f + pyro * 120.0f + progress * 70.0f + sheetCore * 1020.0f;
    const float flameEmitter = fade * sstep(820.0f, 1880.0f, tempK) * sstep(0.010f, 0.68f, fuel + pyro * 0.30f + heat * 0.040f) *
                               sstep(0.030f, 0.82f, oxygen) * reactionFront * sheetGate * (0.24f + sheetCore * 1.42f);
    const float emberEmitter = sstep(0.002f, 0.22f, charMass + ash * 0.18f) * sstep(0.09f, 0.00f, v) * (0.10f + heat * 0.012f + pyro * 0.050f);
    const float sootAtt = expf(-(sootOptics * 0.25f + soot * 0.090f + ash * 0.035f));
    const F3 flame = mul3(blackbody(tempK), flameEmitter * sootAtt * 3.40f);
    const F3 coal = mul3(f3(1.0f, 0.22f, 0.040f), emberEmitter * p.reflectionGain * 0.20f);
    const float extinction = sat(sootOptics * (0.26f + plumeShadow * 0.42f) + soot * (0.065f + plumeShadow * 0.20f) + ash * 0.050f +
                                 progress * lowerShadow * 0.020f);
    const size_t idx = (static_cast<size_t>(lz) * p.lny + ly) * p.lnx + lx;
    L[idx] = make_float4(flame.x + coal.x, flame.y + coal.y, flame.z + coal.z, extinction);
    const float vertDepth = v * (1.0f + plumeShadow * 0.70f);
    const float radial = expf(-((u - 0.50f) * (u - 0.50f) * 5.5f + (w - 0.50f) * (w - 0.50f) * 7.0f));
    const float optical = sootOptics * (0.62f + plumeShadow * 1.15f) + soot * (0.22f + plumeShadow * 0.46f) + ash * 0.072f;
    SH[idx] = expf(-(optical * (0.72f + vertDepth * 0.42f) + radial * optical * 0.34f + progress * 0.035f));
}

template <bool B>
__global__ __launch_bounds__(256) void k_light_build(float4* __restrict__ L, float* __restrict__ SH, Scalars<B> S, DevParams p) {
    light_build_at<B>(L, SH, S, p, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y * blockDim.y + threadIdx.y, blockIdx.z * blockDim.z + threadIdx.z);
}

// ---------------------------------------------------------------- K18 ----
__device__ __forceinline__ void light_propagate_at(float4* __restrict__ out, const float4* __restrict__ in, const float* __restrict__ SH,
                                                   const DevParams& p, float diffusion, float decay, int lx, int ly, int lz) {
    if (lx >= p.lnx || ly >= p.lny || lz >= p.lnz) return;
    auto li = [&](int x, int y, int z) { return (static_cast<size_t>(z) * p.lny + y) * p.lnx + x; };
    const size_t idx = li(lx, ly, lz);
    const float4 center = in[idx];
    const int nbx[6] = {max(lx - 1, 0), min(lx + 1, p.lnx - 1), lx, lx, lx, lx};
    const int nby[6] = {ly, ly, max(ly - 1, 0), min(ly + 1, p.lny - 1), ly, ly};
    const int nbz[6] = {lz, lz, lz, lz, max(lz - 1, 0), min(lz + 1, p.lnz - 1)};
    F3 nb = f3(0.0f, 0.0f, 0.0f);
    float weight = 0.0f;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const float4 s = in[li(nbx[i], nby[i], nbz[i])];
        const float occ = expf(-s.w * 0.42f);
        nb = add3(nb, mul3(f3(s.x, s.y, s.z), occ));
        weight += occ;
    }
    nb = mul3(nb, 1.0f / fmaxf(0.0001f, weight));
    const float shadow = SH[idx];
    const F3 direct = f3(center.x, center.y, center.z);
    const F3 bounced = lerp3(direct, nb, diffusion);
    const float floorLift = sstep(0.10f, 0.00f, (static_cast<float>(ly) + 0.5f) / static_cast<float>(p.lny));
    const F3 contact = mul3(f3(1.0f, 0.40f, 0.12f), floorLift * lum3(direct) * 0.18f);
    const F3 next = add3(mul3(bounced, decay * (0.66f + shadow * 0.34f)), contact);
    out[idx] = make_float4(next.x, next.y, next.z, center.w);
}

__global__ __launch_bounds__(256) void k_light_propagate(float4* __restrict__ out, const float4* __restrict__ in, const float* __restrict__ SH,
                                                         DevParams p, float diffusion, float decay) {
    light_propagate_at(out, in, SH, p, diffusion, decay, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y * blockDim.y + threadIdx.y,
                       blockIdx.z * blockDim.z + threadIdx.z);
}

// ----------------------------------------------------- directional light ----
// An extension (no reference counterpart): a key light travelling along the
// unit direction dirLightDir. Its transmittance to each light cell is
// Beer-Lambert along the ray toward the light, exp(-kDirKappa * sum_k ds * w_k
// * ext(s_k)), with samples s_k = (k + 1/2) ds (ds is half the smallest light
// cell) and weights w_k = sat(1/2 + (len - s_k) / ds), len being the ray's
// length inside the volume box. sum_k ds * w_k is exactly len (the weights are
// the box tiling of [0, len)), so a uniform extinction e gives exp(-kDirKappa e len),
// and the weights make the sum continuous in len. The smoke scatters the light
// into the camera and the room surfaces receive it diffusely.
constexpr float kDirKappa = 3.0f;    // extinction per world unit at light-volume extinction 1
constexpr float kDirScatter = 2.0f;  // in-scatter gain of the smoke's single-scattering albedo per step
constexpr float kDirDiffuse = 0.10f; // diffuse gain on the room surfaces
constexpr int kDirMaxSamples = 208;  // box diagonal 3.35 / smallest ds 0.0171 = 196.1, + 1
__device__ __forceinline__ float dir_step(const DevParams& p) {
    return 0.5f * fminf(fminf(2.10f / static_cast<float>(p.lnx), 2.03f / static_cast<float>(p.lny)), 1.64f / static_cast<float>(p.lnz));
}
// Distance from x0 along d to the slab [lo, hi] (x0 inside); huge when d ~ 0.
__device__ __forceinline__ float slab_exit(float x0, float d, float lo, float hi) {
    if (d > 1.0e-6f) return (hi - x0) / d;
    if (d < -1.0e-6f) return (lo - x0) / d;
    return 1.0e30f;
}
__device__ __forceinline__ void dir_trans_at(float* __restrict__ DT, const float4* __restrict__ L, const DevParams& p, int lx, int ly, int lz) {
    if (lx >= p.lnx || ly >= p.lny || lz >= p.lnz) return;
    const float u = (static_cast<float>(lx) + 0.5f) / static_cast<float>(p.lnx);
    const float v = (static_cast<float>(ly) + 0.5f) / static_cast<float>(p.lny);
    const float w = (static_cast<float>(lz) + 0.5f) / static_cast<float>(p.lnz);
    const float x0 = u * 2.10f - 1.05f, y0 = v * 2.03f + 0.02f, z0 = w * 1.64f - 0.82f;
    const float dx = -p.dirLightDir[0], dy = -p.dirLightDir[1], dz = -p.dirLightDir[2];
    const float len = fminf(fminf(slab_exit(x0, dx, -1.05f, 1.05f), slab_exit(y0, dy, 0.02f, 2.05f)), slab_exit(z0, dz, -0.82f, 0.82f));
    const float ds = dir_step(p);
    float tau = 0.0f;
    for (int k = 0; k < kDirMaxSamples; ++k) {
        const float s = ds * (static_cast<float>(k) + 0.5f);
        const float wk = fminf(1.0f, 0.5f + (len - s) / ds);
        if (wk <= 0.0f) break;
        const float qx = x0 + dx * s, qy = y0 + dy * s, qz = z0 + dz * s;
        tau += wk * sample_light(L, p, (qx + 1.05f) / 2.10f, (qy - 0.02f) / 2.03f, (qz + 0.82f) / 1.64f).w;
    }
    DT[(static_cast<size_t>(lz) * p.lny + ly) * p.lnx + lx] = expf(-kDirKappa * ds * tau);
}
__global__ __launch_bounds__(256) void k_dir_trans(float* __restrict__ DT, const float4* __restrict__ L, DevParams p) {
    dir_trans_at(DT, L, p, blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y * blockDim.y + threadIdx.y, blockIdx.z * blockDim.z + threadIdx.z);
}
// Trilinear sample of the transmittance volume (the light grid's mapping).
__device__ __forceinline__ float sample_dir_trans(const float* __restrict__ DT, const DevParams& p, float u, float v, float w) {
    return tri_plain(DT, tri_setup(p.lnx, p.lny, p.lnz, static_cast<float>(p.lnx - 1), static_cast<float>(p.lny - 1), static_cast<float>(p.lnz - 1), u,
                                   v, w));
}

// ------------------------------------------------------------- render ----
struct Cam {
    F3 eye, fwd, right, up;
    float tanHalfFov, aspect;
};
__device__ __forceinline__ Cam cam_of(const DevParams& p) {
    Cam c;
    c.eye = f3(p.camEye[0], p.camEye[1], p.camEye[2]);
    c.fwd = f3(p.camForward[0], p.camForward[1], p.camForward[2]);
    c.right = f3(p.camRight[0], p.camRight[1], p.camRight[2]);
    c.up = f3(p.camUp[0], p.camUp[1], p.camUp[2]);
    c.tanHalfFov = p.tanHalfFov;
    c.aspect = p.aspect;
    return c;
}

// projectPoint, src/fire_cuda.cu:348-357: frame uv and view depth (z < 0 behind the camera)
__device__ __forceinline__ F3 project_point(const Cam& cam, F3 point) {
    const F3 view = sub3(point, cam.eye);
    const float depth = dot3(view, cam.fwd);
    if (depth <= 0.02f) return f3(-10.0f, -10.0f, -1.0f);
    const float px = dot3(view, cam.right) / (depth * cam.tanHalfFov * cam.aspect);
    const float py = dot3(view, cam.up) / (depth * cam.tanHalfFov);
    return f3(0.5f + px * 0.5f, 0.5f - py * 0.5f, depth);
}

__device__ __forceinline__ F3 uv_from_world(F3 q) {
    return f3(sat((q.x + 1.05f) / 2.10f), sat((q.y - 0.02f) / 2.03f), sat((q.z + 0.82f) / 1.64f));
}
__device__ __forceinline__ F3 world_from_uv(F3 uv) { return f3(uv.x * 2.10f - 1.05f, uv.y * 2.03f + 0.02f, uv.z * 1.64f - 0.82f); }

// src/fire_cuda.cu:1589-1626
__device__ F3 gather_irradiance(F3 hit, F3 normal, const float4* L, const float* SH, const DevParams& p) {
    F3 irr = f3(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const F3 suv = i == 0 ? f3(0.50f, 0.16f, 0.50f) : (i == 1 ? f3(0.34f, 0.26f, 0.43f) : f3(0.66f, 0.26f, 0.57f));
        const float4 light = sample_light(L, p, suv.x, suv.y, suv.z);
        const F3 radiance = f3(light.x, light.y, light.z);
        if (lum3(radiance) <= 0.00001f) continue;
        const F3 toLight = sub3(world_from_uv(suv), hit);
        const float dist2 = fmaxf(0.045f, dot3(toLight, toLight));
        const F3 wi = mul3(toLight, rsqrtf(dist2));
        const float ndotl = fmaxf(0.0f, dot3(normal, wi));
        const float heightOcc = expf(-fmaxf(0.0f, hit.y - world_from_uv(suv).y) * light.w * 1.15f);
        const float vis = sample_shadow(SH, p, suv.x, suv.y, suv.z) * expf(-light.w * 0.92f) * heightOcc;
        const float geometric = (0.12f + ndotl * 0.88f) * expf(-dist2 * 0.22f) / (0.12f + dist2);
        irr = add3(irr, mul3(radiance, vis * geometric));
    }
    const float grazing = 1.0f + sstep(0.0f, 0.45f, 1.0f - fmaxf(0.0f, normal.y)) * 0.36f;
    return mul3(irr, 4.20f * grazing);
}

// src/fire_cuda.cu:1628-1645
__device__ float volume_shadow_ray(F3 hit, const float4* L, const float* SH, const DevParams& p) {
    const F3 target = uv_from_world(hit);
    const F3 source = f3(0.50f, 0.12f, 0.50f);
    float od = 0.0f;
#pragma unroll
    for (int i = 1; i <= 5; ++i) {
        const float t = static_cast<float>(i) / 6.0f;
        const F3 uv = lerp3(source, target, t);
        const float4 light = sample_light(L, p, uv.x, uv.y, uv.z);
        const float shadow = sample_shadow(SH, p, uv.x, uv.y, uv.z);
        od += light.w * (0.18f + t * 0.22f) + (1.0f - shadow) * (0.08f + t * 0.16f);
    }
    return expf(-od);
}

__device__ __forceinline__ float rect_mask(float px, float py, float cx, float cy, float hx, float hy, float feather) {
    return sstep(feather, -feather, fmaxf(fabsf(px - cx) - hx, fabsf(py - cy) - hy));
}
__device__ __forceinline__ float line_mask(float px, float py, float ax, float ay, float bx, float by, float thickness) {
    const float pax = px - ax, pay = py - ay, bax = bx - ax, bay = by - ay;
    const float h = sat((pax * bax + pay * bay) / fmaxf(0.000001f, bax * bax + bay * bay));
    const float dx = pax - bax * h, dy = pay - bay * h;
    return sstep(thickness, thickness * 0.25f, sqrtf(dx * dx + dy * dy));
}

// src/fire_cuda.cu:348-357 and :1661-1681
__device__ void source_model_overlay(float uvx, float uvy, const Cam& cam, const DevParams& p, F3& color) {
    if (uvx < 0.24f || uvx > 0.76f || uvy < 0.42f || uvy > 0.88f) return;
    if (p.sceneId == 1) return;
    const bool burner = p.sceneId == 2 && p.burnerCount > 0;
    const F3 src = f3(burner ? p.burnerX[0] : p.emitterCenterX, burner ? p.burnerY[0] : p.emitterHeightNorm * 2.03f + 0.02f,
                      burner ? p.burnerZ[0] : p.emitterCenterZ);
    const F3 view = sub3(src, cam.eye);
    const float depth = dot3(view, cam.fwd);
    F3 c = f3(-10.0f, -10.0f, -1.0f);
    if (depth > 0.02f) {
        const float px = dot3(view, cam.right) / (depth * cam.tanHalfFov * cam.aspect);
        const float py = dot3(view, cam.up) / (depth * cam.tanHalfFov);
        c = f3(0.5f + px * 0.5f, 0.5f - py * 0.5f, depth);
    }
    const float tray = sstep(0.135f, 0.105f, fabsf(uvx - c.x)) * sstep(0.050f, 0.036f, fabsf(uvy - c.y));
    const float inner = sstep(0.112f, 0.086f, fabsf(uvx - c.x)) * sstep(0.039f, 0.027f, fabsf(uvy - c.y));
    const float alpha = sat((tray - inner * 0.55f) * 0.65f);
    const F3 model = lerp3(f3(0.015f, 0.013f, 0.011f), f3(0.13f, 0.12f, 0.10f), alpha);
    if (alpha > 0.001f) color = lerp3(color, model, sat(alpha));
}

// roomBackgroundRay src/fire_cuda.cu:1683-1865 (floorGrid :1567-1573 is 0,
// so the terms it scales are omitted).
__device__ F3 room_background(const Cam& cam, float uvx, float uvy, F3 rd, float glow, const float4* L, const float* SH, const float* DT,
                              const DevParams& p, float time) {
    const float roomHalf = 2.55f, ceilingY = 2.18f;
    const bool cinematic = p.cinematic != 0;
    float bestT = 1.0e6f;
    F3 hit = f3(0.0f, 0.0f, 0.0f);
    F3 normal = f3(0.0f, 1.0f, 0.0f);
    int surface = 0;
    if (rd.y < -0.0001f) {
        const float t = -cam.eye.y / rd.y;
        const F3 q = add3(cam.eye, mul3(rd, t));
        if (t > 0.0f && fabsf(q.x) <= roomHalf && fabsf(q.z) <= roomHalf && t < bestT) { bestT = t; hit = q; normal = f3(0, 1, 0); surface = 1; }
    }
    if (rd.y > 0.0001f) {
        const float t = (ceilingY - cam.eye.y) / rd.y;
        const F3 q = add3(cam.eye, mul3(rd, t));
        if (t > 0.0f && fabsf(q.x) <= roomHalf && fabsf(q.z) <= roomHalf && t < bestT) { bestT = t; hit = q; normal = f3(0, -1, 0); surface = 2; }
    }
    if (fabsf(rd.x) > 0.0001f) {
        const float planeX = rd.x > 0.0f ? roomHalf : -roomHalf;
        const float t = (planeX - cam.eye.x) / rd.x;
        const F3 q = add3(cam.eye, mul3(rd, t));
        if (t > 0.0f && q.y >= 0.0f && q.y <= ceilingY && fabsf(q.z) <= roomHalf && t < bestT) {
            bestT = t; hit = q; normal = f3(planeX > 0.0f ? -1.0f : 1.0f, 0, 0); surface = 3;
        }
    }
    if (fabsf(rd.z) > 0.0001f) {
        const float planeZ = rd.z > 0.0f ? roomHalf : -roomHalf;
        const float t = (planeZ - cam.eye.z) / rd.z;
        const F3 q = add3(cam.eye, mul3(rd, t));
        if (t > 0.0f && q.y >= 0.0f && q.y <= ceilingY && fabsf(q.x) <= roomHalf && t < bestT) {
            bestT = t; hit = q; normal = f3(0, 0, planeZ > 0.0f ? -1.0f : 1.0f); surface = 4;
        }
    }
    F3 color = f3(0.014f, 0.014f, 0.014f);
    F3 albedo = f3(0.26f, 0.26f, 0.25f);
    float giScale = 0.24f;
    if (surface == 1) {
        const float slabTone = sinf(hit.x * 2.20f + hit.z * 1.40f) * 0.50f + sinf(hit.z * 3.10f - hit.x * 0.80f) * 0.50f;
        color = cinematic ? f3(0.016f, 0.016f, 0.015f) : f3(0.014f, 0.014f, 0.013f);
        color = add3(color, mul3(f3(0.014f, 0.014f, 0.013f), slabTone * 0.018f));
        albedo = f3(0.26f, 0.26f, 0.25f);
        giScale = 0.22f;
        if (p.sceneId == 1) {
            const float logMat = bed_material(hit.x, hit.z, p);
            const float coal = expf(-(hit.x * hit.x * 1.25f + hit.z * hit.z * 2.45f)) * sstep(0.28f, 0.90f, fbm2(hit.x * 13.0f, hit.z * 15.0f, p.saltNoise2));
            const float scorch = expf(-(hit.x * hit.x * 0.66f + hit.z * hit.z * 1.15f));
            color = lerp3(color, f3(0.018f, 0.015f, 0.012f), sat(logMat * 0.78f + scorch * 0.20f));
            color = add3(color, mul3(f3(0.090f, 0.020f, 0.004f), coal * glow * 0.026f));
            color = add3(color, mul3(f3(0.17f, 0.11f, 0.055f), logMat * 0.050f));
            albedo = lerp3(albedo, f3(0.040f, 0.030f, 0.022f), logMat * 0.86f);
            giScale = lerpf(giScale, 0.44f, logMat * 0.80f);
        } else if (p.sceneId == 2) {
            const float radius = sqrtf(hit.x * hit.x + hit.z * hit.z);
            const float plate = sstep(0.78f, 0.70f, radius);
            const float burner = bed_material(hit.x, hit.z, p);
            color = lerp3(color, f3(0.018f, 0.019f, 0.020f), plate * 0.70f);
            color = lerp3(color, f3(0.005f, 0.005f, 0.005f), burner * 0.72f);
            color = add3(color, mul3(f3(0.090f, 0.032f, 0.006f), burner * glow * 0.012f));
            albedo = lerp3(albedo, f3(0.032f, 0.033f, 0.034f), plate * 0.82f);
            giScale = lerpf(giScale, 0.42f, plate * 0.72f);
        } else {
            const float trayBody = rect_mask(hit.x, hit.z, 0.0f, 0.0f, 1.08f, 0.58f, 0.022f);
            const float trayInner = rect_mask(hit.x, hit.z, 0.0f, 0.0f, 0.96f, 0.47f, 0.020f);
            const float trayRim = sat(trayBody - trayInner * 0.74f);
            const float emberBed = trayInner * expf(-(hit.x * hit.x * 0.74f + hit.z * hit.z * 2.20f));
            const float trayContact = trayInner * expf(-(hit.x * hit.x * 1.35f + hit.z * hit.z * 3.35f));
            const float scorch = expf(-(hit.x * hit.x * 0.82f + hit.z * hit.z * 1.75f));
            color = lerp3(color, f3(0.017f, 0.014f, 0.012f), trayBody * 0.82f);
            color = lerp3(color, f3(0.020f, 0.019f, 0.018f), scorch * (1.0f - trayBody) * 0.16f);
            color = lerp3(color, f3(0.010f, 0.009f, 0.008f), trayContact * 0.22f);
            color = add3(color, mul3(f3(0.056f, 0.014f, 0.003f), emberBed * glow * 0.26f));
            color = add3(color, mul3(f3(0.30f, 0.28f, 0.24f), trayRim * 0.10f));
            albedo = lerp3(albedo, f3(0.030f, 0.024f, 0.020f), trayBody * 0.75f);
            giScale = lerpf(giScale, 0.48f, trayBody * 0.85f);
        }
        const float reflectCore = expf(-(hit.x * hit.x * 2.45f + (hit.z + 0.18f) * (hit.z + 0.18f) * 1.95f)) * glow;
        color = add3(color, mul3(f3(0.55f, 0.11f, 0.023f), reflectCore * p.reflectionGain * 0.48f));
    } else if (surface == 2) {
        color = cinematic ? f3(0.012f, 0.012f, 0.012f) : f3(0.011f, 0.011f, 0.010f);
        albedo = f3(0.24f, 0.24f, 0.23f);
        giScale = 0.20f;
        const float panelA = line_mask(hit.x, hit.z, -1.2f, -2.0f, -1.2f, 2.0f, 0.010f);
        const float panelB = line_mask(hit.x, hit.z, 0.0f, -2.0f, 0.0f, 2.0f, 0.008f);
        const float fixL = expf(-((hit.x + 1.50f) * (hit.x + 1.50f) + (hit.z - 0.92f) * (hit.z - 0.92f)) * 120.0f);
        const float fixR = expf(-((hit.x - 1.72f) * (hit.x - 1.72f) + (hit.z - 0.80f) * (hit.z - 0.80f)) * 120.0f);
        color = add3(color, mul3(f3(0.10f, 0.096f, 0.088f), (panelA + panelB) * 0.08f));
        color = add3(color, mul3(f3(0.26f, 0.20f, 0.13f), (fixL + fixR) * 0.10f));
    } else {
        color = cinematic ? f3(0.015f, 0.015f, 0.014f) : color;
        albedo = f3(0.25f, 0.25f, 0.24f);
        giScale = 0.24f;
        const float baseShadow = sstep(0.34f, 0.02f, hit.y);
        const float crownShadow = sstep(ceilingY - 0.32f, ceilingY - 0.02f, hit.y);
        color = lerp3(color, f3(0.012f, 0.012f, 0.011f), sat(baseShadow * 0.30f + crownShadow * 0.20f));
        if (surface == 4 && hit.z > 0.0f) {
            const float recess = rect_mask(hit.x, hit.y, 0.0f, 1.00f, 0.92f, 0.72f, 0.055f);
            const float recessEdge = line_mask(hit.x, hit.y, -0.92f, 0.28f, -0.92f, 1.72f, 0.010f) + line_mask(hit.x, hit.y, 0.92f, 0.28f, 0.92f, 1.72f, 0.010f);
            color = lerp3(color, f3(0.015f, 0.015f, 0.014f), recess * 0.42f);
            color = add3(color, mul3(f3(0.050f, 0.046f, 0.038f), recessEdge * 0.030f));
        }
        const float soot = expf(-(hit.x * hit.x * 0.82f + hit.z * hit.z * 1.12f)) * sstep(0.36f, 1.92f, hit.y);
        color = lerp3(color, f3(0.010f, 0.011f, 0.012f), soot * 0.82f * p.smokeDarkness);
        albedo = lerp3(albedo, f3(0.030f, 0.032f, 0.034f), soot * 0.75f);
        const float wallGlow = expf(-(hit.x * hit.x * 1.1f + hit.z * hit.z * 1.3f + (hit.y - 0.70f) * (hit.y - 0.70f) * 1.0f)) * glow;
        color = add3(color, mul3(f3(0.15f, 0.050f, 0.018f), wallGlow * 0.090f));
        if (surface == 3 && hit.x < 0.0f) {
            const float window = rect_mask(hit.z, hit.y, -1.28f, 0.92f, 0.06f, 0.62f, 0.030f);
            const float glass = 0.55f + 0.25f * sinf(hit.y * 23.0f + hit.z * 15.0f);
            color = lerp3(color, f3(0.009f, 0.011f, 0.012f), window * 0.92f);
            color = add3(color, mul3(f3(0.10f, 0.12f, 0.13f), window * glass * 0.16f));
        }
    }
    const F3 vuv = uv_from_world(hit);
    const float4 gi = sample_light(L, p, vuv.x, vuv.y, vuv.z);
    const float sceneShadow = sample_shadow(SH, p, vuv.x, vuv.y, vuv.z);
    const float volVis = volume_shadow_ray(hit, L, SH, p);
    const F3 gathered = gather_irradiance(hit, normal, L, SH, p);
    const F3 indirect = mul3(add3(mul3(f3(gi.x, gi.y, gi.z), 0.42f), gathered), volVis);
    const float grazingFalloff = surface == 2 ? 0.62f : 1.0f;
    const float contactOcc = sstep(0.005f, 0.42f, sceneShadow) * expf(-gi.w * 0.72f) * (0.30f + volVis * 0.70f);
    const float lift = 0.66f + sstep(0.006f, 0.13f, lum3(indirect)) * 0.72f;
    const float warmth = surface == 1 ? 1.72f : (surface == 2 ? 0.88f : 1.16f);
    color = add3(color, mul3(mul3v(indirect, albedo), giScale * grazingFalloff * contactOcc * lift * warmth));
    if (p.dirLight) {
        const float ndl = fmaxf(0.0f, -(normal.x * p.dirLightDir[0] + normal.y * p.dirLightDir[1] + normal.z * p.dirLightDir[2]));
        const float td = sample_dir_trans(DT, p, vuv.x, vuv.y, vuv.z);
        color = add3(color, mul3(mul3v(albedo, f3(p.dirLightColor[0], p.dirLightColor[1], p.dirLightColor[2])), ndl * td * kDirDiffuse));
    }
    const float shadowedSoot = (1.0f - volVis) * sstep(0.18f, 0.82f, 1.0f - sceneShadow);
    color = lerp3(color, f3(0.006f, 0.007f, 0.008f), shadowedSoot * (surface == 2 ? 0.38f : 0.26f));
    const float wallEdge = surface == 3 ? roomHalf - fabsf(hit.z) : roomHalf - fabsf(hit.x);
    const float corner = (surface == 1 || surface == 2) ? fminf(roomHalf - fabsf(hit.x), roomHalf - fabsf(hit.z))
                                                        : fminf(wallEdge, fminf(hit.y, ceilingY - hit.y));
    color = mul3(color, 0.84f + sstep(0.0f, 0.54f, corner) * 0.16f);
    const float gx = floorf(hit.x * 19.0f + hit.z * 13.0f), gy = floorf(hit.y * 15.0f + time * 0.08f);
    const float grain = (hash2f(gx, gy, static_cast<int>(gx), static_cast<int>(gy), p.saltGrain) - 0.5f) * 0.006f;
    color = add3(color, f3(grain, grain, grain));
    source_model_overlay(uvx, uvy, cam, p, color);
    const float fog = sstep(3.8f, 0.2f, bestT);
    const float vignette = sstep(0.78f, 0.20f, sqrtf((uvx - 0.50f) * (uvx - 0.50f) + (uvy - 0.46f) * (uvy - 0.46f)));
    return mul3(color, (0.12f + fog * 0.56f) * (0.58f + vignette * 0.30f));
}

// src/fire_cuda.cu:1867-1907
__device__ __forceinline__ bool intersect_box(F3 ro, F3 rd, float& tn, float& tf) {
    const float ro3[3] = {ro.x, ro.y, ro.z}, rd3[3] = {rd.x, rd.y, rd.z};
    const float lo3[3] = {-1.05f, 0.02f, -0.82f}, hi3[3] = {1.05f, 2.05f, 0.82f};
    float tNear = 0.0f, tFar = 1.0e6f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float d = rd3[a];
        const float inv = 1.0f / (fabsf(d) < 0.00001f ? copysignf(0.00001f, d) : d);
        float t0 = (lo3[a] - ro3[a]) * inv, t1 = (hi3[a] - ro3[a]) * inv;
        if (t0 > t1) { const float tmp = t0; t0 = t1; t1 = tmp; }
        tNear = fmaxf(tNear, t0);
        tFar = fminf(tFar, t1);
    }
    tn = tNear;
    tf = tFar;
    return tFar > tNear;
}

__device__ __forceinline__ F3 debug_heat_color(float value) {
    value = sat(value);
    const float red = sstep(0.05f, 0.45f, value), orange = sstep(0.24f, 0.78f, value), white = sstep(0.72f, 1.0f, value);
    return f3(sat(red + white * 0.70f), sat(orange * 0.58f + white * 0.78f), sat(white * 0.62f));
}

template <bool B>
__global__ __launch_bounds__(256) void k_render(float4* __restrict__ out, Scalars<B> S, const float4* __restrict__ L, const float* __restrict__ SH,
                                                const float* __restrict__ DT, const float* __restrict__ U, const float* __restrict__ V,
                                                const float* __restrict__ W, DevParams p) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= p.fw || y >= p.fh) return;
    const Cam cam = cam_of(p);
    const float uvx = (static_cast<float>(x) + 0.5f) / static_cast<float>(p.fw);
    const float uvy = (static_cast<float>(y) + 0.5f) / static_cast<float>(p.fh);
    const float sx = (uvx * 2.0f - 1.0f) * cam.aspect * cam.tanHalfFov;
    const float sy = (1.0f - uvy * 2.0f) * cam.tanHalfFov;
    const F3 rd = normalize3(add3(add3(cam.fwd, mul3(cam.right, sx)), mul3(cam.up, sy)));
    const float baseGlow = sat(sample_scalar<HEAT, B>(S, p, 0.50f, 0.06f, 0.50f) * 0.42f);
    float boxNear = 0.0f, boxFar = 0.0f;
    F3 accum = f3(0.0f, 0.0f, 0.0f);
    float trans = 1.0f, flameGlow = 0.0f, smokeOcc = 0.0f;
    float dbgFlame = 0.0f, dbgOD = 0.0f, dbgTemp = 0.0f, dbgFuel = 0.0f, dbgVel = 0.0f;
    const int sid = p.sceneId;
    const float t = ld_at(p.clock, 1u);
    if (intersect_box(cam.eye, rd, boxNear, boxFar)) {
        const int steps = min(max(p.raymarchSteps, 24), 256);
        const float t0 = fmaxf(0.0f, boxNear);
        const float t1 = fminf(boxFar, t0 + 3.05f);
        const float stepT = (t1 - t0) / static_cast<float>(steps);
        const float jitter = hash2f((static_cast<float>(x) + 0.5f) * 0.754877f, (static_cast<float>(y) + 0.5f) * 0.569840f, x, y, p.saltJitter);
        const float cellX = 1.0f / static_cast<float>(p.nx);
        const float cellY = 1.0f / static_cast<float>(p.ny);
        const float cellZ = 1.0f / static_cast<float>(p.nz);
        for (int i = 0; i < steps && trans > 0.018f; ++i) {
            const float tt = t0 + (static_cast<float>(i) + jitter) * stepT;
            const F3 wp = add3(cam.eye, mul3(rd, tt));
            const float fu = sat((wp.x + 1.05f) / 2.10f), fv = sat((wp.y - 0.02f) / 2.03f), fw = sat((wp.z + 0.82f) / 1.64f);
            const float fade = domain_fade(fu, fv, fw);
            if (fade <= 0.001f) continue;
            const Tri tc = tri_scalar_setup(p, fu, fv, fw);
            const float cHeat = tri_scalar<HEAT, B>(S, tc);
            const float cFuel = tri_scalar<FUEL, B>(S, tc);
            const float cSoot = tri_scalar<SOOT, B>(S, tc);
            const float cProg = tri_scalar<PROGRESS, B>(S, tc);
            const float coarse = (cHeat + cFuel * 0.42f + cSoot * 0.34f + cProg * 0.32f) * fade;
            if (coarse <= 0.00150f) {
                ++i;
                continue;
            }
            const float warpA = fbm3(fu * 7.0f + t * 0.18f, fv * 5.0f - t * 0.11f, fw * 9.0f + t * 0.07f, p.saltNoise3) - 0.5f;
            const float warpB = fbm3(fw * 8.0f - t * 0.14f, fu * 8.0f + t * 0.10f, fv * 6.0f - t * 0.08f, p.saltNoise3) - 0.5f;
            const float warpGain = 0.028f + fv * 0.060f;
            const float wu = sat(fu + warpA * warpGain), wv = sat(fv + (warpA - warpB) * 0.020f), ww = sat(fw + warpB * warpGain);
            const Tri tw = tri_scalar_setup(p, wu, wv, ww);
            const float heat = tri_scalar<HEAT, B>(S, tw);
            const float fuel = tri_scalar<FUEL, B>(S, tw);
            const float oxygen = tri_scalar<OXYGEN, B>(S, tw);
            const float soot = tri_scalar<SOOT, B>(S, tw);
            const float charMass = tri_scalar<CHARF, B>(S, tw);
            const float ash = tri_scalar<ASH, B>(S, tw);
            const float pyro = tri_scalar<PYRO, B>(S, tw);
            const float progress = tri_scalar<PROGRESS, B>(S, tw);
            const float turb = tri_scalar<TURB, B>(S, tw);
            const float sootOptics = tri_scalar<SOPT, B>(S, tw);
            const float4 sceneLight = sample_light(L, p, wu, wv, ww);
            const float sceneShadow = sample_shadow(SH, p, wu, wv, ww);
            const float active = (heat + fuel * 0.42f + soot * 0.34f + progress * 0.32f + pyro * 0.18f) * fade;
            if (active <= 0.0025f) continue;
            dbgFuel = fmaxf(dbgFuel, fade * sat(fuel * 0.18f + charMass * 0.22f + pyro * 0.30f));
            if (p.debugMode == 6) {
                const F3 vel = sample_vel(U, V, W, p, wu, wv, ww);
                dbgVel = fmaxf(dbgVel, sat(sqrtf(dot3(vel, vel)) * 0.22f));
            }
            const float gx = sample_scalar<HEAT, B>(S, p, wu + cellX, wv, ww) - sample_scalar<HEAT, B>(S, p, wu - cellX, wv, ww);
            const float gy = sample_scalar<HEAT, B>(S, p, wu, wv + cellY, ww) - sample_scalar<HEAT, B>(S, p, wu, wv - cellY, ww);
            const float gz = sample_scalar<HEAT, B>(S, p, wu, wv, ww + cellZ) - sample_scalar<HEAT, B>(S, p, wu, wv, ww - cellZ);
            const float frontGradient = sqrtf(gx * gx + gy * gy + gz * gz) + fabsf(fuel - cFuel) * 0.62f;
            const float reactionFront = sstep(0.020f, 0.46f, frontGradient * 1.35f + fuel * oxygen * 0.040f + progress * 0.18f + pyro * 0.080f);
            const float fine = fbm3(wu * 42.0f + t * 0.72f, wv * 54.0f - t * 1.28f, ww * 43.0f + t * 0.38f, p.saltNoise3);
            const float shear = fbm3(wu * 15.0f - t * 0.72f, wv * 13.0f - t * 1.08f, ww * 17.0f + t * 0.32f, p.saltNoise3);
            const float plume = sat(fine * 0.58f + shear * 0.42f);
            const float hole = sat((1.0f - fine) * 0.36f + shear * 0.64f);
            const float les = sat(turb * 0.48f + progress * 0.28f);
            const float ragged = sstep(0.18f, 0.92f, fine * (0.46f + les * 0.30f) + shear * 0.38f + plume * 0.32f + sat(active * 0.08f));
            const float fieldEdge = sstep(0.05f, 2.30f, heat + fuel * 0.32f) * (0.10f + reactionFront * 0.90f);
            const float filament = sstep(0.60f, 1.0f, fine + les * 0.10f) * sstep(0.38f, 0.92f, shear + (0.72f - fv) * 0.12f + reactionFront * 0.10f) *
                                   sstep(1.0f, 0.02f, fv) * fieldEdge;
            const float srcV = sid == 2 ? fmaxf(0.0f, fv - p.emitterHeightNorm) : fv;
            const float lowJet = sid == 2 ? sstep(fmaxf(0.012f, p.emitterHeightBandNorm * 0.68f), fmaxf(0.003f, p.emitterHeightBandNorm * 0.10f),
                                                  fabsf(fv - p.emitterHeightNorm)) *
                                                sstep(0.05f, 0.58f, fuel + heat * 0.025f) * sstep(0.50f, 1.0f, oxygen)
                                          : 0.0f;
            const float portJet = sid == 2 ? lowJet * sstep(0.68f, 0.98f, fine + shear * 0.18f) : 0.0f;
            const float campBias = sid == 1 ? sstep(0.035f, 0.72f, fuel + pyro * 0.50f + charMass * 0.12f) * sstep(0.76f, 0.035f, fv) : 0.0f;
            const float traySheet = sid == 0 ? sstep(0.018f, 0.42f, fuel + pyro * 0.72f + charMass * 0.08f) * sstep(0.30f, 0.015f, fv) : 0.0f;
            const float convective = sstep(0.42f, 0.88f, shear + heat * 0.040f + plume * 0.14f + les * 0.10f + portJet * 0.56f + campBias * 0.20f + traySheet * 0.24f) *
                                     (sid == 2 ? sstep(p.emitterHeightNorm + p.emitterHeightBandNorm * 1.20f, p.emitterHeightNorm - p.emitterHeightBandNorm * 0.08f, fv)
                                               : sstep(sid == 1 ? 0.82f : 0.88f, 0.030f, fv)) *
                                     fieldEdge;
            const float breakup = sstep(0.34f, 0.88f, fine * 0.48f + shear * 0.42f + hole * 0.24f + heat * 0.026f);
            const float verticalFade = sstep(0.96f, 0.025f, fv);
            const float heightFade = sid == 2 ? sstep(p.emitterHeightNorm + p.emitterHeightBandNorm * 0.72f, p.emitterHeightNorm - p.emitterHeightBandNorm * 0.05f, fv)
                                              : sstep(sid == 1 ? 0.72f : 0.76f, 0.10f, fv);
            const float lowerWhite = sstep(0.060f, 0.00f, srcV) * sstep(0.05f, 1.30f, pyro + fuel * 0.20f);
            const float tempK = 293.0f + heat * 360.0f + fuel * 44.0f + pyro * 90.0f + filament * 560.0f + convective * 420.0f + lowerWhite * 110.0f + progress * 64.0f;
            const float combustion = sstep(740.0f, 1630.0f, tempK) * sstep(0.016f, 0.62f, fuel + heat * 0.055f) * sstep(0.045f, 0.62f, oxygen) * verticalFade * heightFade;
            const float thinFront = sstep(0.04f, 0.56f, frontGradient + progress * 0.10f) * sstep(0.02f, 0.92f, reactionFront);
            const float tear = sstep(0.52f, 0.88f, hole + shear * 0.18f + les * 0.30f + sat(1.0f - oxygen) * 0.12f);
            const float sheetHole = sstep(0.54f, 0.90f, hole + les * 0.26f + sat(1.0f - oxygen) * 0.12f);
            const float flameSheet = fmaxf(sstep(0.55f, 0.94f, fine * 0.46f + shear * 0.44f + plume * 0.18f + reactionFront * 0.15f + heat * 0.014f - fv * 0.050f) *
                                               (1.0f - sheetHole * (0.74f + les * 0.12f)) * (1.0f - tear * 0.18f),
                                           sstep(0.18f, 0.030f, srcV) * sstep(0.75f, 2.0f, heat) * thinFront * 0.030f);
            const float coherent = sat(filament * 0.86f + convective * 0.55f + thinFront * 0.74f);
            const float confinement = sat(coherent + reactionFront * 0.08f);
            const float coverage = sid == 0 ? (0.72f + traySheet * 1.24f) : 1.0f;
            const float flameDensity = fade * combustion * reactionFront * flameSheet * (0.010f + filament * 0.92f + convective * 0.24f + thinFront * 1.22f) *
                                       (0.052f + breakup * 0.78f + ragged * 0.56f + les * 0.34f) * (0.18f + thinFront * 2.55f) *
                                       (0.060f + confinement * 0.94f + portJet * 0.72f + campBias * 0.34f) * heightFade * coverage;
            const float plumeVoid = sstep(0.54f, 0.90f, hole + fine * 0.26f + shear * 0.22f + fv * 0.20f);
            const float topDissolve = sstep(sid == 1 ? 0.82f : 0.88f, sid == 1 ? 0.34f : 0.42f, fv);
            const float raggedPlume = 1.0f - plumeVoid * sstep(0.18f, 0.74f, fv) * 0.96f;
            const float upperPlume = sstep(0.14f, 0.48f, fv) * sstep(0.030f, 1.20f, sootOptics + soot * 0.42f) *
                                     (0.025f + fine * 0.035f + shear * 0.034f + plume * 0.34f + ragged * 0.12f + turb * 0.035f) * topDissolve * raggedPlume;
            const float smokeScale = sid == 2 ? 0.035f : (sid == 1 ? 0.56f : 1.0f);
            const float rawSoot = fade * smokeScale *
                                  (sstep(0.020f, 1.05f, sootOptics + soot * 0.32f) * sstep(0.16f, 0.42f, fv) *
                                       (0.028f + fine * 0.040f + shear * 0.052f + plume * 0.38f + ragged * 0.16f + ash * 0.016f) +
                                   upperPlume * (0.22f + plume * 0.30f) + sstep(0.004f, 0.34f, sootOptics) * (0.020f + plume * 0.034f));
            const float frontMask = sat(flameDensity * 4.8f + combustion * reactionFront * flameSheet * (0.08f + thinFront * 0.56f + filament * 0.18f));
            const float emissive = sat(frontMask);
            const float resolved = sstep(0.030f, 0.32f, emissive) * sstep(0.035f, 0.72f, oxygen);
            const float smokeOnly = 1.0f - sstep(0.025f, 0.20f, resolved);
            const float scatterSep = sat(smokeOnly * (1.0f - sstep(0.02f, 0.16f, emissive) * 0.58f));
            const float cleanout = 1.0f - sstep(0.025f, 0.34f, resolved) * 0.82f;
            const float smokeOnlyHeight = sstep(0.20f, 0.70f, fv) * (1.0f - sstep(0.02f, 0.20f, emissive));
            const float neutrality = sstep(0.18f, 0.78f, fv) * scatterSep;
            const float absorption = rawSoot * cleanout * (1.10f + smokeOnlyHeight * 2.20f + neutrality * 2.40f);
            const float scatterDensity = rawSoot * scatterSep * sstep(0.10f, 0.55f, fv + sootOptics * 0.10f) * (0.24f + (1.0f - smokeOnlyHeight) * 0.34f);
            const float radiusP = sat(0.08f + sootOptics * 0.028f + ash * 0.085f + sat(1.0f - oxygen) * 0.12f);
            const float sootAbs = absorption * (3.10f + radiusP * 4.60f + sootOptics * 0.70f + neutrality * 2.85f) * p.smokeDarkness;
            const float sootScat = scatterDensity * (0.018f + (1.0f - radiusP) * 0.050f) * (0.20f + p.smokeGain * 0.10f);
            const float sootExt = sootAbs + sootScat;
            const float flameExt = flameDensity * (0.12f + (1.0f - resolved) * 0.12f);
            const float extinction = (sootExt + flameExt) * stepT;
            const float smokeAlpha = 1.0f - expf(-sootAbs * stepT);
            const float scatterAlpha = 1.0f - expf(-sootScat * stepT);
            const float lightVis = sceneShadow * (0.35f + sceneShadow * 0.65f);
            const F3 localIrr = f3(sceneLight.x, sceneLight.y, sceneLight.z);
            const float sootLoad = sat(sootOptics * 0.74f + soot * 0.20f + upperPlume * 0.46f);
            const float whiteCore = sstep(2450.0f, 3450.0f, tempK) * combustion * sstep(0.32f, 0.90f, oxygen) * sstep(0.72f, 0.02f, fv) *
                                    (0.02f + thinFront * 0.48f + filament * 0.18f) * (1.0f - sstep(0.16f, 0.74f, sootLoad) * 0.68f);
            const float radiantBase = sat((tempK - 760.0f) / 2000.0f);
            const float radiantCurve = radiantBase * radiantBase * (0.82f + radiantBase * 0.18f);
            const float radiantPower = radiantCurve * (1.42f + filament * 3.15f + convective * 1.60f + progress * 0.26f + whiteCore * 0.92f);
            const float coreStructure = sat(whiteCore * (0.080f + filament * 0.28f + thinFront * 0.32f));
            const float whiteFil = coreStructure * (0.55f + fine * 0.30f + shear * 0.18f);
            const float orangeEdge = (1.0f - whiteFil * 0.72f) * thinFront * reactionFront * flameSheet * (0.52f + breakup * 0.95f) * sstep(0.08f, 0.80f, oxygen);
            F3 flameColor = lerp3(f3(1.32f, 0.27f, 0.038f), blackbody(tempK), sat(whiteCore * 0.92f + whiteFil * 0.82f));
            if (sid == 2) {
                const float blueBase = portJet * sstep(0.010f, 0.11f, flameDensity + combustion * 0.18f);
                flameColor = lerp3(f3(0.020f, 0.18f, 2.40f), f3(0.36f, 0.58f, 1.95f), sat(whiteCore * 0.34f + blueBase * 0.20f));
            } else if (sid == 1) {
                flameColor = lerp3(flameColor, f3(1.52f, 0.24f, 0.028f), sat(campBias * (1.0f - whiteCore) * 0.58f));
            }
            F3 emission = mul3(flameColor, flameDensity * radiantPower * stepT * (1.10f + whiteCore * 0.38f));
            if (sid == 2) {
                emission = mul3(emission, 0.42f);
                emission = add3(emission, mul3(f3(0.012f, 0.28f, 5.20f), portJet * combustion * stepT * 4.40f));
            }
            emission = add3(emission, mul3(f3(1.10f, 1.02f, 0.86f), whiteFil * whiteFil * flameDensity * radiantPower * stepT * 0.82f));
            emission = add3(emission, mul3(f3(1.70f, 0.30f, 0.040f), orangeEdge * flameDensity * radiantPower * stepT * (sid == 2 ? 0.06f : (sid == 1 ? 3.20f : 3.10f))));
            emission = add3(emission, mul3(f3(1.36f, 0.070f, 0.008f), filament * (1.0f + tear * 0.55f) * flameDensity * radiantPower * stepT *
                                                                          (sid == 2 ? 0.035f : (sid == 1 ? 1.55f : 1.10f)) * (1.0f - whiteCore * 0.62f)));
            const float ashVeil = sat(ash * 0.014f + sstep(0.30f, 0.88f, fv) * (1.0f - sootLoad) * 0.003f);
            F3 smokeColor = lerp3(f3(0.008f, 0.0092f, 0.0115f), f3(0.0018f, 0.0020f, 0.0024f), sootLoad);
            smokeColor = lerp3(smokeColor, f3(0.014f, 0.0165f, 0.0210f), ashVeil);
            smokeColor = mul3(smokeColor, smokeOnly);
            const float irrLuma = lum3(localIrr);
            const float warmDamp = 1.0f - neutrality * 0.985f;
            const float warmLeak = (1.0f - neutrality) * (1.0f - smokeOnlyHeight) * 0.18f;
            const float warmScatter = sid == 2 ? 0.010f : (sid == 1 ? 0.18f : 0.30f);
            const float backScatter = scatterAlpha * lightVis * (0.00085f + irrLuma * 0.0011f * warmDamp + baseGlow * 0.00010f * warmScatter) *
                                      (0.018f + (1.0f - radiusP) * 0.018f) * scatterSep;
            smokeColor = add3(smokeColor, mul3(f3(0.010f, 0.012f, 0.016f), backScatter));
            const float scatterGain = scatterAlpha * (0.00046f + lightVis * 0.00072f * warmDamp + flameDensity * 0.00008f * warmScatter + turb * 0.00018f) *
                                      (0.038f + (1.0f - radiusP) * 0.045f) * scatterSep;
            const float sceneScatter = scatterAlpha * scatterSep * (0.0065f + (1.0f - radiusP) * 0.014f) * lightVis;
            const F3 neutralLight = f3(irrLuma * 0.052f, irrLuma * 0.064f, irrLuma * 0.086f);
            const F3 smokeLight = lerp3(neutralLight, localIrr, warmLeak);
            const float coalGlow = sstep(0.004f, 0.18f, charMass) * sstep(0.54f, 0.02f, fv) * (0.070f + pyro * 0.12f + heat * 0.009f);
            dbgFlame = fmaxf(dbgFlame, sat(flameDensity * radiantPower * 0.26f));
            dbgOD = sat(dbgOD + sootExt * stepT);
            dbgTemp = fmaxf(dbgTemp, sat((tempK - 520.0f) / 1900.0f));
            accum = add3(accum, mul3(emission, trans));
            accum = add3(accum, mul3(f3(1.0f, 0.24f, 0.035f), trans * coalGlow * stepT));
            accum = add3(accum, mul3(smokeColor, trans * scatterGain * smokeOnly));
            accum = add3(accum, mul3(smokeLight, trans * sceneScatter * (0.18f + warmDamp * 0.42f)));
            if (p.dirLight) {
                const float td = sample_dir_trans(DT, p, wu, wv, ww);
                accum = add3(accum, mul3(f3(p.dirLightColor[0], p.dirLightColor[1], p.dirLightColor[2]), trans * scatterAlpha * td * kDirScatter));
            }
            flameGlow = fmaxf(flameGlow, flameDensity * radiantPower * (0.12f + filament * 0.18f + whiteFil * 0.20f));
            smokeOcc = sat(smokeOcc + smokeAlpha * trans);
            trans *= expf(-extinction);
        }
    }
    const size_t o = static_cast<size_t>(y) * p.fw + x;
    if (p.debugMode != 0) {
        F3 c;
        if (p.debugMode == 1) c = debug_heat_color(dbgFlame);
        else if (p.debugMode == 2) { const float od = sat(dbgOD); c = f3(od, od * 0.90f, od * 0.78f); }
        else if (p.debugMode == 3) c = f3(trans, trans, trans);
        else if (p.debugMode == 4) c = debug_heat_color(dbgTemp);
        else if (p.debugMode == 5) c = f3(dbgFuel * 0.82f, dbgFuel * 0.48f, dbgFuel * 0.18f);
        else c = f3(dbgVel * 0.20f, dbgVel * 0.72f, dbgVel);
        out[o] = make_float4(c.x, c.y, c.z, 1.0f);
        return;
    }
    F3 color = accum;
    if (trans > 0.006f) color = add3(mul3(room_background(cam, uvx, uvy, rd, baseGlow, L, SH, DT, p, t), trans), accum);
    const float bloom = sat(flameGlow * 0.10f) * (1.0f - sat(smokeOcc * 0.95f));
    color = add3(color, mul3(f3(1.0f, 0.25f, 0.045f), bloom * bloom * 0.017f * p.reflectionGain));
    out[o] = make_float4(fmaxf(0.0f, color.x), fmaxf(0.0f, color.y), fmaxf(0.0f, color.z), 1.0f);
}

// ---------------------------------------------------------------- K21 ----
__device__ __forceinline__ float aces(float v) {
    return sat((v * (2.51f * v + 0.03f)) / (v * (2.43f * v + 0.59f) + 0.14f));
}

// Gizmos (drawGizmos, src/fire_cuda.cu:1909-1951): a ring and dot on the
// projected source, then the world x/y/z axes from a corner anchor, alpha-
// blended over the 8-bit colour in frame uv. The reference runs them as a
// second kernel over the packed frame (overlayKernel :2468-2494, launched
// after the tonemap at :3419-3425) that unpacks each pixel to value/255; the
// tonemap below does the same unpack from its own packed bytes, so one pass
// gives the same bytes without a second read and write of the frame.
__device__ __forceinline__ float seg_alpha(float px, float py, float ax, float ay, float bx, float by, float thickness) {
    const float pax = px - ax, pay = py - ay, bax = bx - ax, bay = by - ay;
    const float h = sat((pax * bax + pay * bay) / fmaxf(0.000001f, bax * bax + bay * bay));
    const float dx = pax - bax * h, dy = pay - bay * h;
    return sstep(thickness, thickness * 0.35f, sqrtf(dx * dx + dy * dy));
}
__device__ __forceinline__ void draw_gizmos(F3& color, float uvx, float uvy, const Cam& cam, const DevParams& p) {
    const bool burner = p.sceneId == 2 && p.burnerCount > 0;
    const F3 src = project_point(cam, f3(burner ? p.burnerX[0] : p.emitterCenterX, burner ? p.burnerY[0] : p.emitterHeightNorm * 2.03f + 0.02f,
                                         burner ? p.burnerZ[0] : p.emitterCenterZ));
    if (src.z > 0.0f) {
        const float dx = uvx - src.x, dy = uvy - src.y;
        const float r = sqrtf(dx * dx + dy * dy);
        const float ring = sstep(0.006f, 0.006f * 0.25f, fabsf(r - 0.125f / src.z));
        const float dot = sstep(0.012f, 0.012f * 0.45f, r);
        const F3 tint = p.activeGizmo == 1 ? f3(1.0f, 0.34f, 0.055f) : f3(0.45f, 0.18f, 0.055f);
        color = lerp3(color, tint, sat((ring + dot) * 0.85f));
    }
    // axis tips: base + 0.07 * (e . right, -(e . up)) for the unit axes e
    const float bx = 0.90f, by = 0.84f;
    color = lerp3(color, f3(1.0f, 0.12f, 0.08f), sat(seg_alpha(uvx, uvy, bx, by, bx + cam.right.x * 0.070f, by - cam.up.x * 0.070f, 0.005f)));
    color = lerp3(color, f3(0.22f, 1.0f, 0.30f), sat(seg_alpha(uvx, uvy, bx, by, bx + cam.right.y * 0.070f, by - cam.up.y * 0.070f, 0.005f)));
    color = lerp3(color, f3(0.20f, 0.48f, 1.0f), sat(seg_alpha(uvx, uvy, bx, by, bx + cam.right.z * 0.070f, by - cam.up.z * 0.070f, 0.005f)));
}

__global__ __launch_bounds__(256) void k_tonemap(uint32_t* __restrict__ out, const float4* __restrict__ hdr, DevParams p) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= p.fw || y >= p.fh) return;
    const size_t i = static_cast<size_t>(y) * p.fw + x;
    const float4 r = hdr[i];
    F3 c = mul3(f3(fmaxf(0.0f, r.x), fmaxf(0.0f, r.y), fmaxf(0.0f, r.z)), p.exposure);
    const float l0 = fmaxf(0.000001f, lum3(c));
    c = mul3(c, aces(l0) / l0);
    const float luma = lum3(c);
    const float toe = sstep(0.000f, 0.055f, luma);
    const float shoulder = sstep(0.78f, 1.08f, luma);
    c = lerp3(mul3(c, 0.82f), c, toe);
    c = lerp3(c, mul3(c, 1.0f / fmaxf(1.0f, luma / 0.94f)), shoulder * 0.10f);
    c = mul3(c, 1.0f + sstep(0.56f, 0.90f, luma) * 0.28f);
    const float g = 1.0f / 2.2f;
    c = f3(powf(sat(c.x), g), powf(sat(c.y), g), powf(sat(c.z), g));
    const float dither = (hash2f(static_cast<float>(x) * 0.754877f, static_cast<float>(y) * 0.569840f, x, y, p.saltDither) - 0.5f) / 255.0f;
    c = add3(c, f3(dither, dither, dither));
    uint32_t rr = static_cast<uint32_t>(sat(c.x) * 255.0f);
    uint32_t gg = static_cast<uint32_t>(sat(c.y) * 255.0f);
    uint32_t bb = static_cast<uint32_t>(sat(c.z) * 255.0f);
    if (p.showGizmos && p.debugMode == 0) {
        F3 g8 = f3(static_cast<float>(rr) / 255.0f, static_cast<float>(gg) / 255.0f, static_cast<float>(bb) / 255.0f);
        draw_gizmos(g8, (static_cast<float>(x) + 0.5f) / static_cast<float>(p.fw), (static_cast<float>(y) + 0.5f) / static_cast<float>(p.fh), cam_of(p), p);
        rr = static_cast<uint32_t>(sat(g8.x) * 255.0f);
        gg = static_cast<uint32_t>(sat(g8.y) * 255.0f);
        bb = static_cast<uint32_t>(sat(g8.z) * 255.0f);
    }
    out[i] = 0xff000000u | (rr << 16) | (gg << 8) | bb;
}

// ------------------------------------------------------------- embers ----
// emberHdrKernel (src/fire_cuda.cu:2323-2452) splats sparks into the HDR frame
// with float atomicAdd, whose summation order is unspecified. Here it is two
// passes: k_ember_prepare computes every spark (one thread each), then
// k_ember_splat has each pixel add the sparks that cover it in spark order, so
// the frame is deterministic and needs no atomics. The reference's sin-hash
// per spark (hash21(i, {3.7, 9.1, 19.4, 31.6})) is the integer hash of (i, k).

template <bool B>
__global__ __launch_bounds__(kMaxEmbers) void k_ember_prepare(Ember* __restrict__ out, Scalars<B> S, const float* __restrict__ U,
                                                             const float* __restrict__ V, const float* __restrict__ W, DevParams p) {
    const int i = threadIdx.x;
    Ember e{};
    e.active = 0;
    if (i >= kMaxEmbers) return;
    const int count = min(max(p.emberCount, 0), kMaxEmbers);
    const float time = ld_at(p.clock, 1u);
    const int sid = p.sceneId;
    const float baseGlow = sat(sample_scalar<HEAT, B>(S, p, 0.50f, 0.06f, 0.50f) * 0.42f);
    const float sparkGate = sstep(0.02f, 0.30f, baseGlow);
    const float seed = static_cast<float>(i);
    const float h0 = hash2f(seed, 3.7f, i, 0, p.saltEmber), h1 = hash2f(seed, 9.1f, i, 1, p.saltEmber);
    const float h2 = hash2f(seed, 19.4f, i, 2, p.saltEmber), h3 = hash2f(seed, 31.6f, i, 3, p.saltEmber);
    bool live = i < count && p.debugMode == 0 && sparkGate > 0.0001f && !(sid == 1 && h2 > 0.34f) && !(sid == 2 && h0 > 0.14f);
    if (live) {
        const float lt = time * (0.20f + h2 * 0.18f + (sid == 2 ? 0.36f : 0.0f)) + h0;
        const float life = lt - floorf(lt);
        const float angle = h0 * 6.2831853f;
        float startX = 0.0f, startZ = 0.0f;
        if (sid == 2) {
            const float radius = fmaxf(0.04f, p.emitterRadius);
            const float cx = p.burnerCount > 0 ? p.burnerX[0] : p.emitterCenterX;
            const float cz = p.burnerCount > 0 ? p.burnerZ[0] : p.emitterCenterZ;
            startX = cx + cosf(angle) * radius * (0.82f + h1 * 0.18f);
            startZ = cz + sinf(angle) * radius * (0.82f + h1 * 0.18f);
        } else if (sid == 1) {
            const float logA = capsule_mask(cosf(angle) * 0.34f, sinf(angle) * 0.25f, -0.06f, 0.01f, 0.38f, 0.34f, 0.095f);
            const float logB = capsule_mask(cosf(angle + 1.7f) * 0.35f, sinf(angle + 1.7f) * 0.25f, 0.07f, -0.02f, -0.58f, 0.33f, 0.092f);
            const float contact = sat(logA + logB + h3 * 0.35f);
            const float radius = lerpf(0.16f, 0.46f, sqrtf(h1));
            startX = cosf(angle) * radius * (0.64f + contact * 0.36f);
            startZ = sinf(angle) * radius * (0.46f + contact * 0.26f);
        } else {
            const float ring = sqrtf(h1);
            startX = cosf(angle) * 0.48f * ring;
            startZ = sinf(angle) * 0.48f * ring;
        }
        const float su = sat(startX / 2.10f + 0.50f), sw = sat(startZ / 1.64f + 0.50f);
        const float sv = sid == 2 ? p.emitterHeightNorm : (sid == 1 ? 0.035f + h2 * 0.055f : 0.045f);
        const Tri tc = tri_scalar_setup<float>(p, su, sv, sw);
        const float heat = tri_scalar<HEAT, B>(S, tc), charMass = tri_scalar<CHARF, B>(S, tc);
        const float pyro = tri_scalar<PYRO, B>(S, tc), turb = tri_scalar<TURB, B>(S, tc);
        const F3 vel = sample_vel<float>(U, V, W, p, su, sv, sw);
        const float materialGate = sstep(0.018f, 0.40f, charMass + pyro * 0.34f + heat * 0.075f);
        if (materialGate > 0.0001f) {
            const float tempK = 620.0f + heat * 330.0f + pyro * 210.0f + charMass * 42.0f;
            const float lift = sid == 1 ? 0.86f : (sid == 2 ? 0.16f : 0.80f);
            const float t = life * (sid == 1 ? 1.24f : (sid == 2 ? 0.34f : 1.10f));
            const float dragK = sid == 2 ? 4.8f : 1.35f;
            const float rise = 0.26f + h3 * 0.62f + turb * 0.11f;
            const float fall = 0.22f + h2 * 0.16f;
            const float driftX = vel.x * 0.34f + (h1 - 0.5f) * 0.20f + p.windX * 0.26f;
            const float driftZ = vel.z * 0.28f + (h2 - 0.5f) * 0.18f;
            const float drag = expf(-t * dragK);
            const F3 world = f3(startX + driftX * t * drag, sv * 2.03f + 0.02f + rise * t * lift - fall * t * t, startZ + driftZ * t * drag);
            const float pt = fmaxf(0.0f, t - 0.045f);
            const float pdrag = expf(-pt * dragK);
            const F3 prevWorld = f3(startX + driftX * pt * pdrag, sv * 2.03f + 0.02f + rise * pt * lift - fall * pt * pt, startZ + driftZ * pt * pdrag);
            const Cam cam = cam_of(p);
            const F3 sc = project_point(cam, world), ps = project_point(cam, prevWorld);
            if (!(sc.z <= 0.0f || sc.x < -0.05f || sc.x > 1.05f || sc.y < -0.05f || sc.y > 1.05f)) {
                const float size = sstep(720.0f, 1580.0f, tempK) * materialGate;
                const float radius = ((sid == 2 ? 0.00042f : 0.00086f) + h3 * (sid == 2 ? 0.00042f : 0.00145f)) * (0.72f + size * 0.90f) / fmaxf(0.55f, sc.z);
                const float cxp = sc.x * static_cast<float>(p.fw) - 0.5f, cyp = sc.y * static_cast<float>(p.fh) - 0.5f;
                const float trx = sc.x - ps.x, try_ = sc.y - ps.y;
                const float trailLen = sqrtf(trx * trx + try_ * try_);
                const float capped = fminf(trailLen, sid == 1 ? 0.016f : 0.005f);
                const bool hasDir = trailLen > 0.000001f;
                const int pad = max(2, static_cast<int>(ceilf((radius + capped) * static_cast<float>(max(p.fw, p.fh)) * 3.0f)));
                const F3 hot = sid == 2 ? f3(0.36f, 0.58f, 1.04f) : f3(1.0f, 0.42f, 0.090f);
                const F3 cool = sid == 2 ? f3(0.055f, 0.10f, 0.18f) : f3(0.18f, 0.026f, 0.008f);
                const float cooling = sstep(0.16f, 0.92f, life) * (1.0f - sat(heat * 0.045f));
                const F3 color = lerp3(hot, cool, cooling);
                const float fade = sstep(1.0f, 0.08f, life) * sstep(0.0f, 0.16f, life) * sparkGate * materialGate * size *
                                   (sid == 2 ? 0.055f : (sid == 1 ? 0.58f : 0.82f));
                e.sx = sc.x;
                e.sy = sc.y;
                e.radius = radius;
                e.streak = radius * (1.0f + capped * 130.0f);
                e.tdx = hasDir ? trx * (1.0f / trailLen) : 0.0f;
                e.tdy = hasDir ? try_ * (1.0f / trailLen) : -1.0f;
                e.fade = fade;
                e.r = color.x;
                e.g = color.y;
                e.b = color.z;
                e.x0 = max(0, static_cast<int>(floorf(cxp)) - pad);
                e.x1 = min(p.fw - 1, static_cast<int>(ceilf(cxp)) + pad);
                e.y0 = max(0, static_cast<int>(floorf(cyp)) - pad);
                e.y1 = min(p.fh - 1, static_cast<int>(ceilf(cyp)) + pad);
                e.active = 1;
            }
        }
    }
    out[i] = e;
}

__global__ __launch_bounds__(256) void k_ember_splat(float4* __restrict__ hdr, const Ember* __restrict__ E, DevParams p) {
    __shared__ Ember se[kMaxEmbers];
    const int tid = threadIdx.x + threadIdx.y * blockDim.x;
    for (int i = tid; i < kMaxEmbers; i += blockDim.x * blockDim.y) se[i] = E[i];
    __syncthreads();
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= p.fw || y >= p.fh) return;
    const int count = min(max(p.emberCount, 0), kMaxEmbers);
    const float ux = (static_cast<float>(x) + 0.5f) / static_cast<float>(p.fw);
    const float uy = (static_cast<float>(y) + 0.5f) / static_cast<float>(p.fh);
    float ar = 0.0f, ag = 0.0f, ab = 0.0f;
    bool any = false;
    for (int i = 0; i < count; ++i) {
        const Ember& e = se[i];
        if (!e.active || x < e.x0 || x > e.x1 || y < e.y0 || y > e.y1) continue;
        const float dx = ux - e.sx, dy = uy - e.sy;
        const float along = dx * e.tdx + dy * e.tdy;
        const float ax = dx - e.tdx * along, ay = dy - e.tdy * along;
        const float across2 = ax * ax + ay * ay;
        const float core = expf(-(across2 / fmaxf(0.0000002f, e.radius * e.radius) + along * along / fmaxf(0.0000002f, e.streak * e.streak)));
        const float halo = expf(-(dx * dx + dy * dy) / fmaxf(0.0000002f, e.radius * e.radius * 8.0f)) * 0.16f;
        const float spark = (core + halo) * e.fade;
        if (spark > 0.00001f) {
            ar += e.r * spark;
            ag += e.g * spark;
            ab += e.b * spark;
            any = true;
        }
    }
    if (any) {
        float4 h = hdr[static_cast<size_t>(y) * p.fw + x];
        h.x += ar;
        h.y += ag;
        h.z += ab;
        hdr[static_cast<size_t>(y) * p.fw + x] = h;
    }
}

// ---------------------------------------------------------- FP16 pack ----
// packFp16SurfaceKernel src/fire_cuda.cu:2299-2321: RGB clamped below at 0 and
// rounded to nearest-even half, alpha 1.0; one 8-byte store per pixel.
__device__ __forceinline__ uint32_t half_bits(float v) {
    return static_cast<uint32_t>(__builtin_bit_cast(uint16_t, static_cast<_Float16>(fmaxf(0.0f, v))));
}
__global__ __launch_bounds__(256) void k_pack_fp16(uint2* __restrict__ out, const float4* __restrict__ hdr, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 c = hdr[i];
    out[i] = make_uint2(half_bits(c.x) | (half_bits(c.y) << 16), half_bits(c.z) | (0x3C00u << 16));
}

// ========================================================== launchers ====
template <bool B>
static Scalars<B> scal(void* const* f) {
    Scalars<B> s;
    for (int i = 0; i < NSCALAR; ++i) s.f[i] = f[i];
    return s;
}

// One wavefront per 8x8 light-cell tile: the quarter-resolution volume has few
// cells (32^3 at 128^3), so smaller workgroups spread it over every CU.
dim3 light_block() { return dim3(8, 8, 1); }
dim3 light_grid(int lnx, int lny, int lnz) { return dim3((lnx + 7) / 8, (lny + 7) / 8, lnz); }

hipError_t launch_light(DeviceScene& s, const DevParams& p, hipStream_t st) {
    const dim3 b = light_block();
    const dim3 g = light_grid(s.lnx, s.lny, s.lnz);
    if (s.bf16) k_light_build<true><<<g, b, 0, st>>>(s.light, s.shadow, scal<true>(s.s), p);
    else k_light_build<false><<<g, b, 0, st>>>(s.light, s.shadow, scal<false>(s.s), p);
    k_light_propagate<<<g, b, 0, st>>>(s.lightNext, s.light, s.shadow, p, 0.48f, 0.92f);
    k_light_propagate<<<g, b, 0, st>>>(s.light, s.lightNext, s.shadow, p, 0.34f, 0.96f);
    if (p.dirLight) k_dir_trans<<<g, b, 0, st>>>(s.dirT, s.light, p);
    return hipGetLastError();
}

template <bool B>
__global__ __launch_bounds__(256) void kb_light_build(const BatchTable* __restrict__ T) {
    batch_blocks(T->lightStart, T->count, [&](int s, uint32_t local) {
        const BatchScene& S = T->scene[s];
        const uint3 b = block_of(local, S.lightG);
        Scalars<B> sc;
#pragma unroll
        for (int i = 0; i < NSCALAR; ++i) sc.f[i] = S.sOut[i];
        light_build_at<B>(S.light, S.shadow, sc, S.p, b.x * blockDim.x + threadIdx.x, b.y * blockDim.y + threadIdx.y, b.z * blockDim.z + threadIdx.z);
    });
}

// pass 0: light -> lightNext with (0.48, 0.92); pass 1: back with (0.34, 0.96).
__global__ __launch_bounds__(256) void kb_light_propagate(const BatchTable* __restrict__ T, int pass) {
    batch_blocks(T->lightStart, T->count, [&](int s, uint32_t local) {
        const BatchScene& S = T->scene[s];
        const uint3 b = block_of(local, S.lightG);
        const int lx = b.x * blockDim.x + threadIdx.x, ly = b.y * blockDim.y + threadIdx.y, lz = b.z * blockDim.z + threadIdx.z;
        if (pass == 0) light_propagate_at(S.lightNext, S.light, S.shadow, S.p, 0.48f, 0.92f, lx, ly, lz);
        else light_propagate_at(S.light, S.lightNext, S.shadow, S.p, 0.34f, 0.96f, lx, ly, lz);
    });
}

// Directional-light transmittance of every member that has the light on.
__global__ __launch_bounds__(256) void kb_dir_trans(const BatchTable* __restrict__ T) {
    batch_blocks(T->lightStart, T->count, [&](int s, uint32_t local) {
        const BatchScene& S = T->scene[s];
        if (!S.p.dirLight) return;
        const uint3 b = block_of(local, S.lightG);
        dir_trans_at(S.dirT, S.light, S.p, b.x * blockDim.x + threadIdx.x, b.y * blockDim.y + threadIdx.y, b.z * blockDim.z + threadIdx.z);
    });
}

hipError_t launch_batch_light(const BatchTable* table, bool bf16, int grid, hipStream_t st) {
    const dim3 b = light_block();
    const int g = grid / 4 < 16 ? 16 : grid / 4 - (grid / 4) % 8;  // the light grid is 1/64 of the cells
    if (bf16) kb_light_build<true><<<g, b, 0, st>>>(table);
    else kb_light_build<false><<<g, b, 0, st>>>(table);
    kb_light_propagate<<<g, b, 0, st>>>(table, 0);
    kb_light_propagate<<<g, b, 0, st>>>(table, 1);
    kb_dir_trans<<<g, b, 0, st>>>(table);
    return hipGetLastError();
}

hipError_t launch_render(DeviceScene& s, const DevParams& p, hipStream_t st) {
    const dim3 b(16, 16, 1);
    const dim3 g((s.fw + 15) / 16, (s.fh + 15) / 16, 1);
    if (s.bf16) k_render<true><<<g, b, 0, st>>>(s.hdr, scal<true>(s.s), s.light, s.shadow, s.dirT, s.u, s.v, s.w, p);
    else k_render<false><<<g, b, 0, st>>>(s.hdr, scal<false>(s.s), s.light, s.shadow, s.dirT, s.u, s.v, s.w, p);
    if (p.emberCount > 0 && p.debugMode == 0) {
        if (s.bf16) k_ember_prepare<true><<<1, kMaxEmbers, 0, st>>>(s.embers, scal<true>(s.s), s.u, s.v, s.w, p);
        else k_ember_prepare<false><<<1, kMaxEmbers, 0, st>>>(s.embers, scal<false>(s.s), s.u, s.v, s.w, p);
        k_ember_splat<<<g, b, 0, st>>>(s.hdr, s.embers, p);
    }
    return hipGetLastError();
}

hipError_t launch_pack_fp16(DeviceScene& s, const DevParams& p, hipStream_t st) {
    const int n = s.fw * s.fh;
    k_pack_fp16<<<(n + 255) / 256, 256, 0, st>>>(reinterpret_cast<uint2*>(s.rgba16), s.hdr, n);
    return hipGetLastError();
}

hipError_t launch_tonemap(DeviceScene& s, const DevParams& p, hipStream_t st) {
    const dim3 b(16, 16, 1);
    const dim3 g((s.fw + 15) / 16, (s.fh + 15) / 16, 1);
    k_tonemap<<<g, b, 0, st>>>(s.bgra, s.hdr, p);
    return hipGetLastError();
}

}  // namespace fs
