// step_kernels.hip - simulation passes of one step on gfx950.
//
// Passes (reference kernels in src/fire_cuda.cu):
//   k_advect_velocity   K5-K7  :880-932   one launch for u, v and w faces
//   k_advect_react      K8     :934-1126  MacCormack x8 + combustion, FP64 heat-release partials
//   k_force_contrib     K9     :1128-1183 per-cell force contributions (gather form)
//   k_force_apply       K9-K11 :1128-1199,:1261-1291 faces += contributions, clamp, divergence
//   k_sor_half          K12    :1201-1227 one red/black half-sweep
//   k_project           K13-14 :1229-1245,:1293-1317 gradient subtract + wall zeroing
//   k_reset             K1-K4  :805-878
//   k_metrics           K15-16 :1247-1259,:1319-1418 FP64 block partials + k_reduce_metrics
#include "fs_launch.hpp"

namespace fs {

dim3 cell_block() { return dim3(64, 4, 1); }
dim3 cell_grid(int nx, int ny, int nz) {
    const dim3 b = cell_block();
    return dim3((nx + b.x - 1) / b.x, (ny + b.y - 1) / b.y, (nz + b.z - 1) / b.z);
}
size_t react_partial_count(int nx, int ny, int nz) {
    const dim3 g = cell_grid(nx, ny, nz);
    return static_cast<size_t>(g.x) * g.y * g.z;
}

__device__ __forceinline__ Idx sidx(const DevParams& p, int x, int y, int z) {
    return (static_cast<Idx>(z) * static_cast<Idx>(p.ny) + static_cast<Idx>(y)) * static_cast<Idx>(p.nx) + static_cast<Idx>(x);
}
__device__ __forceinline__ Idx uidx(const DevParams& p, int x, int y, int z) {
    return (static_cast<Idx>(z) * static_cast<Idx>(p.ny) + static_cast<Idx>(y)) * static_cast<Idx>(p.nx + 1) + static_cast<Idx>(x);
}
__device__ __forceinline__ Idx vidx(const DevParams& p, int x, int y, int z) {
    return (static_cast<Idx>(z) * static_cast<Idx>(p.ny + 1) + static_cast<Idx>(y)) * static_cast<Idx>(p.nx) + static_cast<Idx>(x);
}
__device__ __forceinline__ Idx widx(const DevParams& p, int x, int y, int z) {
    return (static_cast<Idx>(z) * static_cast<Idx>(p.ny) + static_cast<Idx>(y)) * static_cast<Idx>(p.nx) + static_cast<Idx>(x);
}

// Sum over the 256-thread block in FP64; thread 0 returns the total.
__device__ __forceinline__ double block_sum_256(double v) {
    __shared__ double lds[4];
    const int lane = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    v = wave_sum(v);
    if ((lane & 63) == 0) lds[lane >> 6] = v;
    __syncthreads();
    double t = 0.0;
    if (lane == 0) t = (lds[0] + lds[1]) + (lds[2] + lds[3]);
    return t;
}

// ------------------------------------------------------------ K5-K7 ----
// Sample coordinates in index units, I + r: I an integer cell index and r a
// small FP32 remainder (the position's offset within the grid spacing plus the
// backtrace displacement in cells). The reference forms a normalized
// coordinate u in [0, 1] and scales it by s (n on a face-aligned axis, n - 1 on
// a cell-aligned one): at x / n that is x on a face axis and x - x/n on a cell
// axis, at (x + 0.5) / n it is x + 0.5 and x + 0.5 - (x + 0.5)/n. Splitting off
// the integer keeps the remainder's rounding at FP32 ulp of a number below a
// few cells (~1e-7 of a cell), where a normalized FP32 coordinate scaled by up
// to 1024 cells would carry ~6e-5 of a cell; so the backtrace needs no FP64.
// Clamped to [0, s] as the reference's saturate of u.
__device__ __forceinline__ Axis axis_at(int d, int s, int I, float r) {
    const float fl = floorf(r);
    int i0 = I + static_cast<int>(fl);
    float t = r - fl;
    if (i0 < 0) {
        i0 = 0;
        t = 0.0f;
    } else if (i0 >= s) {
        i0 = s;
        t = 0.0f;
    }
    return Axis{i0, min(i0 + 1, d - 1), t};
}

// Bilinear sample of a plane-aligned position: the component's own axis lies
// on a grid node (t = 0 there), so only four corners are read; lerp(a, b, 0)
// is a, so the value equals the trilinear one.
__device__ __forceinline__ float bilerp(const float* __restrict__ f, Idx i00, Idx o1, Idx o2, float t1, float t2) {
    const float a = lerpf(ld_at(f, i00), ld_at(f, i00 + o1), t1);
    const float b = lerpf(ld_at(f, i00 + o2), ld_at(f, i00 + o2 + o1), t1);
    return lerpf(a, b, t2);
}

// Thread (x, y, z) of a face grid of (nx, ny+1, nz+1) writes the u, v and w
// faces at (x, y, z); the thread at x = nx - 1 also writes the wall u face at
// x = nx (0), so no wavefront is spent on that one column.
__device__ __forceinline__ void advect_velocity_at(float* __restrict__ uo, float* __restrict__ vo, float* __restrict__ wo,
                                                   const float* __restrict__ U, const float* __restrict__ V, const float* __restrict__ W,
                                                   const DevParams& p, int x, int y, int z) {
    if (x >= p.nx || y > p.ny || z > p.nz) return;
    if (x == p.nx - 1 && y < p.ny && z < p.nz) st_at(uo, uidx(p, p.nx, y, z), 0.0f);
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const float k = p.dt * kAdvectionScale;
    // remainders of the face (a) and centre (c) coordinates on the
    // cell-aligned axes; on face-aligned axes they are 0 and 0.5
    const float axr = -static_cast<float>(x) * p.invN[0], cxr = 0.5f - (static_cast<float>(x) + 0.5f) * p.invN[0];
    const float ayr = -static_cast<float>(y) * p.invN[1], cyr = 0.5f - (static_cast<float>(y) + 0.5f) * p.invN[1];
    const float azr = -static_cast<float>(z) * p.invN[2], czr = 0.5f - (static_cast<float>(z) + 0.5f) * p.invN[2];
    const Idx rowU = static_cast<Idx>(nx + 1), planeU = rowU * static_cast<Idx>(ny);
    const Idx rowV = static_cast<Idx>(nx), planeV = rowV * static_cast<Idx>(ny + 1);
    const Idx rowW = static_cast<Idx>(nx), planeW = rowW * static_cast<Idx>(ny);
    // Axis setups shared by the nine velocity samples: per coordinate, the
    // face-aligned (F) and cell-aligned (C) array axes.
    const Axis axC = axis_at(nx, nx - 1, x, axr), cxF = axis_at(nx + 1, nx, x, 0.5f), cxC = axis_at(nx, nx - 1, x, cxr);
    const Axis ayC = axis_at(ny, ny - 1, y, ayr), cyF = axis_at(ny + 1, ny, y, 0.5f), cyC = axis_at(ny, ny - 1, y, cyr);
    const Axis azC = axis_at(nz, nz - 1, z, azr), czF = axis_at(nz + 1, nz, z, 0.5f), czC = axis_at(nz, nz - 1, z, czr);
    if (y < ny && z < nz) {
        float out = 0.0f;
        if (x != 0 && x != nx) {
            // velocity at the u face (x, y + 1/2, z + 1/2): u on its x node
            const float vu = bilerp(U, (static_cast<Idx>(czC.i0) * static_cast<Idx>(ny) + cyC.i0) * rowU + x,
                                    cyC.i1 != cyC.i0 ? rowU : 0u, czC.i1 != czC.i0 ? planeU : 0u, cyC.t, czC.t);
            const float vv = tri_plain(V, tri_of(axC, cyF, czC, nx, ny + 1));
            const float vw = tri_plain(W, tri_of(axC, cyC, czF, nx, ny));
            const Tri tb = tri_of(axis_at(nx + 1, nx, x, -vu * k * static_cast<float>(nx)),
                                  axis_at(ny, ny - 1, y, cyr - vv * k * static_cast<float>(ny - 1)),
                                  axis_at(nz, nz - 1, z, czr - vw * k * static_cast<float>(nz - 1)), nx + 1, ny);
            out = tri_plain(U, tb) * 0.995f;
        }
        st_at(uo, uidx(p, x, y, z), out);
    }
    if (x < nx && z < nz) {
        float out = 0.0f;
        if (y != 0 && !(p.closedTop && y == ny)) {
            // velocity at the v face (x + 1/2, y, z + 1/2): v on its y node
            const float vu = tri_plain(U, tri_of(cxF, ayC, czC, nx + 1, ny));
            const float vv = bilerp(V, (static_cast<Idx>(czC.i0) * static_cast<Idx>(ny + 1) + y) * rowV + cxC.i0, static_cast<Idx>(cxC.i1 - cxC.i0),
                                    czC.i1 != czC.i0 ? planeV : 0u, cxC.t, czC.t);
            const float vw = tri_plain(W, tri_of(cxC, ayC, czF, nx, ny));
            const Tri tb = tri_of(axis_at(nx, nx - 1, x, cxr - vu * k * static_cast<float>(nx - 1)),
                                  axis_at(ny + 1, ny, y, -vv * k * static_cast<float>(ny)),
                                  axis_at(nz, nz - 1, z, czr - vw * k * static_cast<float>(nz - 1)), nx, ny + 1);
            out = tri_plain(V, tb) * 0.996f;
        }
        st_at(vo, vidx(p, x, y, z), out);
    }
    if (x < nx && y < ny) {
        float out = 0.0f;
        if (z != 0 && z != nz) {
            // velocity at the w face (x + 1/2, y + 1/2, z): w on its z node
            const float vu = tri_plain(U, tri_of(cxF, cyC, azC, nx + 1, ny));
            const float vv = tri_plain(V, tri_of(cxC, cyF, azC, nx, ny + 1));
            const float vw = bilerp(W, (static_cast<Idx>(z) * static_cast<Idx>(ny) + cyC.i0) * rowW + cxC.i0, static_cast<Idx>(cxC.i1 - cxC.i0),
                                    cyC.i1 != cyC.i0 ? rowW : 0u, cxC.t, cyC.t);
            const Tri tb = tri_of(axis_at(nx, nx - 1, x, cxr - vu * k * static_cast<float>(nx - 1)),
                                  axis_at(ny, ny - 1, y, cyr - vv * k * static_cast<float>(ny - 1)),
                                  axis_at(nz + 1, nz, z, -vw * k * static_cast<float>(nz)), nx, ny);
            out = tri_plain(W, tb) * 0.995f;
        }
        st_at(wo, widx(p, x, y, z), out);
    }
    (void)planeW;
}

__device__ __forceinline__ int3 cell_of(uint3 blk) {
    return make_int3(static_cast<int>(blk.x * blockDim.x + threadIdx.x), static_cast<int>(blk.y * blockDim.y + threadIdx.y),
                     static_cast<int>(blk.z * blockDim.z + threadIdx.z));
}

// With a non-null `clock`, one thread also advances the device clock by one
// step (as k_advance_clock does): this pass does not read the clock, and every
// later pass of the step sees the advanced value.
__global__ __launch_bounds__(256) void k_advect_velocity(float* __restrict__ uo, float* __restrict__ vo, float* __restrict__ wo,
                                                         const float* __restrict__ U, const float* __restrict__ V,
                                                         const float* __restrict__ W, DevParams p, float* __restrict__ clock) {
    if (clock != nullptr && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
        clock[0] += p.dt;
        clock[1] += p.dt;
    }
    const int3 c = cell_of(xcd_block());
    advect_velocity_at(uo, vo, wo, U, V, W, p, c.x, c.y, c.z);
}

// --------------------------------------------------------------- K8 ----
// MacCormack step sharing the backtrace and forward-trace cells of all fields
// (sampleScalarMacCormack src/fire_cuda.cu:713-732).
template <int K, bool B>
__device__ __forceinline__ float maccormack(const Scalars<B>& S, const Tri& tp, const Tri& tf, float current, float blend) {
    const float c000 = S.template ld<K>(tp.i000), c100 = S.template ld<K>(tp.i100);
    const float c010 = S.template ld<K>(tp.i010), c110 = S.template ld<K>(tp.i110);
    const float c001 = S.template ld<K>(tp.i001), c101 = S.template ld<K>(tp.i101);
    const float c011 = S.template ld<K>(tp.i011), c111 = S.template ld<K>(tp.i111);
    const float first = tri_blend(tp, c000, c100, c010, c110, c001, c101, c011, c111);
    const float lo = fminf(fminf(fminf(c000, c100), fminf(c010, c110)), fminf(fminf(c001, c101), fminf(c011, c111)));
    const float hi = fmaxf(fmaxf(fmaxf(c000, c100), fmaxf(c010, c110)), fmaxf(fmaxf(c001, c101), fmaxf(c011, c111)));
    const float round = tri_scalar<K, B>(S, tf);
    const float corrected = fminf(hi, fmaxf(lo, first + 0.5f * (current - round)));
    return lerpf(first, corrected, sat(blend));
}

template <bool B>
__device__ float react_cell(const Scalars<B>& out, const Scalars<B>& in, const float* __restrict__ U, const float* __restrict__ V,
                            const float* __restrict__ W, const DevParams& p, int x, int y, int z) {
    const Idx idx = sidx(p, x, y, z);
    const float u = (static_cast<float>(x) + 0.5f) * p.invN[0];
    const float v = (static_cast<float>(y) + 0.5f) * p.invN[1];
    const float w = (static_cast<float>(z) + 0.5f) * p.invN[2];
    const float dt = p.dt;
    const float time = ld_at(p.clock, 0u);
    const int sid = p.sceneId;
    const float h0 = in.template ld<HEAT>(idx), f0 = in.template ld<FUEL>(idx), o0 = in.template ld<OXYGEN>(idx);
    const float s0 = in.template ld<SOOT>(idx), ch0 = in.template ld<CHARF>(idx), a0 = in.template ld<ASH>(idx);
    const float py0 = in.template ld<PYRO>(idx), pr0 = in.template ld<PROGRESS>(idx), k0 = in.template ld<TURB>(idx);
    const float so0 = in.template ld<SOPT>(idx);
    const float activity = h0 + f0 * 0.18f + s0 * 0.24f + py0 * 0.20f + pr0 * 0.14f + so0 * 0.08f + k0 * 0.06f;
    const float sdx = u - p.mouseX, sdy = v - p.mouseY, sdz = w - 0.5f;
    const bool outsideInteractive = p.leftDown == 0 || (sdx * sdx + sdy * sdy + sdz * sdz) > 0.090f;
    uint32_t inside = 0u;
    const float cx = static_cast<float>(x) + 0.5f, cy = static_cast<float>(y) + 0.5f, cz = static_cast<float>(z) + 0.5f;
    for (int e = 0; e < p.emitterCount; ++e) {
        const float ddx = cx - p.emitter[e][0], ddy = cy - p.emitter[e][1], ddz = cz - p.emitter[e][2];
        if (ddx * ddx + ddy * ddy + ddz * ddz <= p.emitter[e][3] * p.emitter[e][3]) inside |= 1u << e;
    }
    if (v > 0.18f && activity < 0.00035f && ch0 < 0.00010f && a0 < 0.00010f && outsideInteractive && inside == 0u && p.rightDown == 0) {
        out.template st<HEAT>(idx, 0.0f);
        out.template st<FUEL>(idx, 0.0f);
        out.template st<OXYGEN>(idx, fminf(1.0f, o0 + (1.0f - o0) * dt * 0.08f));
        out.template st<SOOT>(idx, 0.0f);
        out.template st<CHARF>(idx, ch0);
        out.template st<ASH>(idx, a0);
        out.template st<PYRO>(idx, 0.0f);
        out.template st<PROGRESS>(idx, 0.0f);
        out.template st<TURB>(idx, k0 * fmaxf(0.0f, 1.0f - dt * 0.54f));
        out.template st<SOPT>(idx, 0.0f);
        return 0.0f;
    }
    // Backtrace and forward trace in index units, as advect_velocity_at: the
    // normalized centre (x + 1/2)/n maps to x + 1/2 on a face-aligned axis
    // (scale n) and to x + r on a cell-aligned one (scale n - 1, r = 1/2 -
    // (x + 1/2)/n); a displacement d in normalized units adds -d * scale to
    // the remainder. The forward trace starts from the unclamped backtrace.
    const int nx = p.nx, ny = p.ny, nz = p.nz;
    const float k = dt * kAdvectionScale;
    const float cxr = 0.5f - u, cyr = 0.5f - v, czr = 0.5f - w;
    const float sxF = static_cast<float>(nx), syF = static_cast<float>(ny), szF = static_cast<float>(nz);
    const float sxC = static_cast<float>(nx - 1), syC = static_cast<float>(ny - 1), szC = static_cast<float>(nz - 1);
    const Axis cxC = axis_at(nx, nx - 1, x, cxr), cyC = axis_at(ny, ny - 1, y, cyr), czC = axis_at(nz, nz - 1, z, czr);
    const float vx = tri_plain(U, tri_of(axis_at(nx + 1, nx, x, 0.5f), cyC, czC, nx + 1, ny));
    const float vy = tri_plain(V, tri_of(cxC, axis_at(ny + 1, ny, y, 0.5f), czC, nx, ny + 1));
    const float vz = tri_plain(W, tri_of(cxC, cyC, axis_at(nz + 1, nz, z, 0.5f), nx, ny));
    const float bxC = cxr - vx * k * sxC, byC = cyr - vy * k * syC, bzC = czr - vz * k * szC;
    const Axis pxC = axis_at(nx, nx - 1, x, bxC), pyC = axis_at(ny, ny - 1, y, byC), pzC = axis_at(nz, nz - 1, z, bzC);
    const float px = tri_plain(U, tri_of(axis_at(nx + 1, nx, x, 0.5f - vx * k * sxF), pyC, pzC, nx + 1, ny));
    const float py = tri_plain(V, tri_of(pxC, axis_at(ny + 1, ny, y, 0.5f - vy * k * syF), pzC, nx, ny + 1));
    const float pz = tri_plain(W, tri_of(pxC, pyC, axis_at(nz + 1, nz, z, 0.5f - vz * k * szF), nx, ny));
    const Tri tp = tri_of(pxC, pyC, pzC, nx, ny);
    const Tri tf = tri_of(axis_at(nx, nx - 1, x, bxC + px * k * sxC), axis_at(ny, ny - 1, y, byC + py * k * syC),
                          axis_at(nz, nz - 1, z, bzC + pz * k * szC), nx, ny);
    const float anti = sat(0.30f + p.detail * 0.34f);
    float heat = maccormack<HEAT, B>(in, tp, tf, h0, anti);
    float fuel = maccormack<FUEL, B>(in, tp, tf, f0, anti * 0.92f);
    float oxygen = maccormack<OXYGEN, B>(in, tp, tf, o0, anti * 0.70f);
    float soot = maccormack<SOOT, B>(in, tp, tf, s0, anti);
    float charMass = ch0;
    float ash = a0;
    float pyro = maccormack<PYRO, B>(in, tp, tf, py0, anti * 0.72f) * 0.55f;
    float progress = maccormack<PROGRESS, B>(in, tp, tf, pr0, anti * 0.82f);
    float turb = maccormack<TURB, B>(in, tp, tf, k0, anti * 0.52f);
    float sootOptics = maccormack<SOPT, B>(in, tp, tf, so0, anti * 0.72f);

    const float worldX = (u - 0.50f) * 2.10f;
    const float worldZ = (w - 0.50f) * 1.64f;
    const float material = bed_material(worldX, worldZ, p);
    const float exposedChar = sat(charMass / fmaxf(0.055f, charMass + ash * 1.75f));
    const float charEdges = sstep(0.06f, 0.78f, material) * sstep(0.92f, 0.02f, ash);
    const float gasFeed = sid == 2 ? 1.18f : exposedChar * charEdges;
    const float bedSrc = bed_source(material, worldX, worldZ, v, p, time);
    // The source noise only matters when a bed is exposed or char can release gas.
    float sourceNoise = 0.70f;
    if (bedSrc * gasFeed != 0.0f || charMass != 0.0f) {
        sourceNoise = 0.70f + 0.48f * fbm2(worldX * 13.0f + worldZ * 3.0f + time * 0.38f, v * 31.0f + worldZ * 11.0f, p.saltNoise2);
    }
    const float pyroGain = sid == 2 ? 0.42f : (sid == 1 ? 1.34f : 1.0f);
    const float fuelGain = sid == 2 ? 0.46f : (sid == 1 ? 1.18f : 1.0f);
    const float heatGain = sid == 2 ? 1.18f : (sid == 1 ? 1.08f : 1.0f);
    const float sootGain = sid == 2 ? 0.18f : (sid == 1 ? 1.22f : 1.0f);
    const float bed = bedSrc * sourceNoise * gasFeed * p.intensity;
    const float ambientK = 293.0f;
    float tempK = ambientK + heat * 335.0f;
    const float radFb = sstep(430.0f, 980.0f, tempK);
    const float solidO2 = sat(oxygen * 1.28f - soot * 0.030f);
    const float heatFlux = sat((tempK - 405.0f) * (1.0f / 900.0f)) + bed * 0.34f + sootOptics * 0.024f;
    const float charRelease = fminf(charMass, charMass * heatFlux * solidO2 * (0.14f + radFb * 1.38f) * dt);
    pyro += (bed * (0.34f + radFb * 1.05f) * dt + charRelease * (2.45f + sourceNoise * 0.72f)) * pyroGain;
    charMass += material * dt * (sid == 2 ? 0.00005f : (sid == 1 ? 0.050f : 0.026f));
    charMass -= charRelease * 0.92f;
    ash += charRelease * (sid == 2 ? 0.025f : 0.16f + 0.22f * sat(1.0f - oxygen));
    heat += (bed * (sid == 2 ? 10.40f : 6.20f) * dt * (1.0f + radFb * 1.12f) + charRelease * 3.35f) * heatGain;
    fuel += pyro * (sid == 2 ? 1.34f : 3.70f) * fuelGain;
    soot += pyro * (0.090f + ash * 0.032f + sat(1.0f - oxygen) * 0.13f) * p.smokeGain * (sid == 2 ? 0.010f : (sid == 1 ? 0.92f : 1.0f)) * sootGain;

    if (p.leftDown != 0) {
        const float mx = p.mouseX, my = p.mouseY;
        const float source = expf(-((u - mx) * (u - mx)) * (1.0f / 0.0180f) - ((w - 0.5f) * (w - 0.5f)) * (1.0f / 0.0550f) - ((v - my) * (v - my)) * (1.0f / 0.0020f));
        heat += source * dt * 2.8f;
        fuel += source * dt * 1.4f;
        pyro += source * dt * 0.45f;
        progress += source * dt * 0.75f;
        turb += source * dt * 0.28f;
        oxygen = fmaxf(oxygen, 0.72f + source * 0.20f);
    }
    if (p.rightDown != 0) {
        const float source = expf(-((u - p.mouseX) * (u - p.mouseX) + (w - 0.5f) * (w - 0.5f)) * (1.0f / 0.0100f) - (v - p.mouseY) * (v - p.mouseY) * (1.0f / 0.0100f));
        soot += source * dt * 8.0f;
        heat *= 1.0f - source * dt * 1.2f;
        fuel *= 1.0f - source * dt * 0.7f;
    }
    for (int e = 0; e < p.emitterCount; ++e) {
        if (inside & (1u << e)) {
            fuel += p.emitter[e][4] * dt;
            heat = fmaxf(heat, p.emitter[e][5]);
        }
    }

    tempK = ambientK + heat * 335.0f;
    const float thermal = sstep(575.0f, 1220.0f, tempK);
    const float arrhenius = expf(-3800.0f / fmaxf(640.0f, tempK));
    const float oxyLim = fminf(fuel, oxygen * 1.42f);
    const float burnMass = fminf(oxyLim, oxyLim * thermal * arrhenius * (118.0f + p.intensity * 34.0f) * dt);
    const float incomplete = sat(1.0f - oxygen * 1.30f + fuel * 0.065f + soot * 0.015f);
    const float release = burnMass * 7.0f;
    const float sootYield = burnMass * (0.035f + incomplete * 0.34f) * p.smokeGain * (sid == 2 ? 0.006f : (sid == 1 ? 0.96f : 1.0f));
    const float sootOxidation = soot * oxygen * sstep(760.0f, 1540.0f, tempK) * dt * 0.82f;
    const float front = burnMass * (9.8f + turb * 3.4f) + pyro * 0.70f + charRelease * 2.10f;
    const float bA = sat(sstep(0.70f, 0.96f, v) + sstep(0.11f, 0.03f, u) + sstep(0.89f, 0.97f, u) + sstep(0.11f, 0.03f, w) + sstep(0.89f, 0.97f, w));
    const float upper = sstep(0.52f, 0.98f, v);

    heat += release;
    fuel -= burnMass * 1.20f;
    oxygen -= burnMass * 0.98f;
    soot += sootYield;
    soot -= sootOxidation;
    progress += front;
    progress *= expf(-dt * (0.72f + v * 0.45f + oxygen * 0.16f + bA * 1.05f + upper * 0.26f));
    oxygen = lerpf(oxygen, 1.0f, bA * dt * 1.7f);
    oxygen += (1.0f - oxygen) * dt * 0.08f;

    tempK = ambientK + heat * 335.0f;
    const float radBase = fmaxf(0.0f, (tempK - ambientK) * (1.0f / 1560.0f));
    const float radSq = radBase * radBase;
    heat -= radSq * radSq * dt * (0.88f + soot * 0.085f + sootOptics * 0.018f);
    heat *= expf(-dt * (0.31f + v * 0.64f + soot * 0.022f + bA * 0.26f + upper * 0.18f));
    fuel *= expf(-dt * (0.58f + heat * 0.20f + v * 0.92f + bA * 0.18f));
    soot *= expf(-dt * (0.060f + v * 0.092f + oxygen * 0.028f + bA * 0.96f + upper * 0.22f));
    pyro *= expf(-dt * (1.55f + v * 0.58f + bA * 0.18f));

    const float uL = ld_at(U, uidx(p, x, y, z)), uR = ld_at(U, uidx(p, x + 1, y, z));
    const float vD = ld_at(V, vidx(p, x, y, z)), vU = ld_at(V, vidx(p, x, y + 1, z));
    const float wB = ld_at(W, widx(p, x, y, z)), wF = ld_at(W, widx(p, x, y, z + 1));
    const float shear = fabsf(uR - uL) + fabsf(vU - vD) + fabsf(wF - wB);
    const float flameActivity = sat(progress * 0.64f + burnMass * 5.2f + heat * 0.065f);
    turb += dt * (shear * (0.72f + progress * 0.06f) + flameActivity * p.turbulence * 0.62f + front * 0.085f + soot * 0.018f);
    turb *= expf(-dt * (0.54f + v * 0.24f + bA * 0.42f + upper * 0.16f));
    const float growth = sat(incomplete * 0.36f + ash * 0.13f + soot * 0.055f + sstep(0.65f, 1.55f, tempK * (1.0f / 1000.0f)) * 0.18f);
    sootOptics = soot * (1.04f + 0.28f * sqrtf(fmaxf(0.0f, heat)) + 0.24f * ash + 0.46f * growth);
    sootOptics = lerpf(sootOptics, sootOptics * 0.48f, sat(sootOxidation * 0.92f));

    out.template st<HEAT>(idx, fminf(8.8f, fmaxf(0.0f, heat)));
    out.template st<FUEL>(idx, fminf(5.5f, fmaxf(0.0f, fuel)));
    out.template st<OXYGEN>(idx, fminf(1.0f, fmaxf(0.0f, oxygen)));
    out.template st<SOOT>(idx, fminf(7.4f, fmaxf(0.0f, soot)));
    out.template st<CHARF>(idx, fminf(2.8f, fmaxf(0.0f, charMass)));
    out.template st<ASH>(idx, fminf(1.9f, fmaxf(0.0f, ash)));
    out.template st<PYRO>(idx, fminf(3.6f, fmaxf(0.0f, pyro)));
    out.template st<PROGRESS>(idx, fminf(3.4f, fmaxf(0.0f, progress)));
    out.template st<TURB>(idx, fminf(4.2f, fmaxf(0.0f, turb)));
    out.template st<SOPT>(idx, fminf(10.0f, fmaxf(0.0f, sootOptics)));
    return release;
}

template <bool B>
__global__ __launch_bounds__(256) void k_advect_react(Scalars<B> out, Scalars<B> in, const float* __restrict__ U, const float* __restrict__ V,
                                                      const float* __restrict__ W, DevParams p, double* __restrict__ partial) {
    const uint3 blk = xcd_block();
    const int3 c = cell_of(blk);
    float release = 0.0f;
    if (c.x < p.nx && c.y < p.ny && c.z < p.nz) release = react_cell<B>(out, in, U, V, W, p, c.x, c.y, c.z);
    const double t = block_sum_256(static_cast<double>(release));
    // one partial per logical tile, so the batched pass fills the same slots
    if (threadIdx.x == 0 && threadIdx.y == 0 && threadIdx.z == 0) partial[(static_cast<size_t>(blk.z) * gridDim.y + blk.y) * gridDim.x + blk.x] = t;
}

// Advances the device clock by one step (FP32 adds, as the host mirror does).
__global__ void k_advance_clock(float* clock, float dt) {
    if (threadIdx.x == 0) {
        clock[0] += dt;
        clock[1] += dt;
    }
}

// Thread t's sum of p[t], p[t + 256], p[t + 512], ... added in that order,
// with eight loads in flight (a loop of one load and one add waits a full
// memory latency per element). t is the thread's linear id in a 256-thread
// workgroup.
__device__ __forceinline__ double strided_sum_256(const double* __restrict__ p, size_t n, unsigned t) {
    constexpr size_t S = 256;
    double s = 0.0;
    size_t i = t;
    for (; i + 7 * S < n; i += 8 * S) {
        double v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = p[i + j * S];
#pragma unroll
        for (int j = 0; j < 8; ++j) s += v[j];
    }
    for (; i < n; i += S) s += p[i];
    return s;
}

// Sums the react partials in a fixed order: acc[0] = step total, acc[1] += it.
__global__ __launch_bounds__(256) void k_finalize_release(const double* __restrict__ partial, size_t n, double* __restrict__ acc) {
    const double t = block_sum_256(strided_sum_256(partial, n, threadIdx.x));
    if (threadIdx.x == 0) {
        acc[0] = t;
        acc[1] += t;
    }
}

// ---------------------------------------------------------- K9-K11 ----
// Divergence-free curl-noise forcing (an extension that replaces the fbm
// noise terms of src/fire_cuda.cu:1164-1165 when curlOn is set). A vector
// potential psi lives on the grid edges of index space: psi_x at
// (i + 1/2, j, k), psi_y at (i, j + 1/2, k), psi_z at (i, j, k + 1/2). Each
// face receives curlNoise * dt times the index-space curl of psi around it
// (u: dpsi_z/dj - dpsi_y/dk, v: dpsi_x/dk - dpsi_z/di, w: dpsi_y/di -
// dpsi_x/dj), whose discrete divergence, uR - uL + vU - vD + wF - wB, cancels
// term by term. The tangential components vanish on the walls (and on the top
// when it is closed), so no flux crosses them. Each component is one octave of
// integer-hash value noise of the edge's normalized position, drifting with
// time, scaled so that its curl is of order 1.
constexpr float kCurlFrequency = 6.0f;  // lattice cells across the domain
// The noise's lattice values at this step, shared by a workgroup. Edge
// positions e in [0, 1] map to e * 6 + drift, so lattice indices run from
// floor(drift) (e = 0 gives the drift exactly) to at most floor(drift) + 7
// (6 cells, plus one when the FP32 sum rounds up to the next integer), and
// the upper corners one further: 9 nodes per axis. Every workgroup hashes the
// 3 x 9^3 nodes once (about ten hashes a thread, against 42 per cell when each
// edge hashes its own eight corners); the values are those vnoise3 would hash.
constexpr int kCurlLat = 9;
struct CurlLattice {
    float v[3][kCurlLat * kCurlLat * kCurlLat];
    uint32_t row[3][kCurlLat * kCurlLat];
};
__device__ __forceinline__ int3 curl_lattice_origin(float t) {
    return make_int3(static_cast<int>(floorf(t * 0.23f)), static_cast<int>(floorf(-(t * 0.17f))), static_cast<int>(floorf(t * 0.19f)));
}
// Called by every thread of the workgroup (it synchronizes).
__device__ void build_curl_lattice(CurlLattice& L, const DevParams& p, float t) {
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nt = blockDim.x * blockDim.y * blockDim.z;
    const int3 o = curl_lattice_origin(t);
    __syncthreads();  // a previous use of L by this workgroup has finished
    for (int i = tid; i < 3 * kCurlLat * kCurlLat; i += nt) {
        const int c = i / (kCurlLat * kCurlLat), r = i % (kCurlLat * kCurlLat);
        const uint32_t z = pcg(p.saltCurl[c].salt + static_cast<uint32_t>(o.z + r / kCurlLat));
        L.row[c][r] = pcg(z + static_cast<uint32_t>(o.y + r % kCurlLat));
    }
    __syncthreads();
    for (int i = tid; i < 3 * kCurlLat * kCurlLat * kCurlLat; i += nt) {
        const int c = i / (kCurlLat * kCurlLat * kCurlLat), r = i % (kCurlLat * kCurlLat * kCurlLat);
        const int zy = r / kCurlLat, ix = o.x + r % kCurlLat;
        L.v[c][r] = p.saltCurl[c].sinHash
                        ? sin_hash31(static_cast<float>(ix), static_cast<float>(o.y + zy % kCurlLat), static_cast<float>(o.z + zy / kCurlLat))
                        : unit24(pcg(L.row[c][zy] + static_cast<uint32_t>(ix)));
    }
    __syncthreads();
}
// vnoise3 of component c from the lattice table (same corners, same blend).
__device__ __forceinline__ float curl_potential(int c, float ex, float ey, float ez, float t, const DevParams& p, const CurlLattice& L) {
    const float scale = static_cast<float>(p.nx + p.ny + p.nz) / (3.0f * kCurlFrequency);
    const float px = ex * kCurlFrequency + t * 0.23f, py = ey * kCurlFrequency - t * 0.17f, pz = ez * kCurlFrequency + t * 0.19f;
    const float fx0 = floorf(px), fy0 = floorf(py), fz0 = floorf(pz);
    const int3 o = curl_lattice_origin(t);
    const int ix = min(max(static_cast<int>(fx0) - o.x, 0), kCurlLat - 2);
    const int iy = min(max(static_cast<int>(fy0) - o.y, 0), kCurlLat - 2);
    const int iz = min(max(static_cast<int>(fz0) - o.z, 0), kCurlLat - 2);
    const float fx = px - fx0, fy = py - fy0, fz = pz - fz0;
    const float ux = fx * fx * (3.0f - 2.0f * fx);
    const float uy = fy * fy * (3.0f - 2.0f * fy);
    const float uz = fz * fz * (3.0f - 2.0f * fz);
    const float* v = L.v[c] + (iz * kCurlLat + iy) * kCurlLat + ix;
    constexpr int DY = kCurlLat, DZ = kCurlLat * kCurlLat;
    const float c00 = lerpf(v[0], v[1], ux);
    const float c10 = lerpf(v[DY], v[DY + 1], ux);
    const float c01 = lerpf(v[DZ], v[DZ + 1], ux);
    const float c11 = lerpf(v[DZ + DY], v[DZ + DY + 1], ux);
    return scale * (lerpf(lerpf(c00, c10, uy), lerpf(c01, c11, uy), uz) - 0.5f);
}
// Arrays of nx * (ny + 1) * nz edges, x fastest; edges at i = nx or k = nz are
// wall edges (0) and are not stored.
__device__ __forceinline__ Idx psi_index(const DevParams& p, int i, int j, int k) {
    return (static_cast<Idx>(k) * static_cast<Idx>(p.ny + 1) + static_cast<Idx>(j)) * static_cast<Idx>(p.nx) + static_cast<Idx>(i);
}
__device__ __forceinline__ float psi_at(const float* __restrict__ psi, const DevParams& p, int i, int j, int k) {
    return (i >= p.nx || k >= p.nz) ? 0.0f : ld_at(psi, psi_index(p, i, j, k));
}
// Cell (x, y, z) writes the three edges it owns; the top row of cells also
// writes the edges of the top face (j = ny).
__device__ __forceinline__ void write_curl_potential(float* __restrict__ px, float* __restrict__ py, float* __restrict__ pz, const DevParams& p, int x,
                                                     int y, int z, float t, const CurlLattice& L) {
    const float fx = static_cast<float>(x) * p.invN[0], cx = (static_cast<float>(x) + 0.5f) * p.invN[0];
    const float fy = static_cast<float>(y) * p.invN[1], cy = (static_cast<float>(y) + 0.5f) * p.invN[1];
    const float fz = static_cast<float>(z) * p.invN[2], cz = (static_cast<float>(z) + 0.5f) * p.invN[2];
    const bool topWall = p.closedTop != 0;
    st_at(px, psi_index(p, x, y, z), (y == 0 || z == 0) ? 0.0f : curl_potential(0, cx, fy, fz, t, p, L));
    st_at(py, psi_index(p, x, y, z), (x == 0 || z == 0) ? 0.0f : curl_potential(1, fx, cy, fz, t, p, L));
    st_at(pz, psi_index(p, x, y, z), (x == 0 || y == 0) ? 0.0f : curl_potential(2, fx, fy, cz, t, p, L));
    if (y == p.ny - 1) {
        st_at(px, psi_index(p, x, p.ny, z), (topWall || z == 0) ? 0.0f : curl_potential(0, cx, 1.0f, fz, t, p, L));
        st_at(pz, psi_index(p, x, p.ny, z), (topWall || x == 0) ? 0.0f : curl_potential(2, fx, 1.0f, cz, t, p, L));
    }
}

template <bool B>
__device__ __forceinline__ void force_contrib_at(float* __restrict__ cu, float* __restrict__ cv, float* __restrict__ cw, float* __restrict__ psix,
                                                 float* __restrict__ psiy, float* __restrict__ psiz, const Scalars<B>& S,
                                                 const float* __restrict__ U, const float* __restrict__ V, const float* __restrict__ W,
                                                 const DevParams& p, const CurlLattice& L, int x, int y, int z) {
    // The two noise fields first, by every lane of the wavefront (lanes are
    // consecutive x cells of one row, so noiseA's lattice row and noiseB's
    // lattice column are shared by the whole wavefront). With curl noise on,
    // the fbm terms are replaced by the potential's curl on the faces.
    const float u = (static_cast<float>(x) + 0.5f) * p.invN[0];
    const float v = (static_cast<float>(y) + 0.5f) * p.invN[1];
    const float w = (static_cast<float>(z) + 0.5f) * p.invN[2];
    const float t = ld_at(p.clock, 0u);
    const bool curl = p.curlOn;
    float noiseA = 0.0f, noiseB = 0.0f;
    if (!curl) {
        noiseA = fbm2_t<1>(u * 9.0f + w * 3.7f + t * 0.22f, v * 13.0f - t * 0.31f, p.saltNoise2) - 0.5f;
        noiseB = fbm2_t<2>(w * 10.0f - t * 0.19f, u * 8.0f + v * 5.5f + t * 0.27f, p.saltNoise2) - 0.5f;
    }
    if (x >= p.nx || y >= p.ny || z >= p.nz) return;
    if (curl) write_curl_potential(psix, psiy, psiz, p, x, y, z, t, L);
    const Idx idx = sidx(p, x, y, z);
    const int sid = p.sceneId;
    const float h = S.template ld<HEAT>(idx), s = S.template ld<SOOT>(idx), k = S.template ld<TURB>(idx), c = S.template ld<PROGRESS>(idx);
    const int xm = max(x - 1, 0), xp = min(x + 1, p.nx - 1);
    const int ym = max(y - 1, 0), yp = min(y + 1, p.ny - 1);
    const int zm = max(z - 1, 0), zp = min(z + 1, p.nz - 1);
    const float gradX = S.template ld<HEAT>(sidx(p, xp, y, z)) - S.template ld<HEAT>(sidx(p, xm, y, z));
    const float gradY = S.template ld<HEAT>(sidx(p, x, yp, z)) - S.template ld<HEAT>(sidx(p, x, ym, z));
    const float gradZ = S.template ld<HEAT>(sidx(p, x, y, zp)) - S.template ld<HEAT>(sidx(p, x, y, zm));
    const float uL = ld_at(U, uidx(p, x, y, z)), uR = ld_at(U, uidx(p, x + 1, y, z));
    const float vD = ld_at(V, vidx(p, x, y, z)), vU = ld_at(V, vidx(p, x, y + 1, z));
    const float wB = ld_at(W, widx(p, x, y, z)), wF = ld_at(W, widx(p, x, y, z + 1));
    const float curlX = (wF - wB) - (vU - vD);
    const float curlY = (uR - uL) - (wF - wB);
    const float curlZ = (vU - vD) - (uR - uL);
    const float curlMag = fabsf(curlX) + fabsf(curlY) + fabsf(curlZ);
    const float thermal = sstep(0.04f, 4.20f, h + s * 0.28f);
    const float upperDrag = sstep(0.26f, 0.94f, v);
    const float lesGain = sqrtf(fmaxf(0.0f, k)) * (0.30f + c * 0.54f);
    const float curlScale = sid == 2 ? 0.38f : (sid == 1 ? 1.18f : 1.0f);
    const float buoyScale = sid == 2 ? 0.42f : (sid == 1 ? 1.06f : 1.0f);
    const float curlGain = p.turbulence * curlScale * (0.18f + thermal * 0.82f + lesGain * 1.12f);
    const float vortexGain = p.turbulence * curlScale * thermal * sat(curlMag * 0.46f) * 0.68f;
    const float buoyFade = sid == 2 ? sstep(0.24f, 0.035f, v) : sstep(0.96f, 0.08f, v);
    const float vertDamp = 0.5f * (vD + vU) * (0.10f + upperDrag * 0.72f + s * 0.020f);
    const float buoyancy = h * buoyScale * (1.62f + p.turbulence * 0.34f + c * 0.16f) * buoyFade - s * 0.095f - gradY * 0.18f - vertDamp;
    const float windProfile = 0.34f + v * 0.96f;
    st_at(cv, idx, p.dt * (buoyancy + curlY * vortexGain * 0.22f));
    st_at(cu, idx, p.dt * (p.windX * windProfile + (-gradZ * 1.36f + noiseA * (0.20f + upperDrag * 0.16f) + curlX * vortexGain) * curlGain));
    st_at(cw, idx, p.dt * (p.windZ * windProfile + (gradX * 1.36f + noiseB * (0.20f + upperDrag * 0.16f) + curlZ * vortexGain) * curlGain));
}

// The force-contribution kernel's grid of 64 x 4 x kForcePlanes cell blocks.
dim3 force_grid(int nx, int ny, int nz) {
    const dim3 b = cell_block();
    return dim3((nx + b.x - 1) / b.x, (ny + b.y - 1) / b.y, (nz + kForcePlanes - 1) / kForcePlanes);
}

// With a non-null `partial`, workgroup (0, 0, 0) first sums the react
// partials as k_finalize_release does (same order, same result: its first 256
// threads are that kernel's four wavefronts), hidden behind the other
// workgroups of this launch. Workgroups are 64 x 4 x kForcePlanes cells, so
// the curl-noise lattice is built once for 1024 cells.
template <bool B>
__global__ __launch_bounds__(1024) void k_force_contrib(float* __restrict__ cu, float* __restrict__ cv, float* __restrict__ cw, float* __restrict__ psix,
                                                        float* __restrict__ psiy, float* __restrict__ psiz, Scalars<B> S,
                                                        const float* __restrict__ U, const float* __restrict__ V, const float* __restrict__ W,
                                                        DevParams p, const double* __restrict__ partial, size_t npartial,
                                                        double* __restrict__ acc) {
    if (partial != nullptr && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0) {
        __shared__ double red[4];
        const unsigned t = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
        if (t < 256) {
            const double v = wave_sum(strided_sum_256(partial, npartial, t));
            if ((t & 63) == 0) red[t >> 6] = v;
        }
        __syncthreads();
        if (t == 0) {
            const double sum = (red[0] + red[1]) + (red[2] + red[3]);
            acc[0] = sum;
            acc[1] += sum;
        }
    }
    __shared__ CurlLattice lat;
    if (p.curlOn) build_curl_lattice(lat, p, ld_at(p.clock, 0u));
    const int3 c = cell_of(xcd_block());
    force_contrib_at<B>(cu, cv, cw, psix, psiy, psiz, S, U, V, W, p, lat, c.x, c.y, c.z);
}

// New face values after the gathered force and the boundary clamp
// (velocityBoundaryKernel :1261-1291). Both cells sharing a face evaluate the
// same expression, so the face is bit-identical from either side. The sum is
// formed in FP64 and rounded once when stored; the divergence below uses the
// unrounded values, since it is a difference of nearly equal face values whose
// FP32 rounding would dominate its relative error.
// The potential psi's curl on each face (zero without curl noise), times dt
// and the amplitude.
struct CurlPsi {
    const float *x, *y, *z;
};
__device__ __forceinline__ float curl_u(const CurlPsi& s, const DevParams& p, int fx, int y, int z) {
    const float c = (psi_at(s.z, p, fx, y + 1, z) - psi_at(s.z, p, fx, y, z)) - (psi_at(s.y, p, fx, y, z + 1) - psi_at(s.y, p, fx, y, z));
    return p.dt * p.curlNoise * c;
}
__device__ __forceinline__ float curl_v(const CurlPsi& s, const DevParams& p, int x, int fy, int z) {
    const float c = (psi_at(s.x, p, x, fy, z + 1) - psi_at(s.x, p, x, fy, z)) - (psi_at(s.z, p, x + 1, fy, z) - psi_at(s.z, p, x, fy, z));
    return p.dt * p.curlNoise * c;
}
__device__ __forceinline__ float curl_w(const CurlPsi& s, const DevParams& p, int x, int y, int fz) {
    const float c = (psi_at(s.y, p, x + 1, y, fz) - psi_at(s.y, p, x, y, fz)) - (psi_at(s.x, p, x, y + 1, fz) - psi_at(s.x, p, x, y, fz));
    return p.dt * p.curlNoise * c;
}

__device__ __forceinline__ double face_u(const float* U, const float* cu, const CurlPsi& psi, const DevParams& p, int fx, int y, int z) {
    if (fx == 0 || fx == p.nx) return 0.0;
    double val = ld_at(U, uidx(p, fx, y, z));
    val += ld_at(cu, sidx(p, fx - 1, y, z));
    val += ld_at(cu, sidx(p, fx, y, z));
    if (p.curlOn) val += curl_u(psi, p, fx, y, z);
    return fmin(4.0, fmax(-4.0, val));
}
__device__ __forceinline__ double face_v(const float* V, const float* cv, const CurlPsi& psi, const DevParams& p, int x, int fy, int z) {
    if (fy == 0 || (p.closedTop && fy == p.ny)) return 0.0;
    double val = ld_at(V, vidx(p, x, fy, z));
    val += ld_at(cv, sidx(p, x, fy - 1, z));
    if (p.curlOn) val += curl_v(psi, p, x, fy, z);
    return fmin(static_cast<double>(6.4f), fmax(static_cast<double>(-0.9f), val));  // the reference's float bounds
}
__device__ __forceinline__ double face_w(const float* W, const float* cw, const CurlPsi& psi, const DevParams& p, int x, int y, int fz) {
    if (fz == 0 || fz == p.nz) return 0.0;
    double val = ld_at(W, widx(p, x, y, fz));
    val += ld_at(cw, sidx(p, x, y, fz - 1));
    val += ld_at(cw, sidx(p, x, y, fz));
    if (p.curlOn) val += curl_w(psi, p, x, y, fz);
    return fmin(4.0, fmax(-4.0, val));
}

// The force pass's output for a 64 x 4 block of one plane (thread (tx, ty)
// owns x = x0 + tx, y = y0 + ty): each cell stores its u, v and w faces (and
// the wall faces past the domain's last column, row and plane), its
// divergence -(uR - uL + vU - vD + wF - wB) from the unrounded face values, and
// p = 0 (the divergence kernel :1185-1199) unless P is null: a step whose
// pressure pass starts from zero without reading p (zeroGuess below) skips
// that store. The u and v faces are computed
// once: a cell's u face x + 1 is the next lane's u face x and its v face y + 1
// the v face y of the row above (through LDS); lane 63, the top row of the
// block and the domain's last column and row compute those faces themselves.
// Called by every thread of the workgroup (it synchronizes); vrow holds 256
// doubles and may be reused on return.
__device__ __forceinline__ void force_apply_block(float* __restrict__ uo, float* __restrict__ vo, float* __restrict__ wo,
                                                  const float* __restrict__ U, const float* __restrict__ V, const float* __restrict__ W,
                                                  const float* __restrict__ cu, const float* __restrict__ cv, const float* __restrict__ cw,
                                                  const CurlPsi& psi, float* __restrict__ D, float* __restrict__ P, const DevParams& p, uint3 blk,
                                                  double* vrow) {
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int x = static_cast<int>(blk.x) * 64 + tx, y = static_cast<int>(blk.y) * 4 + ty, z = static_cast<int>(blk.z);
    const bool in = x < p.nx && y < p.ny && z < p.nz;
    double uL = 0.0, vD = 0.0, wB = 0.0, wF = 0.0;
    if (in) {
        uL = face_u(U, cu, psi, p, x, y, z);
        vD = face_v(V, cv, psi, p, x, y, z);
        wB = face_w(W, cw, psi, p, x, y, z);
        wF = face_w(W, cw, psi, p, x, y, z + 1);
    }
    double uR = __shfl_down(uL, 1, 64);
    vrow[ty * 64 + tx] = vD;
    __syncthreads();
    double vU = ty < 3 ? vrow[(ty + 1) * 64 + tx] : 0.0;
    __syncthreads();
    if (!in) return;
    if (tx == 63 || x == p.nx - 1) uR = face_u(U, cu, psi, p, x + 1, y, z);
    if (ty == 3 || y == p.ny - 1) vU = face_v(V, cv, psi, p, x, y + 1, z);
    st_at(uo, uidx(p, x, y, z), static_cast<float>(uL));
    if (x == p.nx - 1) st_at(uo, uidx(p, x + 1, y, z), static_cast<float>(uR));
    st_at(vo, vidx(p, x, y, z), static_cast<float>(vD));
    if (y == p.ny - 1) st_at(vo, vidx(p, x, y + 1, z), static_cast<float>(vU));
    st_at(wo, widx(p, x, y, z), static_cast<float>(wB));
    if (z == p.nz - 1) st_at(wo, widx(p, x, y, z + 1), static_cast<float>(wF));
    const float div = static_cast<float>(uR - uL + vU - vD + wF - wB);
    const Idx idx = sidx(p, x, y, z);
    st_at(D, idx, -div);
    if (P != nullptr) st_at(P, idx, 0.0f);
}

__global__ __launch_bounds__(256) void k_force_apply(float* __restrict__ uo, float* __restrict__ vo, float* __restrict__ wo,
                                                     const float* __restrict__ U, const float* __restrict__ V, const float* __restrict__ W,
                                                     const float* __restrict__ cu, const float* __restrict__ cv, const float* __restrict__ cw,
                                                     CurlPsi psi, float* __restrict__ D, float* __restrict__ P, DevParams p) {
    __shared__ double vrow[256];
    force_apply_block(uo, vo, wo, U, V, W, cu, cv, cw, psi, D, P, p, xcd_block(), vrow);
}

// -------------------------------------------------------------- K12 ----
// The SOR cell update of sorPressureParityKernel (src/fire_cuda.cu:1201-1227):
// neighbours summed in the order L, R, D, U, B, F after d. Every SOR kernel
// calls this one function, so all of them evaluate the same FP32 expression.
__device__ __forceinline__ float sor_update(float dv, float pc, float pL, float pR, float pD, float pU, float pB, float pF, float omega) {
    const float projected = (dv + pL + pR + pD + pU + pB + pF) * (1.0f / 6.0f);
    return pc + omega * (projected - pc);
}

// One red/black half-sweep, in place: the updated colour (x+y+z)&1 == parity
// reads only the other colour, so every cell of a half-sweep is independent.
// Thread (tx, y, z) of a ((nx+1)/2, ny, nz) grid updates cell x = 2tx + colour.
__device__ __forceinline__ void sor_half_at(float* __restrict__ P, const float* __restrict__ D, int parity, float omega, int nx, int ny, int nz,
                                            int tx, int y, int z) {
    const int x = 2 * tx + ((parity ^ y ^ z) & 1);
    if (x >= nx || y >= ny || z >= nz) return;
    const Idx row = static_cast<Idx>(nx);
    const Idx plane = row * static_cast<Idx>(ny);
    const Idx idx = static_cast<Idx>(z) * plane + static_cast<Idx>(y) * row + static_cast<Idx>(x);
    const float pc = ld_at(P, idx);
    const float pL = x > 0 ? ld_at(P, idx - 1) : pc;
    const float pR = x < nx - 1 ? ld_at(P, idx + 1) : pc;
    const float pD = y > 0 ? ld_at(P, idx - row) : pc;
    const float pU = y < ny - 1 ? ld_at(P, idx + row) : pc;
    const float pB = z > 0 ? ld_at(P, idx - plane) : pc;
    const float pF = z < nz - 1 ? ld_at(P, idx + plane) : pc;
    st_at(P, idx, sor_update(ld_at(D, idx), pc, pL, pR, pD, pU, pB, pF, omega));
}

__global__ __launch_bounds__(256) void k_sor_half(float* __restrict__ P, const float* __restrict__ D, int parity, float omega, int nx, int ny, int nz) {
    const int3 c = cell_of(xcd_block());
    sor_half_at(P, D, parity, omega, nx, ny, nz, c.x, c.y, c.z);
}

// Temporally blocked red-black SOR. A workgroup owns a T^3 tile, loads p on
// the tile plus a K-cell halo (edge E = T + 2K) into LDS, and runs `count`
// <= K consecutive half-sweeps there in place. After half-sweep s only cells
// at least s+1 from an interior-facing face of the loaded box have correct
// inputs, so the computed box shrinks by one per half-sweep (faces on the
// domain boundary do not shrink: their neighbours are the cells themselves, as
// in the reference's clamped indices). The interior T^3 is exact after K
// half-sweeps and is the only part written, to a second pressure buffer
// (tiles overlap through their halos, so a launch must not overwrite its own
// input). Each cell update is the same FP32 expression, in the same order, as
// k_sor_half, so the result is bit-identical to 2N plain half-sweeps.
//
// Thread layout (E^2 threads): for loads, thread t owns LDS column
// (x = t % E, y = t / E) and walks z, so addresses only increment. For the
// half-sweeps, thread t owns the colour cell pair k = t % (E/2) of row
// y = (t / (E/2)) % E and every other z plane starting at z = t / (E^2/2);
// along its planes the colour parity of its x position is fixed. Every
// neighbour is read at a fixed LDS offset: at a domain wall the slot past the
// wall holds a copy of the wall cell (the reference's clamped-index neighbour).
// d is read from global memory (L2) inside the sweep, which keeps the kernel
// near 35 VGPRs. The tile shapes trade halo redundancy against occupancy:
// 16^3 tiles (55 KB of LDS, two workgroups per CU) suit small grids, 24x24
// tiles with a 32x32 loaded face (1024 threads, one or two per CU) larger ones.
// The 26x26 tiles run K = 3 on the same 32x32 face: less halo per cell than
// 24x24 and five of them span 128 cells (six 24-wide tiles span 144). The
// 22x22 tiles run K = 5: the fewest launches, for small, launch-bound grids.
// Unroll depth of the sweep loop: 4 for deep tiles (one workgroup per CU,
// register room to spare), 2 where two workgroups share a CU.
// With STREAM set (the one-scene kernel), the core is written with
// nontemporal stores: no workgroup of the launch reads it back. At 128^3 that
// takes a 26x26x13 launch's load-and-store phase from 5.95 to 5.08 us and the
// 80-iteration solve from 0.584 to 0.567 ms. The projection reading the solved
// field then runs 3 us slower, but storing the last launch normally costs that
// launch 4.5 us, so every launch streams. The batched kernel stores normally:
// streaming cost the served request stream 2.4% of its rate.
constexpr int sor_unroll(int tz) { return tz >= 24 ? 4 : 2; }

template <int T, int K, int TZ, bool STREAM, int UNR = sor_unroll(TZ)>
__device__ __forceinline__ void sor_tile(const float* __restrict__ P, float* __restrict__ Pout, const float* __restrict__ D, float omega, int h0,
                                         int count, int nx, int ny, int nz, float* __restrict__ sp, uint3 blk, bool zeroIn = false) {
    constexpr int E = T + 2 * K;    // loaded edge in x and y
    constexpr int EZ = TZ + 2 * K;  // loaded depth in z
    constexpr int NT = E * E;  // threads per workgroup
    constexpr int HX = E / 2;  // cells of one colour per LDS row
    static_assert(E % 2 == 0 && NT % 64 == 0, "tile edge plus halo must be even and E^2 a multiple of 64");
    const int tid = threadIdx.x;
    const int ox = blk.x * T, oy = blk.y * T, oz = blk.z * TZ;
    const int bx = ox - K, by = oy - K, bz = oz - K;  // global coordinate of local 0
    const Idx row = static_cast<Idx>(nx), plane = row * static_cast<Idx>(ny);
    if (zeroIn) {
        // The solve's first launch from p = 0: nothing to load.
#pragma unroll
        for (int lz = 0; lz < EZ; ++lz) sp[lz * NT + tid] = 0.0f;
    } else {
        // Box cells outside the domain take the value of the nearest wall
        // cell (indices clamped per axis). The slot just past a wall then
        // holds that wall cell's value: the reference's missing neighbour
        // (the cell itself, by its clamped indices) read at a fixed offset.
        const int gx = min(max(bx + tid % E, 0), nx - 1), gy = min(max(by + tid / E, 0), ny - 1);
        float v[EZ];
        const Idx col = static_cast<Idx>(gy) * row + static_cast<Idx>(gx);
#pragma unroll
        for (int lz = 0; lz < EZ; ++lz) v[lz] = ld_at(P, static_cast<Idx>(min(max(bz + lz, 0), nz - 1)) * plane + col);
#pragma unroll
        for (int lz = 0; lz < EZ; ++lz) sp[lz * NT + tid] = v[lz];
    }
    __syncthreads();
    // local index range of real cells
    const int rx0 = max(0, -bx), ry0 = max(0, -by), rz0 = max(0, -bz);
    const int rx1 = min(E, nx - bx), ry1 = min(E, ny - by), rz1 = min(EZ, nz - bz);
    const int k = tid % HX, ly = (tid / HX) % E;
    // The z group is the same for a whole wavefront when E^2/2 threads fill
    // whole wavefronts: the plane loop then runs on scalar registers.
    int zg = tid / (HX * E);
    if constexpr ((HX * E) % 64 == 0) zg = __builtin_amdgcn_readfirstlane(zg);
    const int gy = by + ly;
    // Per colour parity: the thread's x in the box, its LDS offset within a
    // plane, its d offset within a plane, and its wall-copy offset (the slot
    // past the wall of an x-wall cell, else 0).
    const int c1 = (bx + gy + bz + zg) & 1;
    int lxP[2], iP[2], cxP[2];
    Idx dP[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        lxP[q] = 2 * k + (q ^ c1);
        const int gx = bx + lxP[q];
        iP[q] = ly * E + lxP[q];
        dP[q] = static_cast<Idx>(gy) * row + static_cast<Idx>(gx);
        cxP[q] = gx == 0 ? -1 : (gx == nx - 1 ? 1 : 0);
    }
    const int cy = gy == 0 ? -E : (gy == ny - 1 ? E : 0);
    // Boxes that reach a domain wall in x or y carry wall cells; their sweeps
    // also write each updated cell to its wall-copy slot (the same slot again
    // for cells off the wall), so no separate refresh pass is needed.
    const bool wallX = bx < 0 || bx + E >= nx, wallY = by < 0 || by + E >= ny;
    for (int s = 0; s < count; ++s) {
        const int q = (h0 + s) & 1;
        const int ax = bx < 0 ? rx0 : s + 1, ay = by < 0 ? ry0 : s + 1, az = bz < 0 ? rz0 : s + 1;
        const int ex = ox + T + K > nx ? rx1 : E - 1 - s;
        const int ey = oy + T + K > ny ? ry1 : E - 1 - s;
        const int ez = oz + TZ + K > nz ? rz1 : EZ - 1 - s;
        const int lx = lxP[q];
        const int lz0 = az + ((az ^ zg) & 1);
        if (lx >= ax && lx < ex && ly >= ay && ly < ey && lz0 < ez) {
            const int i0 = lz0 * NT + iP[q];
            const int n = (ez - lz0 + 1) >> 1;  // planes lz0, lz0 + 2, ... below ez
            const int cx = cxP[q];
            Idx di = static_cast<Idx>(bz + lz0) * plane + dP[q];
            int i = i0;
            // The cell above plane lz (other colour, unchanged in this
            // half-sweep) is the cell below plane lz + 2: read it once.
            float pB = sp[i0 - NT];
            float v = 0.0f;
            auto sweep = [&](auto wx, auto wy) {
#pragma unroll UNR
                for (int j = 0; j < n; ++j, i += 2 * NT, di += 2 * plane) {
                    const float pF = sp[i + NT];
                    v = sor_update(ld_at(D, di), sp[i], sp[i - 1], sp[i + 1], sp[i - E], sp[i + E], pB, pF, omega);
                    sp[i] = v;
                    if constexpr (decltype(wx)::value) sp[i + cx] = v;
                    if constexpr (decltype(wy)::value) sp[i + cy] = v;
                    pB = pF;
                }
            };
            if (wallX) {
                if (wallY) sweep(std::true_type{}, std::true_type{});
                else sweep(std::true_type{}, std::false_type{});
            } else {
                if (wallY) sweep(std::false_type{}, std::true_type{});
                else sweep(std::false_type{}, std::false_type{});
            }
            // z-wall cells refresh the slot past their wall, which only they
            // read (in a later half-sweep).
            if (bz + lz0 == 0) sp[i0 - NT] = sp[i0];
            if (bz + lz0 + 2 * (n - 1) == nz - 1) sp[i0 + (2 * n - 1) * NT] = v;
        }
        __syncthreads();
    }
    if (tid < T * T) {
        const int tx = tid % T, ty = tid / T;
        const int gx = ox + tx, gy2 = oy + ty;
        if (gx < nx && gy2 < ny) {
            const Idx dst = static_cast<Idx>(gy2) * row + static_cast<Idx>(gx);
            const int src = (K * E + ty + K) * E + tx + K;
#pragma unroll
            for (int tz = 0; tz < TZ; ++tz) {
                if (oz + tz >= nz) continue;
                const Idx o = static_cast<Idx>(oz + tz) * plane + dst;
                if constexpr (STREAM) st_stream(Pout, o, sp[src + tz * NT]);
                else st_at(Pout, o, sp[src + tz * NT]);
            }
        }
    }
}

template <int T, int K, int TZ>
__global__ __launch_bounds__((T + 2 * K) * (T + 2 * K)) void k_sor_tb(const float* __restrict__ P, float* __restrict__ Pout, const float* __restrict__ D,
                                                                     float omega, int h0, int count, int nx, int ny, int nz, int zeroIn) {
    __shared__ float sp[(T + 2 * K) * (T + 2 * K) * (TZ + 2 * K)];
    sor_tile<T, K, TZ, true>(P, Pout, D, omega, h0, count, nx, ny, nz, sp, xcd_block(), zeroIn != 0);
}

// ----------------------------------------------------------- K13-K14 ----
// Face grid of (nx, ny+1, nz+1), as advect_velocity_at (x = nx - 1 also zeroes the wall u face at x = nx).
__device__ __forceinline__ void project_at(float* __restrict__ U, float* __restrict__ V, float* __restrict__ W, const float* __restrict__ P,
                                           const DevParams& p, int x, int y, int z) {
    if (x >= p.nx || y > p.ny || z > p.nz) return;
    if (x == p.nx - 1 && y < p.ny && z < p.nz) st_at(U, uidx(p, p.nx, y, z), 0.0f);
    if (y < p.ny && z < p.nz) {
        const Idx f = uidx(p, x, y, z);
        if (x == 0 || x == p.nx) st_at(U, f, 0.0f);
        else st_at(U, f, ld_at(U, f) - (ld_at(P, sidx(p, x, y, z)) - ld_at(P, sidx(p, x - 1, y, z))));
    }
    if (x < p.nx && z < p.nz) {
        const Idx f = vidx(p, x, y, z);
        if (y == 0) st_at(V, f, 0.0f);
        else if (y < p.ny) st_at(V, f, ld_at(V, f) - (ld_at(P, sidx(p, x, y, z)) - ld_at(P, sidx(p, x, y - 1, z))));
    }
    if (x < p.nx && y < p.ny) {
        const Idx f = widx(p, x, y, z);
        if (z == 0 || z == p.nz) st_at(W, f, 0.0f);
        else st_at(W, f, ld_at(W, f) - (ld_at(P, sidx(p, x, y, z)) - ld_at(P, sidx(p, x, y, z - 1))));
    }
}

__global__ __launch_bounds__(256) void k_project(float* __restrict__ U, float* __restrict__ V, float* __restrict__ W, const float* __restrict__ P,
                                                 DevParams p) {
    const int3 c = cell_of(xcd_block());
    project_at(U, V, W, P, p, c.x, c.y, c.z);
}

// -------------------------------------------------------------- reset ----
template <bool B>
__global__ __launch_bounds__(256) void k_reset(Scalars<B> S, float* __restrict__ P, float* __restrict__ D, DevParams p) {
    const uint3 blk = xcd_block();
    const int x = blk.x * blockDim.x + threadIdx.x;
    const int y = blk.y * blockDim.y + threadIdx.y;
    const int z = blk.z * blockDim.z + threadIdx.z;
    if (x >= p.nx || y >= p.ny || z >= p.nz) return;
    const Idx idx = sidx(p, x, y, z);
    const float u = (static_cast<float>(x) + 0.5f) * p.invN[0];
    const float v = (static_cast<float>(y) + 0.5f) * p.invN[1];
    const float w = (static_cast<float>(z) + 0.5f) * p.invN[2];
    const float wx = (u - 0.50f) * 2.10f, wz = (w - 0.50f) * 1.64f;
    const float material = bed_material(wx, wz, p);
    float heat = 0, fuel = 0, soot = 0, ch = 0, ash = 0, pyro = 0, prog = 0, k = 0, so = 0;
    if (material != 0.0f) {
        const int sid = p.sceneId;
        const float chunk = fbm2(wx * 18.0f + wz * 2.0f, wz * 22.0f - wx * 5.0f, p.saltNoise2);
        const float bed = bed_source(material, wx, wz, v, p, 0.0f);
        const float charScale = sid == 2 ? 0.004f : (sid == 1 ? 1.18f : 1.0f);
        const float sootScale = sid == 2 ? 0.004f : (sid == 1 ? 0.92f : 1.0f);
        heat = bed * (sid == 2 ? 1.34f : (sid == 1 ? 0.78f + chunk * 0.26f : 0.86f + chunk * 0.24f));
        fuel = bed * (sid == 2 ? 1.42f : (sid == 1 ? 1.05f + chunk * 0.30f : 1.08f + chunk * 0.26f));
        soot = bed * (sid == 2 ? 0.00065f : 0.018f) * sootScale;
        ch = material * (0.92f + chunk * 0.72f) * charScale;
        ash = material * (0.018f + sstep(0.56f, 0.86f, chunk) * 0.055f) * (sid == 2 ? 0.10f : 1.0f);
        pyro = bed * (sid == 2 ? 0.018f : (sid == 1 ? 0.060f + chunk * 0.034f : 0.058f + chunk * 0.028f));
        prog = bed * (sid == 2 ? 0.11f : (sid == 1 ? 0.24f : 0.21f));
        k = bed * (sid == 2 ? 0.012f : (sid == 1 ? 0.082f : 0.065f));
        so = bed * (sid == 2 ? 0.00050f : 0.018f) * sootScale;
    }
    S.template st<HEAT>(idx, heat);
    S.template st<FUEL>(idx, fuel);
    S.template st<OXYGEN>(idx, 1.0f);
    S.template st<SOOT>(idx, soot);
    S.template st<CHARF>(idx, ch);
    S.template st<ASH>(idx, ash);
    S.template st<PYRO>(idx, pyro);
    S.template st<PROGRESS>(idx, prog);
    S.template st<TURB>(idx, k);
    S.template st<SOPT>(idx, so);
    st_at(P, idx, 0.0f);
    st_at(D, idx, 0.0f);
}

// ------------------------------------------------------------ metrics ----
template <bool B>
__global__ __launch_bounds__(256) void k_metrics(double* __restrict__ partial, Scalars<B> S, const float* __restrict__ U, const float* __restrict__ V,
                                                 const float* __restrict__ W, const float* __restrict__ P, const float* __restrict__ D,
                                                 const float4* __restrict__ L, const float* __restrict__ SH, DevParams p) {
    double acc[kMetricSlots];
#pragma unroll
    for (int i = 0; i < kMetricSlots; ++i) acc[i] = 0.0;
    const size_t cells = static_cast<size_t>(p.nx) * p.ny * p.nz;
    for (size_t idx = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; idx < cells; idx += static_cast<size_t>(gridDim.x) * blockDim.x) {
        const int x = static_cast<int>(idx % p.nx);
        const int y = static_cast<int>((idx / p.nx) % p.ny);
        const int z = static_cast<int>(idx / (static_cast<size_t>(p.nx) * p.ny));
        const float heat = S.template ld<HEAT>(idx), fuel = S.template ld<FUEL>(idx), oxygen = S.template ld<OXYGEN>(idx);
        const float soot = S.template ld<SOOT>(idx), ch = S.template ld<CHARF>(idx), ash = S.template ld<ASH>(idx);
        const float pyro = S.template ld<PYRO>(idx), prog = S.template ld<PROGRESS>(idx), turb = S.template ld<TURB>(idx);
        const float so = S.template ld<SOPT>(idx);
        const float divBefore = fabsf(ld_at(D, idx));
        const float div = ld_at(U, uidx(p, x + 1, y, z)) - ld_at(U, uidx(p, x, y, z)) + ld_at(V, vidx(p, x, y + 1, z)) - ld_at(V, vidx(p, x, y, z)) + ld_at(W, widx(p, x, y, z + 1)) -
                          ld_at(W, widx(p, x, y, z));
        const float divAfter = fabsf(-div);
        const float pc = ld_at(P, idx);
        const double resid = static_cast<double>(ld_at(D, idx)) + ld_at(P, sidx(p, max(x - 1, 0), y, z)) + ld_at(P, sidx(p, min(x + 1, p.nx - 1), y, z)) +
                             ld_at(P, sidx(p, x, max(y - 1, 0), z)) + ld_at(P, sidx(p, x, min(y + 1, p.ny - 1), z)) + ld_at(P, sidx(p, x, y, max(z - 1, 0))) +
                             ld_at(P, sidx(p, x, y, min(z + 1, p.nz - 1))) - 6.0 * pc;
        const float u = (static_cast<float>(x) + 0.5f) / static_cast<float>(p.nx);
        const float v = (static_cast<float>(y) + 0.5f) / static_cast<float>(p.ny);
        const float w = (static_cast<float>(z) + 0.5f) / static_cast<float>(p.nz);
        const float4 Lc = sample_light(L, p, u, v, w);
        const float shadow = sample_shadow(SH, p, u, v, w);
        if (!isfinite(heat) || !isfinite(fuel) || !isfinite(oxygen) || !isfinite(soot) || !isfinite(ch) || !isfinite(ash) || !isfinite(pyro) ||
            !isfinite(prog) || !isfinite(turb) || !isfinite(so) || !isfinite(divBefore) || !isfinite(divAfter)) {
            acc[31] += 1.0;
            continue;
        }
        const float tempK = 293.0f + heat * 335.0f;
        const float hrr = (fuel + pyro * 0.42f) * oxygen * sstep(610.0f, 1280.0f, tempK) * expf(-5200.0f / fmaxf(650.0f, tempK));
        const float od = so + soot * (0.24f + 0.05f * sqrtf(fmaxf(0.0f, heat)));
        const float nh = v;
        const float flameMass = sstep(0.72f, 2.10f, heat + prog * 0.62f + pyro * 0.24f) * sstep(0.035f, 0.66f, oxygen) * sstep(0.012f, 0.52f, fuel + pyro * 0.36f);
        const float smokeMass = sstep(0.035f, 1.16f, so + soot * 0.38f + ash * 0.12f) * sstep(0.08f, 0.88f, nh);
        const float volShadow = shadow * expf(-Lc.w * 0.80f);
        const float lum = lum3(f3(Lc.x, Lc.y, Lc.z));
        const float marker = heat + fuel * 0.24f + prog * 0.55f;
        const float sums[21] = {heat, fuel, oxygen, soot, ch, ash, pyro, prog, turb, so, lum, shadow, flameMass, smokeMass,
                                flameMass * smokeMass, volShadow, lum * volShadow, od, hrr, 0.0f, 0.0f};
#pragma unroll
        for (int i = 0; i < 19; ++i) acc[i] += static_cast<double>(sums[i]);
        acc[19] += static_cast<double>(divBefore) * divBefore;
        acc[20] += static_cast<double>(divAfter) * divAfter;
        acc[21] += resid * resid;
        const float mx[8] = {heat, fuel, soot, pyro, prog, turb, divBefore, divAfter};
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (mx[i] > 0.0f) acc[22 + i] = fmax(acc[22 + i], static_cast<double>(mx[i]));
        if (marker > 0.74f) acc[30] = fmax(acc[30], static_cast<double>(nh * 2.03f));
    }
    // block reduction: sums for 0..21 and 31, maxima for 22..30
    __shared__ double lds[4][kMetricSlots];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < kMetricSlots; ++i) {
        const bool isMax = i >= 22 && i <= 30;
        const double r = isMax ? wave_max(acc[i]) : wave_sum(acc[i]);
        if (lane == 0) lds[wave][i] = r;
    }
    __syncthreads();
    if (threadIdx.x < kMetricSlots) {
        const int i = threadIdx.x;
        const bool isMax = i >= 22 && i <= 30;
        const double r = isMax ? fmax(fmax(lds[0][i], lds[1][i]), fmax(lds[2][i], lds[3][i])) : (lds[0][i] + lds[1][i]) + (lds[2][i] + lds[3][i]);
        partial[static_cast<size_t>(blockIdx.x) * kMetricSlots + i] = r;
    }
}

__global__ __launch_bounds__(256) void k_reduce_metrics(const double* __restrict__ partial, int blocks, double* __restrict__ out) {
    // 8 threads per slot, fixed order
    const int slot = threadIdx.x >> 3, sub = threadIdx.x & 7;
    const bool isMax = slot >= 22 && slot <= 30;
    double r = 0.0;
    for (int b = sub; b < blocks; b += 8) {
        const double val = partial[static_cast<size_t>(b) * kMetricSlots + slot];
        r = isMax ? fmax(r, val) : r + val;
    }
#pragma unroll
    for (int off = 4; off > 0; off >>= 1) {
        const double o = __shfl_down(r, off, 8);
        r = isMax ? fmax(r, o) : r + o;
    }
    if (sub == 0) out[slot] = r;
}

// --------------------------------------------------------- pack/unpack ----
__global__ void k_pack_bf16(uint16_t* __restrict__ dst, const float* __restrict__ src, size_t n) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * blockDim.x) {
        uint32_t b = __float_as_uint(src[i]);
        b += 0x7FFFu + ((b >> 16) & 1u);
        dst[i] = static_cast<uint16_t>(b >> 16);
    }
}
__global__ void k_unpack_bf16(float* __restrict__ dst, const uint16_t* __restrict__ src, size_t n) {
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += static_cast<size_t>(gridDim.x) * blockDim.x)
        dst[i] = __uint_as_float(static_cast<uint32_t>(src[i]) << 16);
}

// ============================================================ batches ====
// Batched passes: one persistent launch covers every scene of a BatchTable.
// Each walks the scenes' block lists and runs the same per-cell body as the
// one-scene kernel above, on the block coordinates the one-scene launch would
// give that tile, so results are identical to stepping each scene alone.
template <bool B>
__device__ __forceinline__ Scalars<B> scal_of(void* const* f) {
    Scalars<B> s;
#pragma unroll
    for (int i = 0; i < NSCALAR; ++i) s.f[i] = f[i];
    return s;
}

// block_sum_256 for use inside a loop: the trailing barrier keeps the next
// iteration's writes from overtaking this iteration's read.
__device__ __forceinline__ double block_sum_256_loop(double v) {
    const double t = block_sum_256(v);
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(256) void kb_advance_clock(const BatchTable* __restrict__ T) {
    for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < T->count; s += gridDim.x * blockDim.x) {
        float* c = T->scene[s].clock;
        const float dt = T->scene[s].p.dt;
        c[0] += dt;
        c[1] += dt;
    }
}

// The persistent loop lifts this kernel to 86 VGPRs (five wavefronts per
// SIMD); held to eight wavefronts it needs 63 and spills nothing.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8))) void kb_advect_velocity(const BatchTable* __restrict__ T) {
    batch_blocks(T->faceStart, T->count, [&](int s, uint32_t local) {
        const BatchScene& S = T->scene[s];
        const int3 c = cell_of(block_of(local, S.faceG));
        advect_velocity_at(S.velB[0], S.velB[1], S.velB[2], S.velA[0], S.velA[1], S.velA[2], S.p, c.x, c.y, c.z);
    });
}

template <bool B>
__global__ __launch_bounds__(256) void kb_advect_react(const BatchTable* __restrict__ T) {
    batch_blocks(T->cellStart, T->count, [&](int s, uint32_t local) {
        const BatchScene& S = T->scene[s];
        const int3 c = cell_of(block_of(local, S.cellG));
        float release = 0.0f;
        if (c.x < S.p.nx && c.y < S.p.ny && c.z < S.p.nz)
            release = react_cell<B>(scal_of<B>(S.sOut), scal_of<B>(S.sIn), S.velB[0], S.velB[1], S.velB[2], S.p, c.x, c.y, c.z);
        const double t = block_sum_256_loop(static_cast<double>(release));
        if (threadIdx.x == 0 && threadIdx.y == 0 && threadIdx.z == 0) S.partials[local] = t;
    });
}

__global__ __launch_bounds__(256) void kb_finalize_release(const BatchTable* __restrict__ T) {
    for (int s = blockIdx.x; s < T->count; s += gridDim.x) {
        const BatchScene& S = T->scene[s];
        const size_t n = static_cast<size_t>(S.cellG[0]) * S.cellG[1] * S.cellG[2];
        const double t = block_sum_256_loop(strided_sum_256(S.partials, n, threadIdx.x));
        if (threadIdx.x == 0) {
            S.acc[0] = t;
            S.acc[1] += t;
        }
    }
}

// Held to eight wavefronts per SIMD as kb_advect_velocity (67 -> 64 VGPRs, no spills).
// A workgroup's consecutive blocks mostly belong to one scene: the curl-noise
// lattice (a function of the scene's salts and clock) is rebuilt only when
// the scene changes.
template <bool B>
__global__ __launch_bounds__(256) void kb_force_contrib(const BatchTable* __restrict__ T) {
    __shared__ CurlLattice lat;
    int built = -1;
    batch_blocks(T->cellStart, T->count, [&](int s, uint32_t local) {
        const BatchScene& S = T->scene[s];
        if (S.p.curlOn && s != built) {
            build_curl_lattice(lat, S.p, ld_at(S.p.clock, 0u));
            built = s;
        }
        const int3 c = cell_of(block_of(local, S.cellG));
        force_contrib_at<B>(S.cu, S.cv, S.cw, S.psi[0], S.psi[1], S.psi[2], scal_of<B>(S.sOut), S.velB[0], S.velB[1], S.velB[2], S.p, lat, c.x, c.y,
                            c.z);
    });
}

// writeP = 0 when the pressure launches start from zero without reading p.
__global__ __launch_bounds__(256) void kb_force_apply(const BatchTable* __restrict__ T, int writeP) {
    __shared__ double vrow[256];
    batch_blocks(T->cellStart, T->count, [&](int s, uint32_t local) {
        const BatchScene& S = T->scene[s];
        force_apply_block(S.velA[0], S.velA[1], S.velA[2], S.velB[0], S.velB[1], S.velB[2], S.cu, S.cv, S.cw, CurlPsi{S.psi[0], S.psi[1], S.psi[2]},
                          S.d, writeP ? S.p0 : nullptr, S.p, block_of(local, S.cellG), vrow);
    });
}

// Pressure launch j: half-sweeps j*K .. j*K+K-1 of every scene that has them
// (launch 0 from p = 0 without reading p when zeroIn is set).
template <int T_, int K, int TZ>
__global__ __launch_bounds__((T_ + 2 * K) * (T_ + 2 * K)) void kb_sor_tb(const BatchTable* __restrict__ T, int j, int zeroIn) {
    __shared__ float sp[(T_ + 2 * K) * (T_ + 2 * K) * (TZ + 2 * K)];
    const int h0 = j * K;
    batch_blocks(T->sorStart, T->sorActive[j], [&](int s, uint32_t local) {
        const BatchScene& S = T->scene[s];
        const float* in = (j & 1) ? S.p1 : S.p0;
        float* out = (j & 1) ? S.p0 : S.p1;
        sor_tile<T_, K, TZ, false>(in, out, S.d, S.p.omega, h0, min(K, S.halfSweeps - h0), S.p.nx, S.p.ny, S.p.nz, sp, block_of(local, S.sorG),
                            zeroIn != 0);
        __syncthreads();  // the next tile reuses sp
    });
}

// Pressure launch j of the plain kernel: half-sweep j (colour j & 1), in place.
__global__ __launch_bounds__(256) void kb_sor_half(const BatchTable* __restrict__ T, int j) {
    batch_blocks(T->sorStart, T->sorActive[j], [&](int s, uint32_t local) {
        const BatchScene& S = T->scene[s];
        const int3 c = cell_of(block_of(local, S.sorG));
        sor_half_at(S.p0, S.d, j & 1, S.p.omega, S.p.nx, S.p.ny, S.p.nz, c.x, c.y, c.z);
    });
}

__global__ __launch_bounds__(256) void kb_project(const BatchTable* __restrict__ T) {
    batch_blocks(T->faceStart, T->count, [&](int s, uint32_t local) {
        const BatchScene& S = T->scene[s];
        const int3 c = cell_of(block_of(local, S.faceG));
        project_at(S.velA[0], S.velA[1], S.velA[2], S.pFinal, S.p, c.x, c.y, c.z);
    });
}

// ========================================================== launchers ====
template <bool B>
static Scalars<B> scal(void* const* f) {
    Scalars<B> s;
    for (int i = 0; i < NSCALAR; ++i) s.f[i] = f[i];
    return s;
}

hipError_t launch_reset(const DeviceScene& s, const DevParams& p, hipStream_t st) {
    const dim3 g = cell_grid(s.nx, s.ny, s.nz), b = cell_block();
    if (s.bf16) k_reset<true><<<g, b, 0, st>>>(scal<true>(s.s), s.p, s.d, p);
    else k_reset<false><<<g, b, 0, st>>>(scal<false>(s.s), s.p, s.d, p);
    return hipGetLastError();
}

hipError_t launch_advance_clock(DeviceScene& s, const DevParams& p, hipStream_t st) {
    k_advance_clock<<<1, 64, 0, st>>>(s.clock, p.dt);
    return hipGetLastError();
}

hipError_t launch_advect_velocity(DeviceScene& s, const DevParams& p, hipStream_t st, bool advanceClock) {
    const dim3 g = cell_grid(s.nx, s.ny + 1, s.nz + 1), b = cell_block();
    k_advect_velocity<<<g, b, 0, st>>>(s.uN, s.vN, s.wN, s.u, s.v, s.w, p, advanceClock ? s.clock : nullptr);
    std::swap(s.u, s.uN);
    std::swap(s.v, s.vN);
    std::swap(s.w, s.wN);
    return hipGetLastError();
}

hipError_t launch_advect_react(DeviceScene& s, const DevParams& p, hipStream_t st, bool finalize) {
    const dim3 g = cell_grid(s.nx, s.ny, s.nz), b = cell_block();
    if (s.bf16) k_advect_react<true><<<g, b, 0, st>>>(scal<true>(s.sNext), scal<true>(s.s), s.u, s.v, s.w, p, s.partials);
    else k_advect_react<false><<<g, b, 0, st>>>(scal<false>(s.sNext), scal<false>(s.s), s.u, s.v, s.w, p, s.partials);
    if (finalize) k_finalize_release<<<1, 256, 0, st>>>(s.partials, react_partial_count(s.nx, s.ny, s.nz), s.acc);
    for (int i = 0; i < NSCALAR; ++i) std::swap(s.s[i], s.sNext[i]);
    return hipGetLastError();
}

hipError_t launch_force(DeviceScene& s, const DevParams& p, hipStream_t st, bool finalizeRelease, bool writeP) {
    const dim3 g = force_grid(s.nx, s.ny, s.nz), b = cell_block(), bc(64, 4, kForcePlanes);
    const double* part = finalizeRelease ? s.partials : nullptr;
    const size_t np = react_partial_count(s.nx, s.ny, s.nz);
    if (s.bf16) k_force_contrib<true><<<g, bc, 0, st>>>(s.cu, s.cv, s.cw, s.psi[0], s.psi[1], s.psi[2], scal<true>(s.s), s.u, s.v, s.w, p, part, np, s.acc);
    else k_force_contrib<false><<<g, bc, 0, st>>>(s.cu, s.cv, s.cw, s.psi[0], s.psi[1], s.psi[2], scal<false>(s.s), s.u, s.v, s.w, p, part, np, s.acc);
    k_force_apply<<<cell_grid(s.nx, s.ny, s.nz), b, 0, st>>>(s.uN, s.vN, s.wN, s.u, s.v, s.w, s.cu, s.cv, s.cw, CurlPsi{s.psi[0], s.psi[1], s.psi[2]}, s.d,
                                                                     writeP ? s.p : nullptr, p);
    std::swap(s.u, s.uN);
    std::swap(s.v, s.vN);
    std::swap(s.w, s.wN);
    return hipGetLastError();
}

// Temporally blocked SOR kernels by enum fs_sor_kernel: x/y tile edge T,
// z depth TZ, half-sweeps per launch K (the loaded box is (T+2K)^2 x (TZ+2K)).
struct SorShape {
    int t, tz, k;
};
static SorShape sor_shape(int sorKernel) {
    switch (sorKernel) {
        case 1: return {16, 16, 4};
        case 2: return {24, 24, 4};
        case 3: return {24, 8, 4};
        case 4: return {24, 16, 4};
        case 5: return {24, 28, 4};
        case 6: return {24, 10, 4};
        case 7: return {24, 12, 4};
        case 8: return {24, 6, 4};
        case 9: return {26, 13, 3};
        case 10: return {26, 9, 3};
        case 11: return {22, 8, 5};
        default: return {0, 0, 1};  // plain half-sweep kernel
    }
}

int sor_depth(int sorKernel) { return sor_shape(sorKernel).k; }
int sor_tile_edge(int sorKernel) { return sor_shape(sorKernel).t; }
int sor_tile_depth(int sorKernel) { return sor_shape(sorKernel).tz; }

// Tiles read the halo of their neighbours, so a launch must not overwrite its
// own input: launch j reads buffer a (j even) or b (j odd) and writes the other.
// With zeroGuess the first launch starts from p = 0 and reads nothing from a.
// Returns the number of launches.
template <int T, int K, int TZ = T>
static int launch_sor_tb(float* a, float* b, const float* d, int nx, int ny, int nz, float omega, int iterations, hipStream_t st, bool zeroGuess) {
    const dim3 g((nx + T - 1) / T, (ny + T - 1) / T, (nz + TZ - 1) / TZ);
    const int half = 2 * iterations;
    constexpr int E = T + 2 * K;
    int launches = 0;
    for (int h0 = 0; h0 < half; h0 += K, ++launches) {
        k_sor_tb<T, K, TZ><<<g, E * E, 0, st>>>(a, b, d, omega, h0, min(K, half - h0), nx, ny, nz, zeroGuess && h0 == 0);
        std::swap(a, b);
    }
    return launches;
}

hipError_t launch_pressure_buffers(float* p, float* p2, const float* d, int nx, int ny, int nz, float omega, int iterations, int sorKernel,
                                   hipStream_t st, bool* resultInP2, bool zeroGuess) {
    if (zeroGuess && sorKernel == 0) return hipErrorInvalidValue;  // the half-sweep kernel updates p in place
    const bool z = zeroGuess;
    int launches = 0;
    switch (sorKernel) {
        case 1: launches = launch_sor_tb<16, 4>(p, p2, d, nx, ny, nz, omega, iterations, st, z); break;
        case 2: launches = launch_sor_tb<24, 4>(p, p2, d, nx, ny, nz, omega, iterations, st, z); break;
        case 3: launches = launch_sor_tb<24, 4, 8>(p, p2, d, nx, ny, nz, omega, iterations, st, z); break;
        case 4: launches = launch_sor_tb<24, 4, 16>(p, p2, d, nx, ny, nz, omega, iterations, st, z); break;
        case 5: launches = launch_sor_tb<24, 4, 28>(p, p2, d, nx, ny, nz, omega, iterations, st, z); break;
        case 6: launches = launch_sor_tb<24, 4, 10>(p, p2, d, nx, ny, nz, omega, iterations, st, z); break;
        case 7: launches = launch_sor_tb<24, 4, 12>(p, p2, d, nx, ny, nz, omega, iterations, st, z); break;
        case 8: launches = launch_sor_tb<24, 4, 6>(p, p2, d, nx, ny, nz, omega, iterations, st, z); break;
        case 9: launches = launch_sor_tb<26, 3, 13>(p, p2, d, nx, ny, nz, omega, iterations, st, z); break;
        case 10: launches = launch_sor_tb<26, 3, 9>(p, p2, d, nx, ny, nz, omega, iterations, st, z); break;
        case 11: launches = launch_sor_tb<22, 5, 8>(p, p2, d, nx, ny, nz, omega, iterations, st, z); break;
        default: {
            const dim3 b = cell_block();
            const dim3 g = sor_grid(0, nx, ny, nz);
            for (int i = 0; i < iterations; ++i) {
                k_sor_half<<<g, b, 0, st>>>(p, d, 0, omega, nx, ny, nz);
                k_sor_half<<<g, b, 0, st>>>(p, d, 1, omega, nx, ny, nz);
            }
        }
    }
    *resultInP2 = (launches & 1) != 0;
    return hipGetLastError();
}

hipError_t launch_pressure(DeviceScene& s, const DevParams& p, hipStream_t st, bool zeroGuess) {
    bool inP2 = false;
    const hipError_t e = launch_pressure_buffers(s.p, s.p2, s.d, s.nx, s.ny, s.nz, p.omega, p.iterations, s.sorKernel, st, &inP2, zeroGuess);
    if (inP2) std::swap(s.p, s.p2);
    return e;
}

__global__ __launch_bounds__(256) void k_copy_f32(const float* __restrict__ a, float* __restrict__ b, Idx n) {
    const Idx i = blockIdx.x * 256u + threadIdx.x;
    if (i < n) st_at(b, i, ld_at(a, i));
}

hipError_t launch_copy_f32(const float* a, float* b, size_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    if (n >= (static_cast<size_t>(1) << 30)) return hipErrorInvalidValue;
    k_copy_f32<<<static_cast<unsigned>((n + 255) / 256), 256, 0, st>>>(a, b, static_cast<Idx>(n));
    return hipGetLastError();
}

dim3 sor_grid(int sorKernel, int nx, int ny, int nz) {
    const int t = sor_tile_edge(sorKernel), tz = sor_tile_depth(sorKernel);
    if (t == 0) {
        const dim3 b = cell_block();
        return dim3(((nx + 1) / 2 + b.x - 1) / b.x, (ny + b.y - 1) / b.y, (nz + b.z - 1) / b.z);
    }
    return dim3((nx + t - 1) / t, (ny + t - 1) / t, (nz + tz - 1) / tz);
}

int pressure_launch_count(const DeviceScene& s, int iterations) {
    const int k = sor_depth(s.sorKernel);
    return (2 * iterations + k - 1) / k;
}

// Streaming float4 copy, one element per thread, for the achievable-bandwidth
// figure (on a 1 GiB copy this beats grid-stride loops of 2K-32K workgroups,
// with or without non-temporal accesses, and the runtime's own copy).
__global__ __launch_bounds__(256) void k_copy4(const float4* __restrict__ a, float4* __restrict__ b, size_t n4) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n4) b[i] = a[i];
}

hipError_t launch_copy4(const void* a, void* b, size_t bytes, hipStream_t st) {
    const size_t n4 = bytes / 16;
    k_copy4<<<static_cast<unsigned>((n4 + 255) / 256), 256, 0, st>>>(static_cast<const float4*>(a), static_cast<float4*>(b), n4);
    return hipGetLastError();
}

hipError_t launch_project(DeviceScene& s, const DevParams& p, hipStream_t st) {
    const dim3 g = cell_grid(s.nx, s.ny + 1, s.nz + 1), b = cell_block();
    k_project<<<g, b, 0, st>>>(s.u, s.v, s.w, s.p, p);
    return hipGetLastError();
}

hipError_t launch_metrics(DeviceScene& s, const DevParams& p, hipStream_t st) {
    if (s.partialCapacity < static_cast<size_t>(kMetricBlocks) * kMetricSlots) return hipErrorInvalidValue;
    if (s.bf16)
        k_metrics<true><<<kMetricBlocks, 256, 0, st>>>(s.partials, scal<true>(s.s), s.u, s.v, s.w, s.p, s.d, s.light, s.shadow, p);
    else
        k_metrics<false><<<kMetricBlocks, 256, 0, st>>>(s.partials, scal<false>(s.s), s.u, s.v, s.w, s.p, s.d, s.light, s.shadow, p);
    k_reduce_metrics<<<1, 256, 0, st>>>(s.partials, kMetricBlocks, s.metricsOut);
    return hipGetLastError();
}

template <int T, int K, int TZ = T>
static void launch_batch_sor_tb(const BatchTable* table, int launches, int grid, hipStream_t st, bool zeroGuess) {
    constexpr int E = T + 2 * K, EZ = TZ + 2 * K;
    // as many workgroups as are resident at once: grid / 8 CUs, each holding
    // as many LDS tiles (and E^2-thread workgroups) as fit
    constexpr int byLds = 163840 / (E * E * EZ * 4), byThreads = 2048 / (E * E);
    constexpr int perCu = byLds < byThreads ? byLds : byThreads;
    const int wgs = (grid / 8) * (perCu > 0 ? perCu : 1);
    for (int j = 0; j < launches; ++j) kb_sor_tb<T, K, TZ><<<wgs, E * E, 0, st>>>(table, j, zeroGuess && j == 0);
}

hipError_t launch_batch_step(const BatchTable* table, bool bf16, int sorKernel, int sorLaunches, int grid, hipStream_t st) {
    if (grid < 16 || grid % 16 != 0 || sorLaunches < 0 || sorLaunches > kMaxSorLaunches) return hipErrorInvalidValue;
    const dim3 b = cell_block();
    kb_advance_clock<<<2, 256, 0, st>>>(table);
    kb_advect_velocity<<<grid, b, 0, st>>>(table);
    if (bf16) kb_advect_react<true><<<grid, b, 0, st>>>(table);
    else kb_advect_react<false><<<grid, b, 0, st>>>(table);
    kb_finalize_release<<<kMaxBatch, 256, 0, st>>>(table);
    if (bf16) kb_force_contrib<true><<<grid, b, 0, st>>>(table);
    else kb_force_contrib<false><<<grid, b, 0, st>>>(table);
    // With a tile kernel the pressure launches start from p = 0 without
    // reading p, so the force pass does not store it.
    const bool zeroGuess = sorKernel != 0;
    kb_force_apply<<<grid, b, 0, st>>>(table, zeroGuess ? 0 : 1);
    hipError_t e = launch_batch_pressure(table, sorKernel, sorLaunches, grid, st, zeroGuess);
    if (e != hipSuccess) return e;
    kb_project<<<grid, b, 0, st>>>(table);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_batch_light(table, bf16, grid, st);
}

hipError_t launch_batch_pressure(const BatchTable* table, int sorKernel, int sorLaunches, int grid, hipStream_t st, bool zeroGuess) {
    if (grid < 16 || grid % 16 != 0 || sorLaunches < 0 || sorLaunches > kMaxSorLaunches) return hipErrorInvalidValue;
    if (zeroGuess && sorKernel == 0) return hipErrorInvalidValue;  // the half-sweep kernel updates p in place
    const bool z = zeroGuess;
    switch (sorKernel) {
        case 1: launch_batch_sor_tb<16, 4>(table, sorLaunches, grid, st, z); break;
        case 2: launch_batch_sor_tb<24, 4>(table, sorLaunches, grid, st, z); break;
        case 3: launch_batch_sor_tb<24, 4, 8>(table, sorLaunches, grid, st, z); break;
        case 4: launch_batch_sor_tb<24, 4, 16>(table, sorLaunches, grid, st, z); break;
        case 5: launch_batch_sor_tb<24, 4, 28>(table, sorLaunches, grid, st, z); break;
        case 6: launch_batch_sor_tb<24, 4, 10>(table, sorLaunches, grid, st, z); break;
        case 7: launch_batch_sor_tb<24, 4, 12>(table, sorLaunches, grid, st, z); break;
        case 8: launch_batch_sor_tb<24, 4, 6>(table, sorLaunches, grid, st, z); break;
        case 9: launch_batch_sor_tb<26, 3, 13>(table, sorLaunches, grid, st, z); break;
        case 10: launch_batch_sor_tb<26, 3, 9>(table, sorLaunches, grid, st, z); break;
        case 11: launch_batch_sor_tb<22, 5, 8>(table, sorLaunches, grid, st, z); break;
        default:
            for (int j = 0; j < sorLaunches; ++j) kb_sor_half<<<grid, cell_block(), 0, st>>>(table, j);
    }
    return hipGetLastError();
}

// Eight 256-thread workgroups per CU: enough resident waves to hide memory
// latency in the cell passes (the SOR tile kernels use half as many).
int batch_grid(int device) {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 2048;
    const int g = prop.multiProcessorCount * 8;
    return g < 16 ? 16 : g - g % 16;
}

hipError_t launch_pack_scalar(const DeviceScene& s, int field, const float* src, hipStream_t st) {
    const size_t n = static_cast<size_t>(s.nx) * s.ny * s.nz;
    k_pack_bf16<<<1024, 256, 0, st>>>(static_cast<uint16_t*>(s.s[field]), src, n);
    return hipGetLastError();
}
hipError_t launch_unpack_scalar(const DeviceScene& s, int field, float* dst, hipStream_t st) {
    const size_t n = static_cast<size_t>(s.nx) * s.ny * s.nz;
    k_unpack_bf16<<<1024, 256, 0, st>>>(dst, static_cast<const uint16_t*>(s.s[field]), n);
    return hipGetLastError();
}

}  // namespace fs
