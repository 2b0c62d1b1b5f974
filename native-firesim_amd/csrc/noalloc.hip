// noalloc.hip - the pressure pass on caller buffers without any allocation
// (fs_pressure_solve_noalloc in include/firesim.h).
//
// The tile SOR kernels ping-pong between two pressure buffers. A scene keeps
// the second one in its pool; this entry point instead uses `scratch`, a
// file-scope device array of FS_NOALLOC_MAX_CELLS floats that every GPU the
// library runs on holds until the process ends. Its device address is looked
// up once per device (hipGetSymbolAddress with that device current) and only
// that address is handed to kernels: the array's name in host code is a
// host-side stand-in, not device memory.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "scene_impl.hpp"

namespace {

static __device__ __align__(256) float scratch[FS_NOALLOC_MAX_CELLS];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
float* g_scratch[kMaxDevices] = {};  // device address of `scratch` per device, once looked up

// Device address of `scratch` on `device`, looked up on the first call for it.
hipError_t scratch_of(int device, float** out) {
    std::lock_guard<std::mutex> lock(g_mu);
    if (!g_scratch[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!ptr) return hipErrorInvalidSymbol;
        g_scratch[device] = static_cast<float*>(ptr);
    }
    *out = g_scratch[device];
    return hipSuccess;
}

}  // namespace

extern "C" int fs_pressure_solve_noalloc(float* p, const float* d, int32_t nx, int32_t ny, int32_t nz, float omega, int32_t iterations,
                                         int32_t sor_kernel, int32_t device, void* stream) {
    if (!p || !d) return fsi::global_fail(FS_ERR_INVALID_ARGUMENT, "null pressure or divergence buffer");
    if (nx < 4 || ny < 4 || nz < 4) return fsi::global_fail(FS_ERR_INVALID_ARGUMENT, "grid dimensions must each be >= 4");
    const uint64_t cells = static_cast<uint64_t>(nx) * static_cast<uint64_t>(ny) * static_cast<uint64_t>(nz);
    if (cells > FS_NOALLOC_MAX_CELLS)
        return fsi::global_fail(FS_ERR_INVALID_ARGUMENT, "grid too large for the no-allocation scratch: nx*ny*nz must not exceed FS_NOALLOC_MAX_CELLS (2^24)");
    if (!(omega >= 0.05f && omega <= 1.95f) || iterations < 1 || iterations > 1000)
        return fsi::global_fail(FS_ERR_INVALID_ARGUMENT, "omega must be in [0.05, 1.95] and iterations in [1, 1000]");
    if (sor_kernel < 0 || sor_kernel >= FS_SOR_KERNEL_COUNT) return fsi::global_fail(FS_ERR_INVALID_ARGUMENT, "unknown SOR kernel");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fsi::global_fail(FS_ERR_HIP, "no HIP device");
    if (device < 0 || device >= count || device >= kMaxDevices) return fsi::global_fail(FS_ERR_INVALID_ARGUMENT, "device ordinal out of range");
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return fsi::global_fail(FS_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
    float* buf = nullptr;
    e = scratch_of(device, &buf);
    if (e != hipSuccess) return fsi::global_fail(FS_ERR_HIP, std::string("scratch address: ") + hipGetErrorString(e));
    const hipStream_t st = static_cast<hipStream_t>(stream);
    bool inScratch = false;
    e = fs::launch_pressure_buffers(p, buf, d, nx, ny, nz, omega, iterations, sor_kernel, st, &inScratch);
    // an odd number of tile launches leaves the solution in the scratch array
    if (e == hipSuccess && inScratch) e = fs::launch_copy_f32(buf, p, static_cast<size_t>(cells), st);
    if (e != hipSuccess) return fsi::global_fail(FS_ERR_HIP, std::string("pressure solve: ") + hipGetErrorString(e));
    return FS_OK;
}
