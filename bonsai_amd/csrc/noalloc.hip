// Static scratch for the entry point that allocates nothing
// (bg_intersect_device_noalloc): the traversal's LDS-stack spill rows and the
// BG_COUNT counters live in one file-scope device array instead of buffers the
// context allocates on demand.
//
// The array's name is only a host-side stand-in for device code; host code
// uses the per-device address hipGetSymbolAddress returns (taken once per
// device, with that device current, from bg_create).
#include <mutex>

#include "bg_internal.h"

namespace {

constexpr int kMaxDevices = 64;

__device__ __align__(256) unsigned char scratch[BG_NOALLOC_SCRATCH_BYTES];

std::mutex g_mu;
void* g_ptr[kMaxDevices] = {};

__global__ void k_noalloc_zero(unsigned long long* c, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) c[i] = 0ull;
}

// Every query misses (an empty scene, bvh.rs:286, 359).
__global__ void k_noalloc_miss(uint4* hits, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) hits[i] = make_uint4(BG_MISS, 0u, 0u, 0u);
}

}  // namespace

namespace bg {

hipError_t noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_mu);
    if (g_ptr[device]) return hipSuccess;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_ptr[device] = p;
    return hipSuccess;
}

void* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_mu);
    return g_ptr[device];
}

hipError_t noalloc_zero_counters(unsigned long long* c, uint32_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_noalloc_zero, dim3(1), dim3(64), 0, s, c, n);
    return hipGetLastError();
}

hipError_t noalloc_fill_miss(void* hits, uint64_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_noalloc_miss, dim3((uint32_t)blocks), dim3(256), 0, s, (uint4*)hits, n);
    return hipGetLastError();
}

}  // namespace bg
