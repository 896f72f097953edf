// Block-wide exclusive prefix count of a per-thread flag (256-thread blocks,
// 64-wide wavefronts): ballot within each wavefront, then the wavefront totals
// through LDS. Every thread of the block must call it.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace bg {
namespace dev {

constexpr int kScanBlock = 256;

__device__ __forceinline__ uint32_t block_exclusive(uint32_t flag, uint32_t* warp_tot, uint32_t* total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long m = __ballot(flag);
    const uint32_t below = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) warp_tot[wave] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t off = 0, t = 0;
    for (int w = 0; w < kScanBlock / 64; ++w) {
        if (w < wave) off += warp_tot[w];
        t += warp_tot[w];
    }
    *total = t;
    return off + below;
}

}  // namespace dev
}  // namespace bg
