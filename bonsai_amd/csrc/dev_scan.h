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

// Exclusive prefix of nblk counts in place, by one block of 1024 threads; returns their sum to
// every thread. Each thread sums a run of consecutive counts (loads issued eight at a time), a
// shuffle scan within each wavefront and one over the 16 wavefront totals (wsum, 16 words of
// LDS) give each run's offset, and each thread then rewrites its run.
__device__ __forceinline__ uint32_t scan_counts(uint32_t nblk, uint32_t* __restrict__ blk, uint32_t* wsum) {
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t per = (nblk + 1023) / 1024;
    const uint32_t b0 = min(t * per, nblk), b1 = min(b0 + per, nblk);
    uint32_t sum = 0, k = b0;
    for (; k + 8 <= b1; k += 8) {
        uint32_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = blk[k + j];
#pragma unroll
        for (int j = 0; j < 8; ++j) sum += v[j];
    }
    for (; k < b1; ++k) sum += blk[k];
    uint32_t x = sum;  // inclusive scan over the wavefront
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= (uint32_t)d) x += y;
    }
    if (lane == 63) wsum[wave] = x;
    __syncthreads();
    uint32_t run = x - sum, all = 0;
    for (uint32_t w = 0; w < 16; ++w) {
        const uint32_t v = wsum[w];
        if (w < wave) run += v;
        all += v;
    }
    for (k = b0; k + 8 <= b1; k += 8) {
        uint32_t v[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = blk[k + j];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            blk[k + j] = run;
            run += v[j];
        }
    }
    for (; k < b1; ++k) {
        const uint32_t v = blk[k];
        blk[k] = run;
        run += v;
    }
    return all;
}

}  // namespace dev
}  // namespace bg
