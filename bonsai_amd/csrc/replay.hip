// Replay ring for the guiding network's online training (no counterpart in
// the reference): after each pass the records with nonzero weight are
// compacted, in record order, onto a ring that minibatches are drawn from.
// Roughly 99.7% of the records of a hard indirect scene carry weight 0 and
// add nothing to the gradient, so training on the compacted ring gives each
// Adam step a full batch of informative samples.
#include <hip/hip_runtime.h>

#include "bg_internal.h"
#include "dev_scan.h"

namespace bg {
namespace {

using dev::block_exclusive;
constexpr int kBlk = dev::kScanBlock;

__global__ __launch_bounds__(kBlk) void k_replay_count(uint64_t n, const float4* __restrict__ tgt, uint32_t* __restrict__ blk) {
    __shared__ uint32_t wt[kBlk / 64];
    const uint64_t i = (uint64_t)blockIdx.x * kBlk + threadIdx.x;
    const uint32_t f = (i < n && tgt[i].w > 0.0f) ? 1u : 0u;
    uint32_t total;
    block_exclusive(f, wt, &total);
    if (threadIdx.x == 0) blk[blockIdx.x] = total;
}

// One block: exclusive scan of the block counts in place (dev::scan_counts), then the ring state.
__global__ __launch_bounds__(1024) void k_replay_scan(uint32_t nblk, uint32_t* __restrict__ blk, uint32_t* __restrict__ state,
                                                      uint32_t cap) {
    __shared__ uint32_t wsum[16];
    const uint32_t run = dev::scan_counts(nblk, blk, wsum);
    if (threadIdx.x == 0) {
        const uint32_t head = state[0];
        state[2] = head;  // base of this append
        state[3] = run;   // size of this append
        state[0] = (uint32_t)(((uint64_t)head + run) % cap);
        state[1] = (uint32_t)min<uint64_t>((uint64_t)state[1] + run, cap);
    }
}

__global__ __launch_bounds__(kBlk) void k_replay_scatter(uint64_t n, const float4* __restrict__ rin,
                                                        const float4* __restrict__ rtgt, const uint32_t* __restrict__ blk,
                                                        const uint32_t* __restrict__ state, uint32_t cap,
                                                        float4* __restrict__ in, float4* __restrict__ tgt) {
    __shared__ uint32_t wt[kBlk / 64];
    const uint64_t i = (uint64_t)blockIdx.x * kBlk + threadIdx.x;
    const float4 t = i < n ? rtgt[i] : make_float4(0, 0, 0, 0);
    const uint32_t f = t.w > 0.0f ? 1u : 0u;
    uint32_t total;
    const uint32_t local = block_exclusive(f, wt, &total);
    if (!f) return;
    const uint32_t g = blk[blockIdx.x] + local;  // index within this append
    const uint32_t size = state[3];
    // when an append exceeds the ring only its last `cap` samples are kept, each slot written once
    if (size > cap && g < size - cap) return;
    const uint32_t slot = (uint32_t)(((uint64_t)state[2] + g) % cap);
    in[3 * (uint64_t)slot] = rin[3 * i];
    in[3 * (uint64_t)slot + 1] = rin[3 * i + 1];
    in[3 * (uint64_t)slot + 2] = rin[3 * i + 2];
    tgt[slot] = t;
}

__global__ __launch_bounds__(256) void k_replay_gather(const float4* __restrict__ rin, const float4* __restrict__ rtgt,
                                                       const uint32_t* __restrict__ state, uint64_t first, uint64_t n,
                                                       float4* __restrict__ in, float4* __restrict__ tgt) {
    const uint64_t k = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const uint64_t count = state[1];
    if (count == 0) return;
    const uint64_t j = ((first + k) % count) * 2654435761ull % count;
    in[3 * k] = rin[3 * j];
    in[3 * k + 1] = rin[3 * j + 1];
    in[3 * k + 2] = rin[3 * j + 2];
    tgt[k] = rtgt[j];
}

}  // namespace

hipError_t launch_replay_append(const PathBuffers& b, uint64_t n_records, const Replay& r, hipStream_t s) {
    if (n_records == 0) return hipSuccess;
    const uint64_t nblk = (n_records + kBlk - 1) / kBlk;
    hipLaunchKernelGGL(k_replay_count, dim3((uint32_t)nblk), dim3(kBlk), 0, s, n_records, b.rec_tgt, b.rec_blk);
    hipLaunchKernelGGL(k_replay_scan, dim3(1), dim3(1024), 0, s, (uint32_t)nblk, b.rec_blk, r.state, r.cap);
    hipLaunchKernelGGL(k_replay_scatter, dim3((uint32_t)nblk), dim3(kBlk), 0, s, n_records, b.rec_in, b.rec_tgt,
                       b.rec_blk, r.state, r.cap, r.in, r.tgt);
    return hipGetLastError();
}

hipError_t launch_replay_gather(const Replay& r, uint64_t first, uint64_t n, float4* in, float4* tgt, hipStream_t s) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_replay_gather, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, r.in, r.tgt, r.state, first,
                       n, in, tgt);
    return hipGetLastError();
}

}  // namespace bg
