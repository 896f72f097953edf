// Host side of the path-guiding MLP: parameter upload, batched evaluation and
// one Adam step per call (kernels in guide.hip). The reference crate has no
// path guiding; these entry points have no counterpart there.
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "bg_internal.h"
#include "../../include/bonsai_scenes.h"

namespace {

int fail(bg_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

#define HIP_TRY(ctx, expr)                                                                      \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail((ctx), BG_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct GuideHostBufs {
    float4* in;
    float4* target;
    float* out;
    float* grad;
};

// Device views of the guide's single allocation.
GuideHostBufs carve(bg_guide* g, size_t* total_out) {
    const uint64_t cap = g->scratch.cap;
    const uint64_t blocks = g->scratch.blocks;
    char* p = (char*)g->mem;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        void* r = p ? p + off : nullptr;
        off += ((bytes + 255) / 256) * 256;
        return r;
    };
    g->dev.wbf = (uint16_t*)take(5 * 4096 * 2);
    g->dev.wbft = (uint16_t*)take(5 * 4096 * 2);
    g->dev.bias = (float*)take(5 * 64 * 4);
    g->dev.master = (float*)take(bg::kParamsTotal * 4);
    g->dev.m = (float*)take(bg::kParamsTotal * 4);
    g->dev.v = (float*)take(bg::kParamsTotal * 4);
    GuideHostBufs b;
    b.grad = (float*)take(bg::kParamsTotal * 4);
    b.in = (float4*)take(48 * cap);
    b.target = (float4*)take(16 * cap);
    b.out = (float*)take(4 * BG_GUIDE_OUTPUTS * cap);
    g->scratch.loss = (float*)take(4 * cap);
    g->scratch.partial = (float*)take(blocks * 5 * 64 * 64 * 4);
    g->scratch.bpartial = (float*)take(blocks * 5 * 64 * 4);
    if (total_out) *total_out = off;
    return b;
}

int grid_of(const bg_guide* g) { return bg::kTrainBlocksPerCu * (g->ctx->num_cus > 0 ? g->ctx->num_cus : 1); }
int infer_grid_of(const bg_guide* g) { return bg::kInferBlocksPerCu * (g->ctx->num_cus > 0 ? g->ctx->num_cus : 1); }

}  // namespace

extern "C" {

int bg_guide_create(bg_ctx* ctx, const float* params, uint64_t max_batch, bg_guide** out) {
    if (!ctx || !params || !out) return fail(ctx, BG_ERR_INVALID, "bg_guide_create: null argument");
    *out = nullptr;
    if (max_batch == 0 || max_batch > (1ull << 26))
        return fail(ctx, BG_ERR_INVALID, "bg_guide_create: max_batch must be in [1, 2^26]");
    bg_guide* g = new (std::nothrow) bg_guide;
    if (!g) return fail(ctx, BG_ERR_NOMEM, "bg_guide_create: out of host memory");
    g->ctx = ctx;
    g->scratch.cap = max_batch;
    g->scratch.blocks = (uint32_t)grid_of(g);
    size_t total = 0;
    carve(g, &total);
    hipError_t e = hipMalloc(&g->mem, total);
    if (e != hipSuccess) {
        delete g;
        return fail(ctx, BG_ERR_NOMEM, std::string("bg_guide_create: hipMalloc: ") + hipGetErrorString(e));
    }
    GuideHostBufs hb = carve(g, nullptr);
    g->batch_in = hb.in;
    g->batch_tgt = hb.target;
    int rc = bg_guide_set_params(g, params);
    if (rc != BG_OK) {
        bg_guide_destroy(g);
        return rc;
    }
    const float lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
    bg_guide_set_bounds(g, lo, hi);
    *out = g;
    return BG_OK;
}

void bg_guide_destroy(bg_guide* g) {
    if (!g) return;
    if (g->mem) hipFree(g->mem);
    if (g->replay_mem) hipFree(g->replay_mem);
    delete g;
}

int bg_guide_set_params(bg_guide* g, const float* params) {
    if (!g || !params) return BG_ERR_INVALID;
    bg_ctx* ctx = g->ctx;
    hipStream_t s = ctx->stream;
    HIP_TRY(ctx, hipMemcpyAsync(g->dev.master, params, bg::kParamsTotal * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemsetAsync(g->dev.m, 0, bg::kParamsTotal * 4, s));
    HIP_TRY(ctx, hipMemsetAsync(g->dev.v, 0, bg::kParamsTotal * 4, s));
    HIP_TRY(ctx, bg::guide_refresh(g->dev, s));
    if (g->replay.state) HIP_TRY(ctx, hipMemsetAsync(g->replay.state, 0, 16, s));  // training starts afresh
    HIP_TRY(ctx, hipStreamSynchronize(s));
    g->step = 0;
    return BG_OK;
}

int bg_guide_get_params(bg_guide* g, float* params, float* adam_m, float* adam_v) {
    if (!g) return BG_ERR_INVALID;
    bg_ctx* ctx = g->ctx;
    hipStream_t s = ctx->stream;
    if (params) HIP_TRY(ctx, hipMemcpyAsync(params, g->dev.master, bg::kParamsTotal * 4, hipMemcpyDeviceToHost, s));
    if (adam_m) HIP_TRY(ctx, hipMemcpyAsync(adam_m, g->dev.m, bg::kParamsTotal * 4, hipMemcpyDeviceToHost, s));
    if (adam_v) HIP_TRY(ctx, hipMemcpyAsync(adam_v, g->dev.v, bg::kParamsTotal * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    return BG_OK;
}

int bg_guide_set_bounds(bg_guide* g, const float lo[3], const float hi[3]) {
    if (!g || !lo || !hi) return BG_ERR_INVALID;
    for (int k = 0; k < 3; ++k)
        if (!(hi[k] > lo[k])) return fail(g->ctx, BG_ERR_INVALID, "bg_guide_set_bounds: empty box");
    for (int k = 0; k < 3; ++k) {
        g->box.lo[k] = lo[k];
        g->box.hi[k] = hi[k];
    }
    return BG_OK;
}

int bg_guide_init_params(uint64_t seed, float* params) {
    if (!params) return BG_ERR_INVALID;
    // layers 64 -> 64 (x4) and 64 -> 40: weights (out x in, row-major), then biases
    const int outs[5] = {64, 64, 64, 64, 40};
    uint64_t k = 0;
    for (int l = 0; l < 5; ++l) {
        const float lim = std::sqrt(6.0f / 64.0f);
        for (int i = 0; i < outs[l] * 64; ++i, ++k) params[k] = (2.0f * bgx_uniform(seed, k) - 1.0f) * lim;
        for (int i = 0; i < outs[l]; ++i, ++k) params[k] = 0.0f;
    }
    return k == bg::kParamsTotal ? BG_OK : BG_ERR_INVALID;
}

namespace {
// Snapshot layout: header, then master / m / v (BG_GUIDE_PARAMS floats each), then the replay
// ring's cursor (4 words) and contents (48 + 16 bytes per sample of its capacity).
struct SnapshotHeader {
    char magic[8];
    uint64_t version;
    uint64_t step;
    float lo[3], hi[3];
    uint64_t replay_cap;
    uint64_t bytes;
};
constexpr char kSnapMagic[8] = {'B', 'G', 'G', 'U', 'I', 'D', 'E', '1'};
uint64_t snapshot_bytes(uint64_t replay_cap) {
    return sizeof(SnapshotHeader) + 3ull * bg::kParamsTotal * 4 + (replay_cap ? 16 + 64ull * replay_cap : 0);
}
}  // namespace

int bg_guide_snapshot(bg_guide* g, void* buf, uint64_t* size) {
    if (!g || !size) return BG_ERR_INVALID;
    bg_ctx* ctx = g->ctx;
    const uint64_t cap = g->replay.state ? g->replay.cap : 0;
    const uint64_t need = snapshot_bytes(cap);
    if (!buf) {
        *size = need;
        return BG_OK;
    }
    if (*size < need) return fail(ctx, BG_ERR_INVALID, "bg_guide_snapshot: buffer too small");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    char* p = static_cast<char*>(buf);
    SnapshotHeader h{};
    std::memcpy(h.magic, kSnapMagic, 8);
    h.version = 1;
    h.step = (uint64_t)g->step;
    for (int k = 0; k < 3; ++k) {
        h.lo[k] = g->box.lo[k];
        h.hi[k] = g->box.hi[k];
    }
    h.replay_cap = cap;
    h.bytes = need;
    std::memcpy(p, &h, sizeof(h));
    char* q = p + sizeof(h);
    const size_t pb = bg::kParamsTotal * 4;
    HIP_TRY(ctx, hipMemcpyAsync(q, g->dev.master, pb, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(q + pb, g->dev.m, pb, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(q + 2 * pb, g->dev.v, pb, hipMemcpyDeviceToHost, s));
    q += 3 * pb;
    if (cap) {
        HIP_TRY(ctx, hipMemcpyAsync(q, g->replay.state, 16, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(q + 16, g->replay.in, 48 * cap, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(q + 16 + 48 * cap, g->replay.tgt, 16 * cap, hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(ctx, hipStreamSynchronize(s));
    *size = need;
    return BG_OK;
}

int bg_guide_restore(bg_guide* g, const void* buf, uint64_t size) {
    if (!g || !buf) return BG_ERR_INVALID;
    bg_ctx* ctx = g->ctx;
    SnapshotHeader h;
    if (size < sizeof(h)) return fail(ctx, BG_ERR_INVALID, "bg_guide_restore: not a guide snapshot");
    std::memcpy(&h, buf, sizeof(h));
    if (std::memcmp(h.magic, kSnapMagic, 8) != 0 || h.version != 1 || h.replay_cap > 0xFFFFFFFFull ||
        h.bytes != snapshot_bytes(h.replay_cap) || size < h.bytes)
        return fail(ctx, BG_ERR_INVALID, "bg_guide_restore: not a guide snapshot of this version");
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const char* q = static_cast<const char*>(buf) + sizeof(h);
    const size_t pb = bg::kParamsTotal * 4;
    HIP_TRY(ctx, hipMemcpyAsync(g->dev.master, q, pb, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(g->dev.m, q + pb, pb, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(g->dev.v, q + 2 * pb, pb, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, bg::guide_refresh(g->dev, s));
    q += 3 * pb;
    if (h.replay_cap) {
        const int rc = bg::guide_ensure_replay(g, (uint32_t)h.replay_cap);
        if (rc) return rc;
        HIP_TRY(ctx, hipMemcpyAsync(g->replay.state, q, 16, hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(g->replay.in, q + 16, 48 * h.replay_cap, hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(g->replay.tgt, q + 16 + 48 * h.replay_cap, 16 * h.replay_cap,
                                    hipMemcpyHostToDevice, s));
    } else if (g->replay.state) {
        HIP_TRY(ctx, hipMemsetAsync(g->replay.state, 0, 16, s));
    }
    HIP_TRY(ctx, hipStreamSynchronize(s));
    g->step = (int)h.step;
    for (int k = 0; k < 3; ++k) {
        g->box.lo[k] = h.lo[k];
        g->box.hi[k] = h.hi[k];
    }
    return BG_OK;
}

int bg_guide_eval(bg_guide* g, const float* samples, uint64_t n, float* out) {
    if (!g || (n && (!samples || !out))) return BG_ERR_INVALID;
    bg_ctx* ctx = g->ctx;
    if (n > g->scratch.cap) return fail(ctx, BG_ERR_INVALID, "bg_guide_eval: n exceeds max_batch");
    if (n == 0) return BG_OK;
    GuideHostBufs b = carve(g, nullptr);
    hipStream_t s = ctx->stream;
    HIP_TRY(ctx, hipMemcpyAsync(b.in, samples, 48 * n, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, bg::guide_infer(g->dev, g->box, b.in, n, b.out, infer_grid_of(g), s));
    HIP_TRY(ctx, hipMemcpyAsync(out, b.out, 4 * BG_GUIDE_OUTPUTS * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    return BG_OK;
}

int bg_guide_train(bg_guide* g, const float* samples, const float* targets, uint64_t n, float lr,
                   double* mean_loss, float* grad) {
    if (!g || (n && (!samples || !targets))) return BG_ERR_INVALID;
    bg_ctx* ctx = g->ctx;
    if (n > g->scratch.cap) return fail(ctx, BG_ERR_INVALID, "bg_guide_train: n exceeds max_batch");
    if (n == 0) return BG_OK;
    GuideHostBufs b = carve(g, nullptr);
    hipStream_t s = ctx->stream;
    HIP_TRY(ctx, hipMemcpyAsync(b.in, samples, 48 * n, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(b.target, targets, 16 * n, hipMemcpyHostToDevice, s));
    ++g->step;
    HIP_TRY(ctx, bg::guide_train_step(g->dev, g->box, b.in, b.target, n, g->scratch, lr, g->step, grid_of(g),
                                      grad ? b.grad : nullptr, s));
    std::vector<float> loss;
    if (mean_loss) {
        loss.resize(n);
        HIP_TRY(ctx, hipMemcpyAsync(loss.data(), g->scratch.loss, 4 * n, hipMemcpyDeviceToHost, s));
    }
    if (grad) HIP_TRY(ctx, hipMemcpyAsync(grad, b.grad, bg::kParamsTotal * 4, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    if (mean_loss) {
        double sum = 0.0;
        for (float v : loss) sum += v;
        *mean_loss = sum / (double)n;
    }
    return BG_OK;
}

int bg_guide_eval_device(bg_guide* g, const void* d_samples, uint64_t n, void* d_out, double* kernel_ms) {
    if (!g || (n && (!d_samples || !d_out))) return BG_ERR_INVALID;
    bg_ctx* ctx = g->ctx;
    hipStream_t s = ctx->stream;
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, s));
    HIP_TRY(ctx, bg::guide_infer(g->dev, g->box, (const float4*)d_samples, n, (float*)d_out, infer_grid_of(g), s));
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, s));
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (kernel_ms) *kernel_ms = ms;
    return BG_OK;
}

int bg_guide_train_device(bg_guide* g, const void* d_samples, const void* d_targets, uint64_t n, float lr,
                          double* kernel_ms) {
    if (!g || (n && (!d_samples || !d_targets))) return BG_ERR_INVALID;
    bg_ctx* ctx = g->ctx;
    if (n > g->scratch.cap) return fail(ctx, BG_ERR_INVALID, "bg_guide_train_device: n exceeds max_batch");
    hipStream_t s = ctx->stream;
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, s));
    if (n) {
        ++g->step;
        HIP_TRY(ctx, bg::guide_train_step(g->dev, g->box, (const float4*)d_samples, (const float4*)d_targets, n,
                                          g->scratch, lr, g->step, grid_of(g), nullptr, s));
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, s));
    HIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
    float ms = 0.0f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    if (kernel_ms) *kernel_ms = ms;
    return BG_OK;
}

}  // extern "C"
