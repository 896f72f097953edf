// Path-guiding MLP on gfx950 matrix cores (no counterpart in the reference).
//
// Network: 64-dim encoding of (position, outgoing direction, normal) -> 4
// hidden layers of 64 ReLU units -> 40 outputs = 8 von Mises-Fisher lobes x
// (mixture logit, direction 3-vector, concentration parameter s with log kappa
// = 1.5 + 5.5 tanh((s - 1.5) / 5.5)). Weights are kept in
// FP32 (master copy, Adam moments) and in BF16 for the matrix cores; every
// product is v_mfma_f32_32x32x8_bf16 with FP32 accumulation.
//
// One wavefront evaluates 32 samples at a time as Y^T = W X^T: the weights
// are the A operand (32 output rows x 8 inputs per instruction) and the
// activations the B operand (8 inputs x 32 samples). The accumulator layout
// (lane l, register r -> row 8*(r/4) + 4*(l/32) + r%4, column l%32) is
// exactly the B-operand layout of the next layer's k-steps, so activations
// stay in registers from layer to layer; only the bias, ReLU and the BF16
// rounding touch them in between. The backward pass reuses the same scheme
// with the transposed weights. Weight gradients (dW = dY X^T summed over the
// batch) are split over sample chunks and reduced in the fused Adam kernel.
#include <hip/hip_runtime.h>

#include "bg_internal.h"
#include "dev_guide.h"

namespace bg {
namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf4v __attribute__((ext_vector_type(4)));

constexpr int kOut = 40, kLayers = 5;
// log kappa = C + R tanh((s - C) / R): kappa in (e^-4, e^7) with a gradient everywhere
constexpr float kLogKappaC = 1.5f, kLogKappaR = 5.5f;
constexpr float kPiF = 3.14159265358979323846f;

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x7FFFFFu)) return 0x7FC0;
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// Four floats rounded to BF16 (nearest even) with gfx950's packed conversion
// (v_cvt_pk_bf16_f32, two values per instruction); finite values round exactly
// as f2bf does.
__device__ __forceinline__ s4 pack4(float a, float b, float c, float d) {
    const bf4v v = __builtin_convertvector((f4v){a, b, c, d}, bf4v);
    return __builtin_bit_cast(s4, v);
}

// Encoding features of a sample (oracle/guide_mlp.py encode()): 0-47 sin and
// cos of 2^k pi q_axis (axis = f / 16, k = (f % 16) / 2), 48-50 wo, 51-53 n,
// 54-59 sin and cos of pi wo_i, 60-63 zero. sincospif takes the argument in
// half-turns, so 2^k q is reduced exactly.
// Weights in LDS: BF16 [layer][64 rows][kLd], rows padded from 64 to 68
// elements (136 B = 34 banks) so the 32 rows one A-operand read touches spread
// over the banks instead of all starting on bank 0. A operand of k-step s for
// output block mb: row 32*mb + lane%32, columns 8s + 4*(lane/32) .. +3.
constexpr int kLd = 68;
constexpr int kLayerLds = 64 * kLd;
__device__ __forceinline__ s4 load_a(const uint16_t* __restrict__ w, int mb, int s, int lane) {
    const uint16_t* p = w + (32 * mb + (lane & 31)) * kLd + 8 * s + 4 * (lane >> 5);
    return *reinterpret_cast<const s4*>(p);
}

// One layer: acc[mb] = W[32mb.., :] * X^T, with X^T given as B operands xb[s].
__device__ __forceinline__ void layer_mm(const uint16_t* __restrict__ w, const s4 xb[8], int lane, f16v acc[2]) {
    for (int mb = 0; mb < 2; ++mb) {
        f16v c = {0};
#pragma unroll
        for (int s = 0; s < 8; ++s) c = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(load_a(w, mb, s, lane), xb[s], c, 0, 0, 0);
        acc[mb] = c;
    }
}

// Row held by accumulator register r of lane `lane` in block mb.
__device__ __forceinline__ int acc_row(int mb, int r, int lane) { return 32 * mb + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3); }

// Output row of the last layer that accumulator row p holds in the inference
// kernel: lanes 0-31 of a wavefront hold accumulator rows 8q + j and 32 + j
// (q, j < 4), lanes 32-63 those rows + 4; ordering the outputs so that the first
// set is outputs 0-19 (lobes 0-3) and the second outputs 20-39 (lobes 4-7) gives
// each half-wavefront four whole lobes.
__device__ __forceinline__ int out_row(int p) {
    if (p >= 40) return p;
    const int h = (p >> 2) & 1, idx = p < 32 ? 4 * (p >> 3) + (p & 3) : 16 + (p & 3);
    return 20 * h + idx;
}

// Bias + optional ReLU, then repack the accumulators as the next layer's B operands.
// permute_out: the accumulator rows hold the outputs out_row(row) (last layer, inference).
__device__ __forceinline__ void epilogue(f16v acc[2], const float* __restrict__ bias, bool relu, int lane, s4 xb[8],
                                         bool permute_out = false) {
    for (int mb = 0; mb < 2; ++mb) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = acc_row(mb, r, lane);
            float v = acc[mb][r] + bias[permute_out ? out_row(row) : row];
            if (relu) v = v > 0.0f ? v : 0.0f;
            acc[mb][r] = v;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            xb[4 * mb + q] = pack4(acc[mb][4 * q], acc[mb][4 * q + 1], acc[mb][4 * q + 2], acc[mb][4 * q + 3]);
    }
}

// B operands of the encoding for the sample of this lane: k-step s holds
// features 8s + 4h .. 8s + 4h + 3, h = lane / 32. On each axis a lane needs
// the frequencies b, b + 1, b + 4, b + 5 (b = 2h): one sincospif at 2^b and
// double-angle steps (sin 2x = 2 sin x cos x, cos 2x = 1 - 2 sin^2 x) for the
// rest; five doublings cost a few ulps, far below the BF16 rounding that follows.
__device__ __forceinline__ void dbl(float& sn, float& cs) {
    const float s2 = 2.0f * sn * cs;
    cs = 1.0f - 2.0f * sn * sn;
    sn = s2;
}
__device__ __forceinline__ void encode_b(const float q[3], const float wo[3], const float n[3], int lane, s4 xb[8]) {
    const int h = lane >> 5;
#pragma unroll
    for (int axis = 0; axis < 3; ++axis) {
        float s0, c0;
        sincospif((float)(1 << (2 * h)) * q[axis], &s0, &c0);  // frequency 2h
        float s1 = s0, c1 = c0;
        dbl(s1, c1);                                            // 2h + 1
        float s4v = s1, c4v = c1;
        dbl(s4v, c4v);
        dbl(s4v, c4v);
        dbl(s4v, c4v);                                          // 2h + 4
        float s5 = s4v, c5 = c4v;
        dbl(s5, c5);                                            // 2h + 5
        xb[2 * axis] = pack4(s0, c0, s1, c1);
        xb[2 * axis + 1] = pack4(s4v, c4v, s5, c5);
    }
    float sw0, cw0, sw1, cw1, sw2, cw2;
    sincospif(wo[0], &sw0, &cw0);
    sincospif(wo[1], &sw1, &cw1);
    sincospif(wo[2], &sw2, &cw2);
    xb[6] = h == 0 ? pack4(wo[0], wo[1], wo[2], n[0]) : pack4(n[1], n[2], sw0, cw0);
    xb[7] = h == 0 ? pack4(sw1, cw1, sw2, cw2) : pack4(0.0f, 0.0f, 0.0f, 0.0f);
}

struct Sample {
    float p[3], wo[3], n[3];
};
__device__ __forceinline__ Sample load_sample(const float4* __restrict__ in, uint64_t i) {
    const float4 a = in[3 * i], b = in[3 * i + 1], c = in[3 * i + 2];
    return Sample{{a.x, a.y, a.z}, {b.x, b.y, b.z}, {c.x, c.y, c.z}};
}

__device__ __forceinline__ void normalise_pos(const Sample& s, const GuideBox& box, float q[3]) {
    for (int k = 0; k < 3; ++k) q[k] = (s.p[k] - box.lo[k]) / (box.hi[k] - box.lo[k]);
}

// Copies the 5 BF16 weight matrices (dense rows of 64) into padded LDS rows,
// 8 bytes per thread and step; with permute_out the last layer's accumulator
// row p takes output row out_row(p).
__device__ void load_weights_lds(const uint16_t* __restrict__ wg, uint16_t* __restrict__ lds, bool permute_out = false) {
    const uint2* src = reinterpret_cast<const uint2*>(wg);
    for (int i = threadIdx.x; i < kLayers * 64 * 16; i += blockDim.x) {
        const int row = i / 16, q = i % 16;  // row over all layers, 4-element quad in the row
        const int from = permute_out && row >= (kLayers - 1) * 64 ? (kLayers - 1) * 64 + out_row(row % 64) : row;
        *reinterpret_cast<uint2*>(lds + row * kLd + 4 * q) = src[from * 16 + q];
    }
    __syncthreads();
}

// ------------------------------------------------------------- inference
// in: 3 float4 per sample (p, wo, n); out: 40 floats per sample.
// MIX: write the 10-float4 mixture record of each sample (dev_guide.h) instead
// of its 40 raw outputs.
// Held to 3 waves per SIMD (156 VGPRs, the accumulators in VGPRs): left alone
// the compiler settles at 2.
template <bool MIX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 8))) void k_guide_infer(GuideDevice g, GuideBox box, const float4* __restrict__ in,
                                                     uint64_t n, const uint32_t* __restrict__ n_dev,
                                                     float* __restrict__ out) {
    __shared__ uint16_t wl[kLayers * kLayerLds];
    if (n_dev) n = *n_dev;
    if ((uint64_t)blockIdx.x * 128 >= n) return;  // the whole block is idle (uniform)
    load_weights_lds(g.wbf, wl, true);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, h = lane >> 5;
    const uint64_t groups = (n + 31) / 32;
    for (uint64_t grp = (uint64_t)blockIdx.x * 4 + wave; grp < groups; grp += (uint64_t)gridDim.x * 4) {
        const uint64_t i = grp * 32 + (lane & 31);
        const bool valid = i < n;
        Sample smp = valid ? load_sample(in, i) : Sample{{0, 0, 0}, {0, 0, 1}, {0, 0, 1}};
        float q[3];
        normalise_pos(smp, box, q);
        s4 xb[8];
        encode_b(q, smp.wo, smp.n, lane, xb);
        f16v acc[2];
        for (int l = 0; l < kLayers; ++l) {
            layer_mm(wl + l * kLayerLds, xb, lane, acc);
            epilogue(acc, g.bias + l * 64, l < kLayers - 1, lane, xb, l == kLayers - 1);
        }
        // this lane holds outputs 20h .. 20h + 19 of sample l % 32 (lobes 4h .. 4h + 3): accumulators
        // 0-15 of block 0, then 0-3 of block 1 (out_row)
        float o[20];
#pragma unroll
        for (int k = 0; k < 16; ++k) o[k] = acc[0][k];
#pragma unroll
        for (int k = 0; k < 4; ++k) o[16 + k] = acc[1][k];
        if (MIX) {
            dev::guide_mixture_half(o, h, valid, reinterpret_cast<float4*>(out) + dev::kMixRecord * i);
        } else if (valid) {
            float4* dst = reinterpret_cast<float4*>(out + i * kOut + 20 * h);
#pragma unroll
            for (int qq = 0; qq < 5; ++qq) dst[qq] = make_float4(o[4 * qq], o[4 * qq + 1], o[4 * qq + 2], o[4 * qq + 3]);
        }
    }
}

// ------------------------------------------------------------- mixture math
struct Lobe8 {
    float pi[8], mu[8][3], kappa[8], nm[8], s[8];  // s: tanh((s_raw - C) / R)
};
__device__ __forceinline__ void mixture_of(const float* o, Lobe8& L) {
    float amax = -3.0e38f;
    for (int j = 0; j < 8; ++j) amax = fmaxf(amax, o[5 * j]);
    float sum = 0.0f;
    for (int j = 0; j < 8; ++j) { L.pi[j] = expf(o[5 * j] - amax); sum += L.pi[j]; }
    for (int j = 0; j < 8; ++j) {
        L.pi[j] /= sum;
        const float mx = o[5 * j + 1], my = o[5 * j + 2], mz = o[5 * j + 3];
        const float nm = sqrtf(mx * mx + my * my + mz * mz + 1e-12f);
        L.nm[j] = nm;
        L.mu[j][0] = mx / nm; L.mu[j][1] = my / nm; L.mu[j][2] = mz / nm;
        L.s[j] = tanhf((o[5 * j + 4] - kLogKappaC) / kLogKappaR);  // kept for the chain rule
        L.kappa[j] = expf(kLogKappaC + kLogKappaR * L.s[j]);
    }
}
__device__ __forceinline__ float vmf(const float* mu, float kappa, const float* w) {
    const float c = kappa / (2.0f * kPiF * (-expm1f(-2.0f * kappa)));
    return c * expf(kappa * (mu[0] * w[0] + mu[1] * w[1] + mu[2] * w[2] - 1.0f));
}

// d(-weight * log p(w)) / d out, scaled by inv_batch (oracle loss_grad_out).
__device__ void out_grad(const float* o, const float* w, float weight, float inv_batch, float* g, float* loss) {
    if (weight == 0.0f) {  // empty record: no loss, no gradient (even where the mixture underflows)
        for (int k = 0; k < kOut; ++k) g[k] = 0.0f;
        *loss = 0.0f;
        return;
    }
    Lobe8 L;
    mixture_of(o, L);
    float f[8], p = 0.0f;
    for (int j = 0; j < 8; ++j) { f[j] = vmf(L.mu[j], L.kappa[j], w); p += L.pi[j] * f[j]; }
    p = fmaxf(p, 1e-30f);  // a direction where every lobe underflows in FP32
    *loss = -weight * logf(p);
    const float sc = -weight * inv_batch;
    for (int j = 0; j < 8; ++j) {
        const float r = L.pi[j] * f[j] / p;
        g[5 * j] = sc * (r - L.pi[j]);
        const float k = L.kappa[j];
        const float muw = L.mu[j][0] * w[0] + L.mu[j][1] * w[1] + L.mu[j][2] * w[2];
        const float dlogk = 1.0f / k - 2.0f * expf(-2.0f * k) / (-expm1f(-2.0f * k)) + (muw - 1.0f);
        g[5 * j + 4] = sc * r * dlogk * k * (1.0f - L.s[j] * L.s[j]);
        for (int c = 0; c < 3; ++c) {
            const float proj = k * w[c] - k * muw * L.mu[j][c];
            g[5 * j + 1 + c] = sc * r * proj / L.nm[j];
        }
    }
}

// ------------------------------------------------------------- training
// Forward + loss gradient + backward for a batch. Stores every layer's input
// activations X_l and output gradients dY_l as BF16, feature-major [64][S],
// for the weight-gradient kernel. Per-sample losses go to loss_out.
__global__ __launch_bounds__(256) void k_guide_train(GuideDevice g, GuideBox box, const float4* __restrict__ in,
                                                     const float4* __restrict__ target, uint64_t n, float inv_batch,
                                                     uint16_t* __restrict__ xs, uint16_t* __restrict__ dys,
                                                     float* __restrict__ loss_out) {
    __shared__ uint16_t wl[kLayers * kLayerLds];
    __shared__ uint16_t wt[kLayers * kLayerLds];
    __shared__ float tile[4][32][65];
    load_weights_lds(g.wbf, wl);
    load_weights_lds(g.wbft, wt);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t groups = (n + 31) / 32;
    // every wave of the block runs the same number of rounds (the rounds hold block barriers)
    for (uint64_t base = (uint64_t)blockIdx.x * 4; base < groups; base += (uint64_t)gridDim.x * 4) {
        const uint64_t i = (base + wave) * 32 + (lane & 31);
        const bool valid = i < n;
        Sample smp = valid ? load_sample(in, i) : Sample{{0, 0, 0}, {0, 0, 1}, {0, 0, 1}};
        float q[3];
        normalise_pos(smp, box, q);
        s4 xb[8];
        encode_b(q, smp.wo, smp.n, lane, xb);
        f16v acc[2];
        for (int l = 0; l < kLayers; ++l) {
            // store X_l (B operand, feature-major) for dW
            if (valid)
                for (int s = 0; s < 8; ++s) {
                    const int f = 8 * s + 4 * (lane >> 5);
                    for (int j = 0; j < 4; ++j) xs[((uint64_t)l * 64 + f + j) * n + i] = (uint16_t)xb[s][j];
                }
            layer_mm(wl + l * kLayerLds, xb, lane, acc);
            epilogue(acc, g.bias + l * 64, l < kLayers - 1, lane, xb);
        }
        // gather the 40 outputs of each sample, form d loss / d out
        for (int mb = 0; mb < 2; ++mb)
            for (int r = 0; r < 16; ++r) tile[wave][lane & 31][acc_row(mb, r, lane)] = acc[mb][r];
        __syncthreads();
        if (lane < 32) {
            float o[kOut], gr[kOut], ls = 0.0f;
            for (int k = 0; k < kOut; ++k) o[k] = tile[wave][lane][k];
            const float4 tg = valid ? target[i] : make_float4(0, 0, 1, 0);
            const float w[3] = {tg.x, tg.y, tg.z};
            out_grad(o, w, valid ? tg.w : 0.0f, inv_batch, gr, &ls);
            for (int k = 0; k < 64; ++k) tile[wave][lane][k] = k < kOut ? gr[k] : 0.0f;
            if (valid) loss_out[i] = ls;
        }
        __syncthreads();
        // backward: d of layer-l input = W_l^T dY_l, masked by the ReLU of that input
        s4 db[8];
        for (int s = 0; s < 8; ++s) {
            const int f = 8 * s + 4 * (lane >> 5);
            db[s] = pack4(tile[wave][lane & 31][f], tile[wave][lane & 31][f + 1], tile[wave][lane & 31][f + 2],
                          tile[wave][lane & 31][f + 3]);
        }
        __syncthreads();
        for (int l = kLayers - 1; l >= 0; --l) {
            if (valid)
                for (int s = 0; s < 8; ++s) {
                    const int f = 8 * s + 4 * (lane >> 5);
                    for (int j = 0; j < 4; ++j) dys[((uint64_t)l * 64 + f + j) * n + i] = (uint16_t)db[s][j];
                }
            if (l == 0) break;
            layer_mm(wt + l * kLayerLds, db, lane, acc);
            // mask by ReLU'(pre-activation of layer l-1) = (X_l > 0)
            for (int mb = 0; mb < 2; ++mb)
                for (int r = 0; r < 16; ++r) {
                    const int row = acc_row(mb, r, lane);
                    const uint16_t xv = valid ? xs[((uint64_t)l * 64 + row) * n + i] : 0;
                    if (!(bf2f(xv) > 0.0f)) acc[mb][r] = 0.0f;
                }
            for (int mb = 0; mb < 2; ++mb)
                for (int qq = 0; qq < 4; ++qq)
                    db[4 * mb + qq] = pack4(acc[mb][4 * qq], acc[mb][4 * qq + 1], acc[mb][4 * qq + 2], acc[mb][4 * qq + 3]);
        }
    }
}

// Weight gradients: for layer l, output block mb, input block nb and sample
// chunk c, partial[c][l][32mb + row][32nb + col] = sum over the chunk of
// dY_l[row][s] * X_l[col][s]; bias partials = sum of dY_l over the chunk.
__global__ __launch_bounds__(64) void k_guide_dw(const uint16_t* __restrict__ xs, const uint16_t* __restrict__ dys,
                                                 uint64_t n, uint64_t chunk, float* __restrict__ partial,
                                                 float* __restrict__ bpartial) {
    const int lane = threadIdx.x;
    const int l = blockIdx.y / 4, mb = (blockIdx.y / 2) & 1, nb = blockIdx.y & 1;
    const uint64_t c = blockIdx.x;
    const uint64_t s0 = c * chunk, s1 = s0 + chunk < n ? s0 + chunk : n;
    f16v acc = {0};
    float bsum = 0.0f;
    const uint64_t rowA = (uint64_t)l * 64 + 32 * mb + (lane & 31);
    const uint64_t rowB = (uint64_t)l * 64 + 32 * nb + (lane & 31);
    for (uint64_t s = s0; s < s1; s += 8) {
        s4 a, b;
        for (int j = 0; j < 4; ++j) {
            const uint64_t k = s + 4 * (lane >> 5) + j;
            a[j] = (short)(k < s1 ? dys[rowA * n + k] : 0);
            b[j] = (short)(k < s1 ? xs[rowB * n + k] : 0);
            if (nb == 0) bsum += bf2f((uint16_t)a[j]);
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
    }
    float* dst = partial + ((c * kLayers + l) * 64) * 64;
    for (int r = 0; r < 16; ++r) dst[(uint64_t)(32 * mb + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)) * 64 + 32 * nb + (lane & 31)] = acc[r];
    if (nb == 0) {
        // lanes l and l+32 hold halves of the same row's sum
        const float other = __shfl_down(bsum, 32);
        if (lane < 32) bpartial[(c * kLayers + l) * 64 + 32 * mb + lane] = bsum + other;
    }
}

// Sums the partials into the gradient of every parameter (oracle flat order),
// takes one Adam step on the FP32 master weights and refreshes the BF16 copies.
__global__ __launch_bounds__(256) void k_guide_adam(GuideDevice g, const float* __restrict__ partial,
                                                    const float* __restrict__ bpartial, uint64_t chunks, float lr,
                                                    float b1, float b2, float eps, int t, float* __restrict__ grad_out) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= kParamsTotal) return;
    // locate (layer, row, col) of flat index idx: per layer W (rows x cols) then b (rows)
    int l = 0, base = 0;
    int rows = 64, cols = 64;
    for (;;) {
        rows = l == kLayers - 1 ? kOut : 64;
        cols = 64;
        const int sz = rows * cols + rows;
        if (idx < base + sz) break;
        base += sz;
        ++l;
    }
    const int off = idx - base;
    float gsum = 0.0f;
    int row, col = -1;
    if (off < rows * cols) {
        row = off / cols;
        col = off % cols;
        for (uint64_t c = 0; c < chunks; ++c) gsum += partial[((c * kLayers + l) * 64 + row) * 64 + col];
    } else {
        row = off - rows * cols;
        for (uint64_t c = 0; c < chunks; ++c) gsum += bpartial[(c * kLayers + l) * 64 + row];
    }
    if (grad_out) grad_out[idx] = gsum;
    // Adam (Kingma & Ba 2015) with bias correction
    const float m = b1 * g.m[idx] + (1.0f - b1) * gsum;
    const float v = b2 * g.v[idx] + (1.0f - b2) * gsum * gsum;
    g.m[idx] = m;
    g.v[idx] = v;
    const float mh = m / (1.0f - powf(b1, (float)t));
    const float vh = v / (1.0f - powf(b2, (float)t));
    const float w = g.master[idx] - lr * mh / (sqrtf(vh) + eps);
    g.master[idx] = w;
    if (col >= 0) {
        g.wbf[l * 4096 + row * 64 + col] = f2bf(w);
        g.wbft[l * 4096 + col * 64 + row] = f2bf(w);
    } else {
        g.bias[l * 64 + row] = w;
    }
}

// BF16 copies and biases from the FP32 master parameters.
__global__ __launch_bounds__(256) void k_guide_refresh(GuideDevice g) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= kLayers * 64 * 64) return;
    const int l = idx / 4096, row = (idx / 64) % 64, col = idx % 64;
    const int rows = l == kLayers - 1 ? kOut : 64;
    int base = 0;
    for (int k = 0; k < l; ++k) base += 64 * 64 + 64;
    const float w = row < rows ? g.master[base + row * 64 + col] : 0.0f;
    g.wbf[l * 4096 + row * 64 + col] = f2bf(w);
    g.wbft[l * 4096 + col * 64 + row] = f2bf(w);
    if (col == 0) g.bias[l * 64 + row] = row < rows ? g.master[base + rows * 64 + row] : 0.0f;
}

}  // namespace

hipError_t guide_refresh(const GuideDevice& g, hipStream_t s) {
    hipLaunchKernelGGL(k_guide_refresh, dim3((kLayers * 4096 + 255) / 256), dim3(256), 0, s, g);
    return hipGetLastError();
}

hipError_t guide_infer(const GuideDevice& g, const GuideBox& box, const float4* in, uint64_t n, float* out,
                       int grid, hipStream_t s, const uint32_t* n_dev, bool mixture) {
    if (n == 0) return hipSuccess;
    const uint64_t need = (n + 127) / 128;
    const int gr = (int)(need < (uint64_t)grid ? need : (uint64_t)grid);
    if (mixture) hipLaunchKernelGGL(k_guide_infer<true>, dim3(gr), dim3(256), 0, s, g, box, in, n, n_dev, out);
    else hipLaunchKernelGGL(k_guide_infer<false>, dim3(gr), dim3(256), 0, s, g, box, in, n, n_dev, out);
    return hipGetLastError();
}

hipError_t guide_train_step(const GuideDevice& g, const GuideBox& box, const float4* in, const float4* target,
                            uint64_t n, const GuideScratch& sc, float lr, int t, int grid, float* grad_out,
                            hipStream_t s) {
    if (n == 0) return hipSuccess;
    const uint64_t need = (n + 127) / 128;
    const int gr = (int)(need < (uint64_t)grid ? need : (uint64_t)grid);
    hipLaunchKernelGGL(k_guide_train, dim3(gr), dim3(256), 0, s, g, box, in, target, n, 1.0f / (float)n, sc.xs,
                       sc.dys, sc.loss);
    const uint64_t chunks = (n + sc.chunk - 1) / sc.chunk;
    hipLaunchKernelGGL(k_guide_dw, dim3((uint32_t)chunks, kLayers * 4), dim3(64), 0, s, sc.xs, sc.dys, n, sc.chunk,
                       sc.partial, sc.bpartial);
    hipLaunchKernelGGL(k_guide_adam, dim3((kParamsTotal + 255) / 256), dim3(256), 0, s, g, sc.partial, sc.bpartial,
                       chunks, lr, 0.9f, 0.999f, 1e-8f, t, grad_out);
    return hipGetLastError();
}

}  // namespace bg
