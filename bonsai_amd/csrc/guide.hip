// Path-guiding MLP on gfx950 matrix cores (no counterpart in the reference).
//
// Network: 64-dim encoding of (position, outgoing direction, normal) -> 4
// hidden layers of 64 ReLU units -> 40 outputs = 8 von Mises-Fisher lobes x
// (mixture logit, direction 3-vector, concentration parameter s with log kappa
// = 1.5 + 5.5 tanh((s - 1.5) / 5.5)). Weights are kept in
// FP32 (master copy, Adam moments) and in BF16 for the matrix cores; every
// product is gfx950's v_mfma_f32_32x32x16_bf16 with FP32 accumulation.
//
// One wavefront evaluates 32 samples at a time as Y^T = W X^T: the weights
// are the A operand (32 output rows x 16 inputs per instruction) and the
// activations the B operand (16 inputs x 32 samples). Lane l supplies 8 of an
// instruction's 16 inputs, half h = l/32 of them, for both operands; which
// feature each of those 8 slots holds is free as long as A and B agree. The
// accumulator layout (lane l, register r -> row 8*(r/4) + 4*(l/32) + r%4,
// column l%32) puts features 8q + 4h .. + 3 in registers 4q .. 4q + 3, so
// k-step t takes the quads q = 2t and 2t + 1 of every lane: the activations
// stay in registers from layer to layer, as B operands, and only the bias,
// ReLU and the BF16 rounding touch them in between. The backward pass reuses the same scheme
// with the transposed weights. Weight gradients (dW = dY X^T summed over the
// batch) are accumulated per block in the training kernel and reduced, in a
// fixed order, by the Adam kernel.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bg_internal.h"
#include "dev_guide.h"

namespace bg {
namespace {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef short s8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
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
// over the banks instead of all starting on bank 0. Quad q of output block mb
// for this lane: row 32*mb + lane%32, columns 8q + 4*(lane/32) .. +3.
constexpr int kLd = 68;
constexpr int kLayerLds = 64 * kLd;
__device__ __forceinline__ s4 load_a(const uint16_t* __restrict__ w, int mb, int q, int lane) {
    const uint16_t* p = w + (32 * mb + (lane & 31)) * kLd + 8 * q + 4 * (lane >> 5);
    return *reinterpret_cast<const s4*>(p);
}

// One 32 x 32 x 16 product step, each operand given as this lane's two quads of inputs.
__device__ __forceinline__ f16v mfma16(s4 a0, s4 a1, s4 b0, s4 b1, f16v c) {
    const s8 a = __builtin_shufflevector(a0, a1, 0, 1, 2, 3, 4, 5, 6, 7);
    const s8 b = __builtin_shufflevector(b0, b1, 0, 1, 2, 3, 4, 5, 6, 7);
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(bf8v, a), __builtin_bit_cast(bf8v, b), c, 0, 0,
                                                   0);
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

// One layer: acc[mb] = b + W[32mb.., :] * X^T, with X^T given as the B-operand quads xb[q]; with
// BIAS the accumulators start from the bias b (16-byte aligned), so no add follows the products.
// Registers 4q .. 4q + 3 hold rows r0 .. r0 + 3 (r0 = acc_row(mb, 4q, lane)), and out_row maps
// such a quad to four consecutive outputs, so each quad of the bias is one 16-byte read.
// permute_out: the accumulator rows hold the outputs out_row(row) (last layer).
template <bool BIAS>
__device__ __forceinline__ void layer_mm(const uint16_t* __restrict__ w, const s4 xb[8], int lane, f16v acc[2],
                                         const float* __restrict__ bias = nullptr, bool permute_out = false) {
#pragma unroll
    for (int mb = 0; mb < 2; ++mb) {
        f16v c = {0};
        if (BIAS)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int row = acc_row(mb, 4 * q, lane);
                const float4 b = *reinterpret_cast<const float4*>(bias + (permute_out ? out_row(row) : row));
                c[4 * q] = b.x;
                c[4 * q + 1] = b.y;
                c[4 * q + 2] = b.z;
                c[4 * q + 3] = b.w;
            }
#pragma unroll
        for (int t = 0; t < 4; ++t)
            c = mfma16(load_a(w, mb, 2 * t, lane), load_a(w, mb, 2 * t + 1, lane), xb[2 * t], xb[2 * t + 1], c);
        acc[mb] = c;
    }
}

// A hidden layer's activations as the next layer's B operands: the accumulators rounded to
// BF16, then the ReLU on the rounded values as a signed 16-bit max with 0 (rounding keeps the
// sign, and a negative BF16 value or -0 is a negative 16-bit integer), two values per
// instruction. The max is written as an instruction: as a vector max on the rounded values the
// compiler splits each two-value conversion into two one-value conversions and a permute.
__device__ __forceinline__ uint32_t pk_max0(uint32_t x) {
    uint32_t r;
    asm("v_pk_max_i16 %0, %1, 0" : "=v"(r) : "v"(x));
    return r;
}
__device__ __forceinline__ void relu_pack(const f16v acc[2], s4 xb[8]) {
    typedef uint32_t u2v __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            u2v u = __builtin_bit_cast(
                u2v, pack4(acc[mb][4 * q], acc[mb][4 * q + 1], acc[mb][4 * q + 2], acc[mb][4 * q + 3]));
            u.x = pk_max0(u.x);
            u.y = pk_max0(u.y);
            xb[4 * mb + q] = __builtin_bit_cast(s4, u);
        }
}

// B operands of the encoding for the sample of this lane: quad s holds
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
// row p takes output row out_row(p), with permute_cols (the transposed copy)
// its column p takes output column out_row(p). out_row keeps groups of 4
// together, so whole quads move. Every load of a thread is issued before the
// first LDS write, so the block waits for one memory latency, not twenty.
constexpr int kBlock = 256;  // threads of the training kernel
template <int BLOCK = kBlock>
__device__ void load_weights_lds(const uint16_t* __restrict__ wg, uint16_t* __restrict__ lds, bool permute_out = false,
                                 bool permute_cols = false) {
    constexpr int kQuadsPerThread = kLayers * 64 * 16 / BLOCK;
    static_assert(kLayers * 64 * 16 % BLOCK == 0, "weight quads split evenly over the block");
    const uint2* src = reinterpret_cast<const uint2*>(wg);
    uint2 v[kQuadsPerThread];
#pragma unroll
    for (int k = 0; k < kQuadsPerThread; ++k) {
        const int i = threadIdx.x + k * BLOCK;
        const int row = i / 16, q = i % 16;  // row over all layers, 4-element quad in the row
        const bool last = row >= (kLayers - 1) * 64;
        const int from = permute_out && last ? (kLayers - 1) * 64 + out_row(row % 64) : row;
        const int fq = permute_cols && last ? out_row(4 * q) / 4 : q;
        v[k] = src[from * 16 + fq];
    }
#pragma unroll
    for (int k = 0; k < kQuadsPerThread; ++k) {
        const int i = threadIdx.x + k * BLOCK;
        *reinterpret_cast<uint2*>(lds + (i / 16) * kLd + 4 * (i % 16)) = v[k];
    }
    __syncthreads();
}

// ------------------------------------------------------------- inference
// in: 3 float4 per sample (p, wo, n); out: 40 floats per sample.
// MIX: write the 10-float4 mixture record of each sample (dev_guide.h) instead
// of its 40 raw outputs.
// Blocks of kInferBlock threads share one LDS copy of the weights (44 KB), so
// LDS leaves room for kInferBlocksPerCu blocks per CU; registers are held to
// that many waves per SIMD.
constexpr int kInferWaves = kInferBlock / 64;
static_assert(kInferBlocksPerCu * kInferWaves % 4 == 0, "whole waves per SIMD");
template <bool MIX>
__global__ __launch_bounds__(kInferBlock) __attribute__((amdgpu_waves_per_eu(kInferBlocksPerCu * kInferWaves / 4, 8))) void k_guide_infer(GuideDevice g, GuideBox box, const float4* __restrict__ in,
                                                     uint64_t n, const uint32_t* __restrict__ n_dev,
                                                     float* __restrict__ out) {
    __shared__ uint16_t wl[kLayers * kLayerLds];
    __shared__ __align__(16) float bl[kLayers * 64];
    if (n_dev) n = *n_dev;
    if ((uint64_t)blockIdx.x * 32 * kInferWaves >= n) return;  // the whole block is idle (uniform)
    for (int k = threadIdx.x; k < kLayers * 64; k += kInferBlock) bl[k] = g.bias[k];
    load_weights_lds<kInferBlock>(g.wbf, wl, true);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, h = lane >> 5;
    const uint64_t groups = (n + 31) / 32;
    for (uint64_t grp = (uint64_t)blockIdx.x * kInferWaves + wave; grp < groups; grp += (uint64_t)gridDim.x * kInferWaves) {
        const uint64_t i = grp * 32 + (lane & 31);
        const bool valid = i < n;
        Sample smp = valid ? load_sample(in, i) : Sample{{0, 0, 0}, {0, 0, 1}, {0, 0, 1}};
        float q[3];
        normalise_pos(smp, box, q);
        s4 xb[8];
        encode_b(q, smp.wo, smp.n, lane, xb);
        f16v acc[2];
        // an offset the compiler cannot see through, so the weight reads stay in this loop
        // instead of being hoisted out of it into registers that would then spill
        int w_off = 0;
        asm volatile("" : "+s"(w_off));
        for (int l = 0; l < kLayers; ++l) {
            layer_mm<true>(wl + w_off + l * kLayerLds, xb, lane, acc, bl + l * 64, l == kLayers - 1);
            if (l < kLayers - 1) relu_pack(acc, xb);
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
__device__ __forceinline__ float vmf(const float* mu, float kappa, const float* w) {
    const float c = kappa / (2.0f * kPiF * (-expm1f(-2.0f * kappa)));
    return c * expf(kappa * (mu[0] * w[0] + mu[1] * w[1] + mu[2] * w[2] - 1.0f));
}

// d(-weight * log p(w)) / d out, scaled by inv_batch (oracle loss_grad_out), for
// the 4 lobes this lane holds (lobes 4h .. 4h + 3 of its sample, o: their 20
// outputs); the two halves of the wavefront exchange the logit maximum,
// the exp sum and the mixture density, each adding its own part first (the sums
// are the same on both halves). g gets d loss / d o for the same 20 outputs and
// *loss the sample's loss. All 64 lanes must call it (it shuffles).
__device__ __forceinline__ void out_grad_half(const float* o, const float* w, float weight, float inv_batch, float* g,
                                              float* loss) {
    float amax = -3.0e38f;
#pragma unroll
    for (int j = 0; j < 4; ++j) amax = fmaxf(amax, o[5 * j]);
    amax = fmaxf(amax, __shfl_xor(amax, 32));
    float pi[4], sum = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        pi[j] = expf(o[5 * j] - amax);
        sum += pi[j];
    }
    sum += __shfl_xor(sum, 32);
    float mu[4][3], nm[4], sv[4], kap[4], f[4], p = 0.0f;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        pi[j] /= sum;
        const float mx = o[5 * j + 1], my = o[5 * j + 2], mz = o[5 * j + 3];
        nm[j] = sqrtf(mx * mx + my * my + mz * mz + 1e-12f);
        mu[j][0] = mx / nm[j];
        mu[j][1] = my / nm[j];
        mu[j][2] = mz / nm[j];
        sv[j] = tanhf((o[5 * j + 4] - kLogKappaC) / kLogKappaR);
        kap[j] = expf(kLogKappaC + kLogKappaR * sv[j]);
        f[j] = vmf(mu[j], kap[j], w);
        p += pi[j] * f[j];
    }
    p += __shfl_xor(p, 32);
    p = fmaxf(p, 1e-30f);  // a direction where every lobe underflows in FP32
    const bool live = weight != 0.0f;  // an empty record has no loss and no gradient
    *loss = live ? -weight * logf(p) : 0.0f;
    const float sc = -weight * inv_batch;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float r = pi[j] * f[j] / p;
        const float k = kap[j];
        const float muw = mu[j][0] * w[0] + mu[j][1] * w[1] + mu[j][2] * w[2];
        const float dlogk = 1.0f / k - 2.0f * expf(-2.0f * k) / (-expm1f(-2.0f * k)) + (muw - 1.0f);
        g[5 * j] = live ? sc * (r - pi[j]) : 0.0f;
        g[5 * j + 4] = live ? sc * r * dlogk * k * (1.0f - sv[j] * sv[j]) : 0.0f;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float proj = k * w[c] - k * muw * mu[j][c];
            g[5 * j + 1 + c] = live ? sc * r * proj / nm[j] : 0.0f;
        }
    }
}

// ------------------------------------------------------------- training
// One fused pass per block over its share of the batch: forward, loss gradient,
// backward and the weight gradients, with nothing but the per-sample loss and
// one partial gradient per block leaving the chip.
//
// Each wave takes 32-sample groups; the 4 waves of a block run in step, one
// group each per round. Every layer input X_l stays in registers (B-operand
// layout) from the forward to the backward pass, where it also gives the ReLU
// masks. The last layer's rows are permuted as in inference (out_row), so each
// half-wavefront holds four whole lobes of its sample and forms their loss
// gradient in registers; the transposed copy of that layer has its columns
// permuted the same way, and its weight and bias gradient rows are put back
// in output order when they are written.
//
// For the weight gradient dW_l = sum over samples of dY_l X_l^T the block's 128
// samples of X_l and dY_l are staged in LDS feature-major ([64 features][128
// samples], BF16), the layout in which both MFMA operands read samples along k;
// wave w accumulates block (w / 2, w % 2) of dW_l (32 x 32, 8 k-steps of 16) in
// registers across rounds, and thread t the bias gradient of row t / 4 over
// samples 32 (t % 4) .. + 31. LDS holds the weights, their transposes (the
// backward products' A operands) and the biases, 88 KB, and the staging, 33 KB.
// Registers (the 5 layer inputs and 5 gradient blocks stay live through the
// backward pass) allow one wave per SIMD: one block per CU. Summation order is
// fixed, so results repeat bit for bit.
constexpr int kStage = 132;  // staging row stride in BF16 (66 dwords: rows two banks apart)
constexpr int kStageTile = 64 * kStage;

// Writes a wave's 64 x 32 B-operand tile (feature 8s + 4h + j, sample lane % 32) into the
// block's staging tile at sample column 32 * wave.
__device__ __forceinline__ void stage(uint16_t* __restrict__ t, const s4 v[8], int lane, int wave) {
    const int col = 32 * wave + (lane & 31), h = lane >> 5;
#pragma unroll
    for (int s = 0; s < 8; ++s)
#pragma unroll
        for (int j = 0; j < 4; ++j) t[(8 * s + 4 * h + j) * kStage + col] = (uint16_t)v[s][j];
}

__global__ __launch_bounds__(kBlock, 1) void k_guide_train(GuideDevice g, GuideBox box, const float4* __restrict__ in,
                                                        const float4* __restrict__ target, uint64_t n, float inv_batch,
                                                        uint32_t rounds, float* __restrict__ partial,
                                                        float* __restrict__ bpartial, float* __restrict__ loss_out) {
    __shared__ uint16_t wl[kLayers * kLayerLds];
    __shared__ uint16_t wt[kLayers * kLayerLds];
    __shared__ uint16_t stg[2 * kStageTile];  // X tile, dY tile
    __shared__ __align__(16) float bl[kLayers * 64];
    uint16_t* xt = stg;
    uint16_t* dt = stg + kStageTile;
    for (int k = threadIdx.x; k < kLayers * 64; k += kBlock) bl[k] = g.bias[k];
    load_weights_lds(g.wbf, wl, true);
    load_weights_lds(g.wbft, wt, false, true);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, h = lane >> 5;
    const int mbw = wave >> 1, nbw = wave & 1;  // this wave's block of every dW_l
    const int brow = threadIdx.x >> 2, bq = threadIdx.x & 3;  // this thread's bias row and sample quarter
    f16v dw[kLayers];
    float bsum[kLayers];
#pragma unroll
    for (int l = 0; l < kLayers; ++l) {
        dw[l] = (f16v){0};
        bsum[l] = 0.0f;
    }
    // sample index of this lane in round r; a round's inputs are loaded one round ahead
    auto index = [&](uint32_t r) { return (((uint64_t)blockIdx.x * rounds + r) * 4 + wave) * 32 + (lane & 31); };
    auto fetch = [&](uint64_t i, Sample& smp, float4& tg) {
        smp = i < n ? load_sample(in, i) : Sample{{0, 0, 0}, {0, 0, 1}, {0, 0, 1}};
        tg = i < n ? target[i] : make_float4(0, 0, 1, 0);
    };
    Sample next_smp;
    float4 next_tg;
    fetch(index(0), next_smp, next_tg);
    for (uint32_t r = 0; r < rounds; ++r) {
        const uint64_t i = index(r);
        const bool valid = i < n;
        const Sample smp = next_smp;
        const float4 tg = next_tg;
        if (r + 1 < rounds) fetch(index(r + 1), next_smp, next_tg);
        float q[3];
        normalise_pos(smp, box, q);
        s4 xs[kLayers][8];
        encode_b(q, smp.wo, smp.n, lane, xs[0]);
        f16v acc[2];
        s4 nx[8];
#pragma unroll
        for (int l = 0; l < kLayers; ++l) {
            layer_mm<true>(wl + l * kLayerLds, xs[l], lane, acc, bl + l * 64, l == kLayers - 1);
            if (l + 1 < kLayers) relu_pack(acc, nx);
            if (l + 1 < kLayers)
#pragma unroll
                for (int s = 0; s < 8; ++s) xs[l + 1][s] = nx[s];
        }
        // outputs 20h .. 20h + 19 of this lane's sample: accumulators 0-15 of block 0, 0-3 of block 1
        float o[20], gr[20], ls;
#pragma unroll
        for (int k = 0; k < 16; ++k) o[k] = acc[0][k];
#pragma unroll
        for (int k = 0; k < 4; ++k) o[16 + k] = acc[1][k];
        const float w[3] = {tg.x, tg.y, tg.z};
        out_grad_half(o, w, tg.w, inv_batch, gr, &ls);
        if (valid && h == 0) loss_out[i] = ls;
        // dY of the last layer in the same (permuted) accumulator layout, as B operands
        s4 db[8];
#pragma unroll
        for (int qq = 0; qq < 4; ++qq) db[qq] = pack4(gr[4 * qq], gr[4 * qq + 1], gr[4 * qq + 2], gr[4 * qq + 3]);
        db[4] = pack4(gr[16], gr[17], gr[18], gr[19]);
        db[5] = db[6] = db[7] = (s4){0, 0, 0, 0};
#pragma unroll
        for (int l = kLayers - 1; l >= 0; --l) {
            __syncthreads();  // the previous reads of the staging tiles are done
            stage(xt, xs[l], lane, wave);
            stage(dt, db, lane, wave);
            __syncthreads();
            // dW_l over the block's 128 samples of this round
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const s4* a = reinterpret_cast<const s4*>(dt + (32 * mbw + (lane & 31)) * kStage + 16 * t + 8 * h);
                const s4* b = reinterpret_cast<const s4*>(xt + (32 * nbw + (lane & 31)) * kStage + 16 * t + 8 * h);
                dw[l] = mfma16(a[0], a[1], b[0], b[1], dw[l]);
            }
            {  // bias: row brow, samples 32 bq .. 32 bq + 31, in order
                const uint2* rp = reinterpret_cast<const uint2*>(dt + brow * kStage + 32 * bq);
                float sum = 0.0f;
#pragma unroll
                for (int k = 0; k < 8; ++k) {
                    const uint2 v = rp[k];
                    sum += __uint_as_float(v.x << 16);
                    sum += __uint_as_float(v.x & 0xFFFF0000u);
                    sum += __uint_as_float(v.y << 16);
                    sum += __uint_as_float(v.y & 0xFFFF0000u);
                }
                bsum[l] += sum;
            }
            if (l == 0) break;
            // d of layer-l input = W_l^T dY_l, masked by the ReLU that made X_l: X_l is +0 or
            // positive in BF16, so the mask is X_l != 0, applied to the rounded values
            layer_mm<false>(wt + l * kLayerLds, db, lane, acc);
#pragma unroll
            for (int mb = 0; mb < 2; ++mb)
#pragma unroll
                for (int qq = 0; qq < 4; ++qq) {
                    const s4 d = pack4(acc[mb][4 * qq], acc[mb][4 * qq + 1], acc[mb][4 * qq + 2], acc[mb][4 * qq + 3]);
                    db[4 * mb + qq] = d & (xs[l][4 * mb + qq] != (s4){0, 0, 0, 0});
                }
        }
    }
    // this block's partial gradients (summed over blocks in the Adam kernel), last layer rows in
    // output order
    float* dst = partial + (uint64_t)blockIdx.x * kLayers * 4096;
#pragma unroll
    for (int l = 0; l < kLayers; ++l)
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int row = acc_row(mbw, k, lane);
            dst[l * 4096 + (l == kLayers - 1 ? out_row(row) : row) * 64 + 32 * nbw + (lane & 31)] = dw[l][k];
        }
#pragma unroll
    for (int l = 0; l < kLayers; ++l) {
        float v = bsum[l];
        v += __shfl_xor(v, 1);  // (q + (q ^ 1)) + ((q ^ 2) + (q ^ 3))
        v += __shfl_xor(v, 2);
        if (bq == 0) bpartial[((uint64_t)blockIdx.x * kLayers + l) * 64 + (l == kLayers - 1 ? out_row(brow) : brow)] = v;
    }
}

// Sums the per-block partials into the gradient of every parameter (oracle flat
// order), takes one Adam step on the FP32 master weights and refreshes the BF16
// copies. A block takes 64 consecutive parameters with 16 slices of threads:
// slice k sums partials k, k + 16, ..., then the 16 slice sums are added in
// slice order, so the result does not depend on timing.
constexpr int kAdamParams = 64, kAdamSlices = 16;
__global__ __launch_bounds__(kAdamParams* kAdamSlices) void k_guide_adam(
    GuideDevice g, const float* __restrict__ partial, const float* __restrict__ bpartial, uint64_t chunks, float lr,
    float b1, float b2, float eps, int t, float* __restrict__ grad_out) {
    __shared__ float red[kAdamSlices][kAdamParams];
    const int p = threadIdx.x % kAdamParams, slice = threadIdx.x / kAdamParams;
    const int idx = blockIdx.x * kAdamParams + p;
    // locate (layer, row, col) of flat index idx: per layer W (rows x cols) then b (rows)
    int l = 0, base = 0;
    int rows = 64;
    const int cols = 64;
    for (; l < kLayers; ++l) {
        rows = l == kLayers - 1 ? kOut : 64;
        const int sz = rows * cols + rows;
        if (idx < base + sz) break;
        base += sz;
    }
    const int off = idx - base;
    int row, col = -1;
    const float* src;
    uint64_t stride;
    if (off < rows * cols) {
        row = off / cols;
        col = off % cols;
        src = partial + ((uint64_t)l * 64 + row) * 64 + col;
        stride = (uint64_t)kLayers * 4096;
    } else {
        row = off - rows * cols;
        src = bpartial + (uint64_t)l * 64 + row;
        stride = (uint64_t)kLayers * 64;
    }
    float part = 0.0f;
    if (idx < kParamsTotal)
        for (uint64_t c = slice; c < chunks; c += kAdamSlices) part += src[c * stride];
    red[slice][p] = part;
    __syncthreads();
    if (slice != 0 || idx >= kParamsTotal) return;
    float gsum = 0.0f;
    for (int k = 0; k < kAdamSlices; ++k) gsum += red[k][p];
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
    const uint64_t need = (n + 32 * kInferWaves - 1) / (32 * kInferWaves);
    const int gr = (int)(need < (uint64_t)grid ? need : (uint64_t)grid);
    if (mixture) hipLaunchKernelGGL(k_guide_infer<true>, dim3(gr), dim3(kInferBlock), 0, s, g, box, in, n, n_dev, out);
    else hipLaunchKernelGGL(k_guide_infer<false>, dim3(gr), dim3(kInferBlock), 0, s, g, box, in, n, n_dev, out);
    return hipGetLastError();
}

hipError_t guide_train_step(const GuideDevice& g, const GuideBox& box, const float4* in, const float4* target,
                            uint64_t n, const GuideScratch& sc, float lr, int t, int grid, float* grad_out,
                            hipStream_t s) {
    if (n == 0) return hipSuccess;
    // blocks of 4 waves x `rounds` groups of 32 samples, at most sc.blocks of them (their partials)
    const uint64_t groups = (n + 31) / 32;
    uint64_t blocks = std::min<uint64_t>((uint64_t)std::min(grid, (int)sc.blocks), (groups + 3) / 4);
    const uint32_t rounds = (uint32_t)((groups + 4 * blocks - 1) / (4 * blocks));
    blocks = (groups + 4ull * rounds - 1) / (4ull * rounds);
    hipLaunchKernelGGL(k_guide_train, dim3((uint32_t)blocks), dim3(kBlock), 0, s, g, box, in, target, n,
                       1.0f / (float)n, rounds, sc.partial, sc.bpartial, sc.loss);
    hipLaunchKernelGGL(k_guide_adam, dim3((kParamsTotal + kAdamParams - 1) / kAdamParams),
                       dim3(kAdamParams * kAdamSlices), 0, s, g, sc.partial, sc.bpartial, blocks, lr, 0.9f, 0.999f,
                       1e-8f, t, grad_out);
    return hipGetLastError();
}

}  // namespace bg
