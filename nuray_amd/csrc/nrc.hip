// Neural radiance cache kernels for gfx950: inference, a fused training pass (forward, backward and
// weight gradient) and Adam, all matrix work on BF16 MFMA with FP32 accumulation.
//
// The reference's cache (NeuralRadianceCache / NNNode*, renderernrc.h:17-43, nnnode.cpp:1-208) is a
// per-sample scalar MLP trained by SGD; this one is the 44 -> 5 x 64 (ReLU) -> 3 network trained
// with Adam on batches of 2^16 samples and the relative L2 loss
//   L = mean_s sum_c (y_sc - t_sc)^2 / (y_sc^2 + 0.01)      (y in the denominator held constant)
// whose output gradient has the reference's form (o - a) / (o^2 + 0.01) (nnnode.cpp:195-196).
//
// Kernels
//   k_nrc_infer      persistent; weights in LDS once per block (16 waves: 4 per SIMD), one
//                    64-sample tile per wave, the encoded inputs through a half-width LDS tile and
//                    the hidden activations in registers (nrc_forward); the next tile's queries are
//                    loaded while the current one runs
//   k_nrc_train      persistent blocks of 8 waves over 64-sample tiles: forward, loss gradient,
//                    backward (ReLU masks read back from the kept activations) and the weight
//                    gradient, with every layer's input and gradient tile kept in LDS in both
//                    sample-major and feature-major layout (nothing per sample goes to global
//                    memory); each block's weight-gradient partial is written once; the weights
//                    come from L2 (W^T for the forward, the transposed copy W for the backward);
//                    per-tile loss slots instead of one device-wide atomic
//   k_nrc_adam       sums the blocks' partials (16 slices per parameter, 8 loads in flight per
//                    thread) and applies Adam to the FP32 master weights, refreshes the BF16
//                    copies (W^T and W)
#include <cmath>
#include <cstring>
#include <type_traits>
#include <stdexcept>
#include <vector>
#include "nrc.h"

namespace nr {

namespace {

// waves per inference block: 48 KB of weights + 16 x 5 KB half-width input tiles = 128 KB of LDS,
// one block per CU, 4 waves per SIMD (both forms of the kernel stay within 128 VGPRs)
constexpr int INF_WAVES = 16;

template <bool ACC>
__global__ __launch_bounds__(INF_WAVES * 64) void k_nrc_infer(const __bf16* __restrict__ w16,
                                                              const NrcQuery* __restrict__ q,
                                                              const int* __restrict__ count, int n_max,
                                                              float4* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) __bf16 sW[NRC_LDS_W];
    __shared__ __attribute__((aligned(16))) __bf16 sX[INF_WAVES][NRC_TILE_HALF];
    nrc_load_weights(w16, sW);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = count ? min(*count, n_max) : n_max;
    __bf16* X = sX[wave];
    const int stride = gridDim.x * INF_WAVES;
    // the lane's query of the wave's next tile is loaded while the current tile runs its layers
    auto fetch = [&](int chunk) { return chunk * 64 + lane < n ? q[chunk * 64 + lane] : NrcQuery{}; };
    NrcQuery nxt = fetch(blockIdx.x * INF_WAVES + wave);
    for (int chunk = blockIdx.x * INF_WAVES + wave; chunk * 64 < n; chunk += stride) {
        const int s = chunk * 64 + lane;
        const NrcQuery rec = nxt;
        // the layer-0 B operands, through the half-width tile: features 0..31, then 32..63
        bf16x8 b[4][2];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            nrc_encode_half(rec, s < n, X, lane, h);
            wave_sync();
#pragma unroll
            for (int nt = 0; nt < 4; ++nt)
                b[nt][h] = *reinterpret_cast<const bf16x8*>(X + (16 * nt + (lane & 15)) * NRC_LDH + 8 * (lane >> 4));
            wave_sync();
        }
        if ((chunk + stride) * 64 < n) nxt = fetch(chunk + stride);
        float4 y[4];
        nrc_forward(sW, b, y, lane);
        // lanes 0..15 finish samples 16 nt + lane, whose records sit in lanes 16 nt + lane
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const int si = chunk * 64 + 16 * nt + lane;
            if (ACC) {
                const int src = 16 * nt + (lane & 15);
                const int slot = __shfl(__float_as_int(rec.c.w), src, 64);
                const float dx = __shfl(rec.d.x, src, 64), dy = __shfl(rec.d.y, src, 64),
                            dz = __shfl(rec.d.z, src, 64);
                if (lane < 16 && si < n && slot >= 0) {
                    float4 r = out[slot];
                    r.x += dx * fmaxf(y[nt].x, 0.0f);
                    r.y += dy * fmaxf(y[nt].y, 0.0f);
                    r.z += dz * fmaxf(y[nt].z, 0.0f);
                    out[slot] = r;
                }
            } else if (lane < 16 && si < n) {
                out[si] = y[nt];
            }
        }
    }
}

// One training pass over the batch: forward, loss gradient, backward and the weight gradient,
// fused.  Persistent blocks of 8 waves (two per SIMD); block b takes the 64-sample tiles b, b + B, ...
//
// Every activation and gradient tile lives in LDS in two layouts (padded rows of NRC_LD bf16):
// sample-major ([sample][feature], the K = feature operand of the forward and backward products)
// and feature-major ([feature][sample], the K = sample operands of the weight gradient).  A
// 16x16x32 MFMA issued with its operands swapped computes D^T = B^T A^T, which lands in the lane
// layout of the transposed tile, so the feature-major copy costs one more MFMA per product and a
// plain 8-byte store.  LDS: X_0..X_5 in both layouts (kept for the backward, whose ReLU masks are
// read back from them) and the gradients G_l in both layouts (two tiles each): 144 KB, one block
// per CU.
//
// Wave (w, h) = (wave & 3, wave >> 2) computes output features [16 w, 16 w + 16) of samples
// [32 h, 32 h + 32) of every hidden layer and, in the backward, input features [16 w, 16 w + 16) of
// the same samples.  It accumulates rows [16 w, 16 w + 16), columns [32 h, 32 h + 32) of every
// hidden layer's weight gradient dW_l = G_l^T X_l over its tiles in registers (waves h = 0 also
// the output layer's rows 0..15, columns [16 w, 16 w + 16), and its loss over samples
// [16 w, 16 w + 16)), and writes them once, at the end, as the block's partial
// part[(l B + b) 4096 + row 64 + col].  The weights come from global memory (L2-resident, one layer
// ahead): W^T rows (w16) for the forward, W rows (w16t, kept by Adam) for the backward.  The next
// tile's records are loaded during the backward.  loss[4 tile + w] = the summed relative L2 loss
// of samples [16 w, 16 w + 16) of the tile (no atomics).
constexpr int FT_WAVES = 8;

// The A operand fragments of one weight layer for lane (r, qd): row `row` of a [64][64] bf16
// matrix, k = 8 qd .. 8 qd + 7 and 32 + 8 qd ...
struct AFrag {
    bf16x8 a0, a1;
};
__device__ __forceinline__ AFrag load_afrag(const __bf16* m, int row, int qd) {
    const __bf16* p = m + row * 64 + 8 * qd;
    return AFrag{*reinterpret_cast<const bf16x8*>(p), *reinterpret_cast<const bf16x8*>(p + 32)};
}

__device__ __forceinline__ bf16x8 lds8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }
__device__ __forceinline__ bf16x4 lds4(const __bf16* p) { return *reinterpret_cast<const bf16x4*>(p); }

// Input features 8 g .. 8 g + 7 of one record (g wave-uniform; each case folds to its features).
template <int F0>
__device__ __forceinline__ bf16x8 encode8(const NrcQuery& rec, bool valid) {
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = (__bf16)(valid ? nrc_feature(rec, F0 + j) : 0.0f);
    return v;
}
__device__ __forceinline__ bf16x8 encode_group(int g, const NrcQuery& rec, bool valid) {
    switch (g) {
        case 0: return encode8<0>(rec, valid);
        case 1: return encode8<8>(rec, valid);
        case 2: return encode8<16>(rec, valid);
        case 3: return encode8<24>(rec, valid);
        case 4: return encode8<32>(rec, valid);
        case 5: return encode8<40>(rec, valid);
        case 6: return encode8<48>(rec, valid);
        default: return encode8<56>(rec, valid);
    }
}

__global__ __launch_bounds__(FT_WAVES * 64) void k_nrc_train(const __bf16* __restrict__ w16,
                                                             const __bf16* __restrict__ w16t,
                                                             const NrcQuery* __restrict__ q,
                                                             const float4* __restrict__ tgt, int n, float inv_n,
                                                             float* __restrict__ part, float* __restrict__ loss) {
    __shared__ __attribute__((aligned(16))) __bf16 sX[6][NRC_TILE];
    __shared__ __attribute__((aligned(16))) __bf16 sXT[6][NRC_TILE];
    __shared__ __attribute__((aligned(16))) __bf16 sG[2][NRC_TILE];
    __shared__ __attribute__((aligned(16))) __bf16 sGT[2][NRC_TILE];
    const int lane = threadIdx.x & 63, r = lane & 15, qd = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int w = wave & 3, h = wave >> 2;
    const int tiles = (n + 63) / 64;
    f32x4 dw[5][2], dw5 = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int l = 0; l < 5; ++l)
#pragma unroll
        for (int j = 0; j < 2; ++j) dw[l][j] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // this lane's record, for the first tile (later tiles: fetched during the previous backward)
    NrcQuery rec{};
    auto fetch = [&](int t) {
        const int s = t * 64 + lane;
        rec = s < n ? q[s] : NrcQuery{};
    };
    fetch(blockIdx.x);
    for (int tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        const int s0 = tile * 64;
        AFrag nxt = load_afrag(w16, 16 * w + r, qd);
        float4 tv = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // waves h = 0, qd = 0: target of sample 16 w + r
        if (h == 0 && qd == 0 && s0 + 16 * w + r < n) tv = tgt[s0 + 16 * w + r];
        // encode: this wave writes features 8 wave .. 8 wave + 7 of sample `lane`, both layouts
        {
            const bf16x8 v = encode_group(wave, rec, s0 + lane < n);
            *reinterpret_cast<bf16x8*>(sX[0] + lane * NRC_LD + 8 * wave) = v;
#pragma unroll
            for (int j = 0; j < 8; ++j) sXT[0][(8 * wave + j) * NRC_LD + lane] = v[j];
        }
        __syncthreads();
        // forward: X_{l+1} = ReLU(W_l X_l), both layouts
#pragma unroll
        for (int l = 0; l < 5; ++l) {
            const AFrag A = nxt;
            nxt = load_afrag(w16 + (l + 1) * 4096, l < 4 ? 16 * w + r : r, qd);
            const __bf16* X = sX[l];
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int nt = 2 * h + j;
                const bf16x8 b0 = lds8(X + (16 * nt + r) * NRC_LD + 8 * qd);
                const bf16x8 b1 = lds8(X + (16 * nt + r) * NRC_LD + 32 + 8 * qd);
                f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f}, ct = {0.0f, 0.0f, 0.0f, 0.0f};
                c = mfma16(A.a0, b0, c);
                ct = mfma16(b0, A.a0, ct);
                c = mfma16(A.a1, b1, c);
                ct = mfma16(b1, A.a1, ct);
                *reinterpret_cast<bf16x4*>(sX[l + 1] + (16 * nt + r) * NRC_LD + 16 * w + 4 * qd) = relu_bf16(c);
                *reinterpret_cast<bf16x4*>(sXT[l + 1] + (16 * w + r) * NRC_LD + 16 * nt + 4 * qd) = relu_bf16(ct);
            }
            __syncthreads();
        }
        // output layer (rows 0..15 of layer 5) on waves h = 0, samples 16 w .. 16 w + 15, and its
        // gradient G_5: sample-major columns 0..2 of sG[0] (columns 3..31 zero: the backward's K of
        // layer 5 spans 32 outputs), feature-major rows 0..2 of sGT[0] (rows 3..15 zero)
        const AFrag A5 = nxt;
        nxt = load_afrag(w16t + 5 * 4096, 16 * w + r, qd);  // the backward's first layer
        if (h == 0) {
            const int nt = w;
            const bf16x8 b0 = lds8(sX[5] + (16 * nt + r) * NRC_LD + 8 * qd);
            const bf16x8 b1 = lds8(sX[5] + (16 * nt + r) * NRC_LD + 32 + 8 * qd);
            f32x4 y = {0.0f, 0.0f, 0.0f, 0.0f};
            y = mfma16(A5.a0, b0, y);
            y = mfma16(A5.a1, b1, y);
            const int s = s0 + 16 * nt + r;
            float loss_acc = 0.0f;
            bf16x4 gv = {(__bf16)0.0f, (__bf16)0.0f, (__bf16)0.0f, (__bf16)0.0f};
            if (qd == 0 && s < n) {
                const float t3[3] = {tv.x, tv.y, tv.z};
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float den = y[c] * y[c] + 0.01f, e = y[c] - t3[c];
                    loss_acc += e * e / den;
                    gv[c] = (__bf16)(2.0f * e / den * inv_n);
                }
            }
            const bf16x4 z = {(__bf16)0.0f, (__bf16)0.0f, (__bf16)0.0f, (__bf16)0.0f};
            *reinterpret_cast<bf16x4*>(sG[0] + (16 * nt + r) * NRC_LD + 4 * qd) = gv;
            *reinterpret_cast<bf16x4*>(sG[0] + (16 * nt + r) * NRC_LD + 16 + 4 * qd) = z;
            if (qd == 0) {
#pragma unroll
                for (int c = 0; c < 3; ++c) sGT[0][c * NRC_LD + 16 * nt + r] = gv[c];
            }
            if (lane < 26) {  // rows 3..15, columns 16 nt .. 16 nt + 15: two 8-sample stores a row
                const bf16x8 z8 = {};
                *reinterpret_cast<bf16x8*>(sGT[0] + (3 + (lane >> 1)) * NRC_LD + 16 * nt + 8 * (lane & 1)) = z8;
            }
            for (int off = 32; off > 0; off >>= 1) loss_acc += __shfl_xor(loss_acc, off);
            if (lane == 0) loss[4 * tile + w] = loss_acc;
        }
        __syncthreads();
        // backward: dW_l += G_l^T X_l (this wave's 16 rows x 32 columns of layer l), then
        // G_{l-1} = (W_l^T G_l) * [X_l > 0] for this wave's 16 input features and 32 samples, both
        // layouts (layer 0 needs G_0 feature-major only)
        int gc = 0;
#pragma unroll
        for (int l = 5; l >= 1; --l) {
            const AFrag A = nxt;
            if (l > 1) nxt = load_afrag(w16t + (l - 1) * 4096, 16 * w + r, qd);
            if (l == 2 && tile + int(gridDim.x) < tiles) fetch(tile + gridDim.x);
            const __bf16* Gl = sG[gc];
            const __bf16* GTl = sGT[gc];
            if (l < 5) {
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) {
                    const bf16x8 a = lds8(GTl + (16 * w + r) * NRC_LD + 32 * ks + 8 * qd);
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        dw[l][j] = mfma16(a, lds8(sXT[l] + (16 * (2 * h + j) + r) * NRC_LD + 32 * ks + 8 * qd), dw[l][j]);
                }
            } else if (h == 0) {
#pragma unroll
                for (int ks = 0; ks < 2; ++ks)
                    dw5 = mfma16(lds8(GTl + r * NRC_LD + 32 * ks + 8 * qd),
                                 lds8(sXT[5] + (16 * w + r) * NRC_LD + 32 * ks + 8 * qd), dw5);
            }
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int nt = 2 * h + j;
                const bf16x8 b0 = lds8(Gl + (16 * nt + r) * NRC_LD + 8 * qd);
                f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f}, ct = {0.0f, 0.0f, 0.0f, 0.0f};
                if (l > 1) c = mfma16(A.a0, b0, c);
                ct = mfma16(b0, A.a0, ct);
                if (l < 5) {
                    const bf16x8 b1 = lds8(Gl + (16 * nt + r) * NRC_LD + 32 + 8 * qd);
                    if (l > 1) c = mfma16(A.a1, b1, c);
                    ct = mfma16(b1, A.a1, ct);
                }
                if (l > 1)
                    *reinterpret_cast<bf16x4*>(sG[gc ^ 1] + (16 * nt + r) * NRC_LD + 16 * w + 4 * qd) =
                        masked_bf16(c, lds4(sX[l] + (16 * nt + r) * NRC_LD + 16 * w + 4 * qd));
                *reinterpret_cast<bf16x4*>(sGT[gc ^ 1] + (16 * w + r) * NRC_LD + 16 * nt + 4 * qd) =
                    masked_bf16(ct, lds4(sXT[l] + (16 * w + r) * NRC_LD + 16 * nt + 4 * qd));
            }
            __syncthreads();
            gc ^= 1;
        }
#pragma unroll
        for (int ks = 0; ks < 2; ++ks) {
            const bf16x8 a = lds8(sGT[gc] + (16 * w + r) * NRC_LD + 32 * ks + 8 * qd);
#pragma unroll
            for (int j = 0; j < 2; ++j)
                dw[0][j] = mfma16(a, lds8(sXT[0] + (16 * (2 * h + j) + r) * NRC_LD + 32 * ks + 8 * qd), dw[0][j]);
        }
        __syncthreads();  // the next tile overwrites X_0 and the gradient tiles
    }
    // this block's partial: hidden layers rows 16 w + 4 qd + i, columns 16 (2 h + j) + r; the
    // output layer (waves h = 0) rows 4 qd + i, columns 16 w + r
    const int B = gridDim.x;
#pragma unroll
    for (int l = 0; l < 5; ++l) {
        float* dst = part + (size_t(l) * B + blockIdx.x) * 4096;
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int i = 0; i < 4; ++i) dst[(16 * w + 4 * qd + i) * 64 + 16 * (2 * h + j) + r] = dw[l][j][i];
    }
    if (h == 0) {
        float* dst = part + (size_t(5) * B + blockIdx.x) * 4096;
#pragma unroll
        for (int i = 0; i < 4; ++i) dst[(4 * qd + i) * 64 + 16 * w + r] = dw5[i];
    }
}

// Adam over the summed partials: a block takes 64 consecutive parameters and sums their P
// partials in 16 interleaved slices (coalesced 256-byte rows, 8 independent loads in flight per
// thread, 16 waves a block: enough reads in flight to stream the partials), then one thread per
// parameter adds the slices in a fixed order and applies the update.
constexpr int ADAM_PARAMS = 64, ADAM_SLICES = 16;

__global__ __launch_bounds__(ADAM_PARAMS* ADAM_SLICES) void k_nrc_adam(
    float* __restrict__ w32, __bf16* __restrict__ w16, __bf16* __restrict__ w16t, float* __restrict__ m,
    float* __restrict__ v, float* __restrict__ wema, __bf16* __restrict__ w16inf, const float* __restrict__ part, int P,
    float lr, float b1, float b2, float eps, float c1, float c2, float ema, float ce, float* __restrict__ grad_out) {
    __shared__ float red[ADAM_SLICES - 1][ADAM_PARAMS];
    const int t = threadIdx.x % ADAM_PARAMS, sl = threadIdx.x / ADAM_PARAMS;
    const int i = blockIdx.x * ADAM_PARAMS + t;
    const int l = i < 5 * 4096 ? i / 4096 : 5, k = i - l * 4096;
    float g = 0.0f;
    if (i < NRC_NPARAM) {
        const float* src = part + size_t(l) * P * 4096 + k;
        int p = sl;
        for (; p + 7 * ADAM_SLICES < P; p += 8 * ADAM_SLICES) {
            float tt[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) tt[j] = src[size_t(p + j * ADAM_SLICES) * 4096];
#pragma unroll
            for (int j = 0; j < 8; ++j) g += tt[j];
        }
        for (; p < P; p += ADAM_SLICES) g += src[size_t(p) * 4096];
    }
    if (sl > 0) red[sl - 1][t] = g;
    __syncthreads();
    if (sl > 0 || i >= NRC_NPARAM) return;
#pragma unroll
    for (int j = 0; j < ADAM_SLICES - 1; ++j) g += red[j][t];
    if (grad_out) {
        grad_out[i] = g;
        return;
    }
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    const float w = w32[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w32[i] = w;
    w16[i] = (__bf16)w;
    w16t[l * 4096 + (k & 63) * 64 + (k >> 6)] = (__bf16)w;  // row = input, column = output
    if (ema > 0.0f) {
        const float e = ema * wema[i] + (1.0f - ema) * w;
        wema[i] = e;
        w16inf[i] = (__bf16)(e / ce);
    } else {
        w16inf[i] = (__bf16)w;
    }
}

__global__ void k_nrc_targets(const NrcQuery* __restrict__ rec, const float4* __restrict__ lb,
                              const float4* __restrict__ radiance, int n, float4* __restrict__ target) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const NrcQuery r = rec[i];
    const float4 L = radiance[__float_as_int(r.c.w)], b = lb[i];
    auto q = [](float l, float l0, float beta) { return beta > 1e-8f ? (l - l0) / beta : 0.0f; };
    target[i] = make_float4(q(L.x, b.x, r.d.x), q(L.y, b.y, r.d.y), q(L.z, b.z, r.d.z), 0.0f);
}

__global__ void k_nrc_interleave(const NrcQuery* __restrict__ in, const float4* __restrict__ tin, int n, int steps,
                                 NrcQuery* __restrict__ out, float4* __restrict__ tout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = i % steps, j = i / steps, base = n / steps, rem = n % steps;
    const int o = k * base + min(k, rem) + j;
    out[o] = in[i];
    tout[o] = tin[i];
}

// Synthetic queries for timing: positions in a 1000-unit box, angles and albedo in [0, 1).
__global__ void k_nrc_fill(NrcQuery* q, float4* t, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t h = uint32_t(i) * 0x9E3779B9u + 0x7F4A7C15u;
    auto u = [&h]() {
        h ^= h >> 16;
        h *= 0x85EBCA6Bu;
        h ^= h >> 13;
        h *= 0xC2B2AE35u;
        h ^= h >> 16;
        return float(h >> 8) * (1.0f / 16777216.0f);
    };
    NrcQuery r;
    r.a = make_float4(1000.0f * u(), 200.0f * u(), 1000.0f * u(), 0.05f + 0.75f * u());
    r.b = make_float4(u(), u(), u(), u());
    r.c = make_float4(u(), u(), u(), __int_as_float(-1));
    r.d = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
    q[i] = r;
    t[i] = make_float4(u(), u(), u(), 0.0f);
}

// BF16 copies of the FP32 weights: W^T as stored (b) and transposed per layer (bt: row = input,
// column = output; layer 5's columns 16..63 stay zero).
__global__ void k_to_bf16(const float* __restrict__ a, __bf16* __restrict__ b, __bf16* __restrict__ bt, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    b[i] = (__bf16)a[i];
    const int l = i < 5 * 4096 ? i / 4096 : 5, k = i - l * 4096;
    bt[l * 4096 + (k & 63) * 64 + (k >> 6)] = (__bf16)a[i];
}

int n_cus(int device) {
    int cus = 0;
    check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "CU count");
    return cus;
}

uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

void nrc_targets(const NrcQuery* rec, const float4* lb, const float4* radiance, int n, float4* target,
                 hipStream_t s) {
    if (n <= 0) return;
    k_nrc_targets<<<(n + 255) / 256, 256, 0, s>>>(rec, lb, radiance, n, target);
    check_hip(hipGetLastError(), "k_nrc_targets");
}

double Nrc::bench(int n, int reps, double* train_ms) {
    DevBuf q, t, o;
    q.ensure(sizeof(NrcQuery) * size_t(n));
    t.ensure(sizeof(float4) * size_t(n));
    o.ensure(sizeof(float4) * size_t(n));
    k_nrc_fill<<<(n + 255) / 256, 256, 0, stream>>>(q.as<NrcQuery>(), t.as<float4>(), n);
    check_hip(hipGetLastError(), "k_nrc_fill");
    hipEvent_t e0, e1;
    check_hip(hipEventCreate(&e0), "event");
    check_hip(hipEventCreate(&e1), "event");
    infer(q.as<NrcQuery>(), n, o.as<float4>());
    check_hip(hipEventRecord(e0, stream), "event");
    for (int r = 0; r < reps; ++r) infer(q.as<NrcQuery>(), n, o.as<float4>());
    check_hip(hipEventRecord(e1, stream), "event");
    check_hip(hipEventSynchronize(e1), "event");
    float ms = 0.0f;
    check_hip(hipEventElapsedTime(&ms, e0, e1), "elapsed");
    const double infer_ms = double(ms) / reps;
    if (train_ms) {
        const int nb = std::min(n, cfg.max_batch);
        train(q.as<NrcQuery>(), t.as<float4>(), nb, false);
        check_hip(hipEventRecord(e0, stream), "event");
        for (int r = 0; r < reps; ++r) train(q.as<NrcQuery>(), t.as<float4>(), nb, false);
        check_hip(hipEventRecord(e1, stream), "event");
        check_hip(hipEventSynchronize(e1), "event");
        check_hip(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        *train_ms = double(ms) / reps;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return infer_ms;
}

void nrc_interleave(const NrcQuery* in, const float4* tin, int n, int steps, NrcQuery* out, float4* tout,
                    hipStream_t s) {
    if (n <= 0) return;
    k_nrc_interleave<<<(n + 255) / 256, 256, 0, s>>>(in, tin, n, steps, out, tout);
    check_hip(hipGetLastError(), "k_nrc_interleave");
}

Nrc::Nrc(uint32_t seed, const NrcConfig& c) : cfg(c) {
    if (cfg.max_batch <= 0 || cfg.max_batch > (1 << 22)) throw std::runtime_error("nrc: max_batch out of range");
    check_hip(hipGetDevice(&device), "hipGetDevice");
    check_hip(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    w32.ensure(sizeof(float) * NRC_NPARAM);
    w16.ensure(sizeof(__bf16) * NRC_NPARAM);
    w16t.ensure(sizeof(__bf16) * 6 * 4096);
    check_hip(hipMemsetAsync(w16t.ptr, 0, sizeof(__bf16) * 6 * 4096, stream), "nrc transposed weights");
    m.ensure(sizeof(float) * NRC_NPARAM);
    v.ensure(sizeof(float) * NRC_NPARAM);
    wema.ensure(sizeof(float) * NRC_NPARAM);
    w16inf.ensure(sizeof(__bf16) * NRC_NPARAM);
    part.ensure(sizeof(float) * 6 * size_t(TRAIN_MAX_BLOCKS) * 4096);  // one partial per training block
    lossp.ensure(sizeof(float) * 4 * size_t((cfg.max_batch + 63) / 64));  // 4 slots a 64-sample tile
    init_weights(seed);
}

Nrc::~Nrc() {
    if (stream) (void)hipStreamDestroy(stream);
}

void Nrc::init_weights(uint32_t seed) {
    std::vector<float> w(NRC_NPARAM, 0.0f);
    uint64_t st = 0x6E7263ull ^ (uint64_t(seed) << 20);
    for (int l = 0; l < 6; ++l) {
        const int fan_in = l == 0 ? NRC_IN + 1 : 64;
        const int rows = l < 5 ? 64 : 3, cols = l == 0 ? NRC_IN + 1 : 64;
        const double bound = std::sqrt(6.0 / fan_in) * (l == 5 ? 0.1 : 1.0);
        for (int o = 0; o < rows; ++o)
            for (int i = 0; i < cols; ++i) {
                const double u = double(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
                w[nrc_layer_offset(l) + o * 64 + i] = float((2.0 * u - 1.0) * bound);
            }
    }
    set_weights(w.data());
}

void Nrc::set_weights(const float* w) {
    check_hip(hipMemcpyAsync(w32.ptr, w, sizeof(float) * NRC_NPARAM, hipMemcpyHostToDevice, stream), "nrc weights");
    check_hip(hipMemsetAsync(m.ptr, 0, sizeof(float) * NRC_NPARAM, stream), "nrc adam m");
    check_hip(hipMemsetAsync(v.ptr, 0, sizeof(float) * NRC_NPARAM, stream), "nrc adam v");
    check_hip(hipMemsetAsync(wema.ptr, 0, sizeof(float) * NRC_NPARAM, stream), "nrc ema");
    k_to_bf16<<<(NRC_NPARAM + 255) / 256, 256, 0, stream>>>(w32.as<float>(), w16.as<__bf16>(), w16t.as<__bf16>(),
                                                            NRC_NPARAM);
    check_hip(hipMemcpyAsync(w16inf.ptr, w16.ptr, sizeof(__bf16) * NRC_NPARAM, hipMemcpyDeviceToDevice, stream),
              "nrc inference weights");
    check_hip(hipGetLastError(), "k_to_bf16");
    check_hip(hipStreamSynchronize(stream), "nrc set_weights");
    step = 0;
}

void Nrc::get_weights(float* w) {
    check_hip(hipMemcpyAsync(w, w32.ptr, sizeof(float) * NRC_NPARAM, hipMemcpyDeviceToHost, stream), "nrc weights");
    check_hip(hipStreamSynchronize(stream), "nrc get_weights");
}

void Nrc::infer(const NrcQuery* q, int n, float4* out, hipStream_t s) {
    if (!s) s = stream;
    if (n <= 0) return;
    const int chunks = (n + 63) / 64;
    const int blocks = std::min((chunks + INF_WAVES - 1) / INF_WAVES, n_cus(device));
    k_nrc_infer<false><<<blocks, INF_WAVES * 64, 0, s>>>(w16inf.as<__bf16>(), q, nullptr, n, out);
    check_hip(hipGetLastError(), "k_nrc_infer");
}

void Nrc::infer_accumulate(const NrcQuery* q, const int* count, int max_n, float4* radiance, hipStream_t s) {
    if (!s) s = stream;
    if (max_n <= 0) return;
    const int chunks = (max_n + 63) / 64;
    const int blocks = std::min((chunks + INF_WAVES - 1) / INF_WAVES, n_cus(device));
    k_nrc_infer<true><<<blocks, INF_WAVES * 64, 0, s>>>(w16inf.as<__bf16>(), q, count, max_n, radiance);
    check_hip(hipGetLastError(), "k_nrc_infer<accumulate>");
}

int Nrc::train_blocks(int n) const {
    // two 64-sample tiles per block, at most one block per CU (144 KB of LDS): each block keeps a
    // full 84 KB weight-gradient partial in registers and writes it once
    const int tiles = (n + 63) / 64;
    return std::max(1, std::min(TRAIN_MAX_BLOCKS, (tiles + 1) / 2));
}

double Nrc::train(const NrcQuery* q, const float4* target, int n, bool want_loss, float* grad_out, hipStream_t s) {
    if (!s) s = stream;
    if (n <= 0) return 0.0;
    if (n > cfg.max_batch) throw std::runtime_error("nrc: training batch larger than max_batch");
    const int chunks = (n + 63) / 64;
    const int blocks = train_blocks(n);
    k_nrc_train<<<blocks, FT_WAVES * 64, 0, s>>>(w16.as<__bf16>(), w16t.as<__bf16>(), q, target, n, 1.0f / float(n),
                                                 part.as<float>(), lossp.as<float>());
    check_hip(hipGetLastError(), "k_nrc_train");
    DevBuf gtmp;
    float* gdev = nullptr;
    if (grad_out) {
        gtmp.ensure(sizeof(float) * NRC_NPARAM);
        gdev = gtmp.as<float>();
    } else {
        ++step;
    }
    const float c1 = 1.0f - std::pow(cfg.beta1, float(std::max(step, 1)));
    const float c2 = 1.0f - std::pow(cfg.beta2, float(std::max(step, 1)));
    const float ce = 1.0f - std::pow(cfg.ema, float(std::max(step, 1)));
    k_nrc_adam<<<(NRC_NPARAM + ADAM_PARAMS - 1) / ADAM_PARAMS, ADAM_PARAMS * ADAM_SLICES, 0, s>>>(w32.as<float>(), w16.as<__bf16>(), w16t.as<__bf16>(),
                                                            m.as<float>(),
                                                            v.as<float>(), wema.as<float>(), w16inf.as<__bf16>(),
                                                            part.as<float>(), blocks, cfg.lr, cfg.beta1,
                                                            cfg.beta2, cfg.eps, c1, c2, cfg.ema, ce, gdev);
    check_hip(hipGetLastError(), "k_nrc_adam");
    double loss = 0.0;
    if (grad_out) {
        check_hip(hipMemcpyAsync(grad_out, gdev, sizeof(float) * NRC_NPARAM, hipMemcpyDeviceToHost, s), "grad");
    }
    if (want_loss || grad_out) {
        std::vector<float> lp(4 * size_t(chunks));
        check_hip(hipMemcpyAsync(lp.data(), lossp.ptr, sizeof(float) * lp.size(), hipMemcpyDeviceToHost, s), "loss");
        check_hip(hipStreamSynchronize(s), "nrc train");
        for (float x : lp) loss += double(x);
        loss /= n;
    }
    return loss;
}

}  // namespace nr
