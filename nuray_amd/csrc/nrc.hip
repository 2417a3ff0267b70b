// Neural radiance cache kernels for gfx950: inference, training forward + backward, split-K weight
// gradient and Adam, all matrix work on BF16 MFMA with FP32 accumulation.
//
// The reference's cache (NeuralRadianceCache / NNNode*, renderernrc.h:17-43, nnnode.cpp:1-208) is a
// per-sample scalar MLP trained by SGD; this one is the 44 -> 5 x 64 (ReLU) -> 3 network trained
// with Adam on batches of 2^16 samples and the relative L2 loss
//   L = mean_s sum_c (y_sc - t_sc)^2 / (y_sc^2 + 0.01)      (y in the denominator held constant)
// whose output gradient has the reference's form (o - a) / (o^2 + 0.01) (nnnode.cpp:195-196).
//
// Kernels
//   k_nrc_infer      persistent; weights in LDS once per block (12 waves: 3 per SIMD, the most
//                    the 138 VGPRs allow), one 64-sample tile per wave
//   k_nrc_fwd_bwd    one 64-sample tile per block of 4 waves, each wave 16 features of every layer:
//                    forward, loss gradient, backward (ReLU masks kept in registers), weights read
//                    from L2 (W^T for the forward, the transposed copy W for the backward), two
//                    LDS tiles (18 KB a block, so many blocks share a CU and hide each other's
//                    latency); writes each layer's input X_l and pre-activation gradient G_l (bf16,
//                    [sample][64]) for the weight gradient, and the tile's loss to its own slot (no
//                    atomics: a device-scope atomic per tile on one address serialises the grid
//                    across the XCDs)
//   k_nrc_wgrad      dW_l^T = sum_s G_l[s]^T X_l[s]: K = samples, split over (split, layer)
//                    blocks of 4 waves that prefetch their next 32-sample chunk; the 4 waves'
//                    partials are summed in LDS, one partial per block (no atomics)
//   k_nrc_adam       sums the blocks' partials (independent loads, 8 in flight) and applies Adam
//                    to the FP32 master weights, refreshes the BF16 copies
#include <cmath>
#include <cstring>
#include <type_traits>
#include <stdexcept>
#include <vector>
#include "nrc.h"

namespace nr {

namespace {

// waves per inference block: 48 KB of weights + 12 x 9 KB tiles = 155 KB of LDS, one block per CU,
// 3 waves per SIMD (the 138 VGPRs of the kernel allow 3)
constexpr int INF_WAVES = 12;

template <bool ACC>
__global__ __launch_bounds__(INF_WAVES * 64) void k_nrc_infer(const __bf16* __restrict__ w16,
                                                              const NrcQuery* __restrict__ q,
                                                              const int* __restrict__ count, int n_max,
                                                              float4* __restrict__ out) {
    __shared__ __attribute__((aligned(16))) __bf16 sW[NRC_LDS_W];
    __shared__ __attribute__((aligned(16))) __bf16 sX[INF_WAVES][NRC_TILE];
    nrc_load_weights(w16, sW);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = count ? min(*count, n_max) : n_max;
    __bf16* X = sX[wave];
    for (int chunk = blockIdx.x * INF_WAVES + wave; chunk * 64 < n; chunk += gridDim.x * INF_WAVES) {
        const int s = chunk * 64 + lane;
        NrcQuery rec;
        const bool valid = s < n;
        if (valid) rec = q[s];
        else rec = NrcQuery{};
        nrc_encode_row(rec, valid, X, lane);
        wave_sync();
        float4 y[4];
        nrc_forward(sW, X, y, lane);
        wave_sync();
        if (lane < 16) {
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const int si = chunk * 64 + 16 * nt + lane;
                if (si >= n) continue;
                if (ACC) {
                    const float4 qc = q[si].c, qd = q[si].d;
                    const int slot = __float_as_int(qc.w);
                    if (slot < 0) continue;
                    float4 r = out[slot];
                    r.x += qd.x * fmaxf(y[nt].x, 0.0f);
                    r.y += qd.y * fmaxf(y[nt].y, 0.0f);
                    r.z += qd.z * fmaxf(y[nt].z, 0.0f);
                    out[slot] = r;
                } else {
                    out[si] = y[nt];
                }
            }
        }
    }
}

// The 16 features [16 w, 16 w + 16) of the 64 rows of tile T to global rows [s0, s0 + 64) of a
// [batch][64] bf16 array (lane = row; two 16-byte stores).
__device__ __forceinline__ void slice_to_global(const __bf16* T, __bf16* g, int s0, int n, int w, int lane) {
    const int s = s0 + lane;
    if (s >= n) return;
    const uint4* src = reinterpret_cast<const uint4*>(T + lane * NRC_LD + 16 * w);
    uint4* dst = reinterpret_cast<uint4*>(g + size_t(s) * 64 + 16 * w);
    dst[0] = src[0];
    dst[1] = src[1];
}

// Forward, loss gradient and backward of one 64-sample tile per block of 4 waves.  Wave w computes
// output features [16 w, 16 w + 16) of every layer (the 16-row output layer: wave 0) and, in the
// backward, input features [16 w, 16 w + 16): exactly the features whose ReLU masks it kept from
// the forward (a lane's backward results sit at the fragment positions of its forward results).
// Activations ping-pong between two LDS tiles, one block barrier per layer; the weights come from
// global memory (L2-resident): W^T rows (w16) as the forward's A operand, W rows (w16t, kept beside
// them by Adam) as the backward's.  Layer inputs X_l and pre-activation gradients G_l go to global
// memory for the weight gradient; loss[tile] = the tile's summed relative L2 loss.
constexpr int FB_WAVES = 4;

// The A operand fragments of one weight layer for lane (r, qd): row `row` of a [64][64] bf16 matrix,
// k = 8 qd .. 8 qd + 7 and 32 + 8 qd ...
struct AFrag {
    bf16x8 a0, a1;
};
__device__ __forceinline__ AFrag load_afrag(const __bf16* m, int row, int qd) {
    const __bf16* p = m + row * 64 + 8 * qd;
    return AFrag{*reinterpret_cast<const bf16x8*>(p), *reinterpret_cast<const bf16x8*>(p + 32)};
}

__global__ __launch_bounds__(FB_WAVES * 64) void k_nrc_fwd_bwd(const __bf16* __restrict__ w16,
                                                               const __bf16* __restrict__ w16t,
                                                               const NrcQuery* __restrict__ q,
                                                               const float4* __restrict__ tgt, int n, float inv_n,
                                                               __bf16* __restrict__ act, __bf16* __restrict__ gz,
                                                               float* __restrict__ loss) {
    __shared__ __attribute__((aligned(16))) __bf16 sT[2][NRC_TILE];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, r = lane & 15, qd = lane >> 4;
    const int s0 = blockIdx.x * 64;
    const size_t ls = size_t(n) * 64;
    // Loads are issued one layer ahead: on gfx950 the vector-memory counter covers stores as well,
    // so a weight load issued after a layer's activation stores would make its wait cover them too.
    AFrag nxt = load_afrag(w16, 16 * w + r, qd);
    float4 tv[4];
    if (w == 0 && qd == 0) {
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const int s = s0 + 16 * nt + r;
            tv[nt] = s < n ? tgt[s] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        }
    }
    // encode: wave w writes features [16 w, 16 w + 16) of the lane's sample
    {
        const int s = s0 + lane;
        const bool valid = s < n;
        const NrcQuery rec = valid ? q[s] : NrcQuery{};
        bf16x8 v[2];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int j = 0; j < 8; ++j) v[h][j] = (__bf16)(valid ? nrc_feature(rec, 16 * w + 8 * h + j) : 0.0f);
        *reinterpret_cast<bf16x8*>(sT[0] + lane * NRC_LD + 16 * w) = v[0];
        *reinterpret_cast<bf16x8*>(sT[0] + lane * NRC_LD + 16 * w + 8) = v[1];
        if (valid) {
            *reinterpret_cast<bf16x8*>(act + size_t(s) * 64 + 16 * w) = v[0];
            *reinterpret_cast<bf16x8*>(act + size_t(s) * 64 + 16 * w + 8) = v[1];
        }
    }
    __syncthreads();
    // forward; bit (4 nt + i) of mask[l]: output feature 16 w + 4 qd + i of sample 16 nt + r is > 0
    uint32_t mask[5];
    int cur = 0;
#pragma unroll
    for (int l = 0; l < 5; ++l) {
        const AFrag A = nxt;
        // next: layer l + 1 (the output layer's 16 rows after layer 4: wave 0 only, others idle)
        nxt = load_afrag(w16 + (l + 1) * 4096, l < 4 ? 16 * w + r : r, qd);
        const __bf16* X = sT[cur];
        __bf16* Y = sT[cur ^ 1];
        uint32_t m = 0;
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const bf16x8 b0 = *reinterpret_cast<const bf16x8*>(X + (16 * nt + r) * NRC_LD + 8 * qd);
            const bf16x8 b1 = *reinterpret_cast<const bf16x8*>(X + (16 * nt + r) * NRC_LD + 32 + 8 * qd);
            f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
            c = mfma16(A.a0, b0, c);
            c = mfma16(A.a1, b1, c);
            bf16x4 y;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                y[i] = (__bf16)fmaxf(c[i], 0.0f);
                if ((float)y[i] > 0.0f) m |= 1u << (4 * nt + i);
            }
            *reinterpret_cast<bf16x4*>(Y + (16 * nt + r) * NRC_LD + 16 * w + 4 * qd) = y;
        }
        mask[l] = m;
        wave_sync();
        slice_to_global(Y, act + (l + 1) * ls, s0, n, w, lane);
        __syncthreads();
        cur ^= 1;
    }
    // output layer (rows 0..15 of layer 5, wave 0) and its gradient G_5 (columns 0..2 of the
    // other tile; columns 3..31 zero: the backward's K of layer 5 spans 32 outputs)
    const AFrag A5 = nxt;
    nxt = load_afrag(w16t + 5 * 4096, 16 * w + r, qd);  // the backward's first layer
    __bf16* G = sT[cur ^ 1];
    if (w == 0) {
        const __bf16* X = sT[cur];
        float loss_acc = 0.0f;
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const bf16x8 b0 = *reinterpret_cast<const bf16x8*>(X + (16 * nt + r) * NRC_LD + 8 * qd);
            const bf16x8 b1 = *reinterpret_cast<const bf16x8*>(X + (16 * nt + r) * NRC_LD + 32 + 8 * qd);
            f32x4 y = {0.0f, 0.0f, 0.0f, 0.0f};
            y = mfma16(A5.a0, b0, y);
            y = mfma16(A5.a1, b1, y);
            const int s = s0 + 16 * nt + r;
            bf16x4 gv = {(__bf16)0.0f, (__bf16)0.0f, (__bf16)0.0f, (__bf16)0.0f};
            if (qd == 0 && s < n) {
                const float t3[3] = {tv[nt].x, tv[nt].y, tv[nt].z};
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float den = y[c] * y[c] + 0.01f, e = y[c] - t3[c];
                    loss_acc += e * e / den;
                    gv[c] = (__bf16)(2.0f * e / den * inv_n);
                }
            }
            const bf16x4 z = {(__bf16)0.0f, (__bf16)0.0f, (__bf16)0.0f, (__bf16)0.0f};
            *reinterpret_cast<bf16x4*>(G + (16 * nt + r) * NRC_LD + 4 * qd) = gv;
            *reinterpret_cast<bf16x4*>(G + (16 * nt + r) * NRC_LD + 16 + 4 * qd) = z;
        }
        for (int off = 32; off > 0; off >>= 1) loss_acc += __shfl_xor(loss_acc, off);
        if (lane == 0) loss[blockIdx.x] = loss_acc;
        wave_sync();
        slice_to_global(G, gz + 5 * ls, s0, n, 0, lane);
    }
    __syncthreads();
    // backward: G_{l-1} = (W_l^T G_l) * [Z_{l-1} > 0] for this wave's 16 input features
    int gc = cur ^ 1;
#pragma unroll
    for (int l = 5; l >= 1; --l) {
        const AFrag A = nxt;
        if (l > 1) nxt = load_afrag(w16t + (l - 1) * 4096, 16 * w + r, qd);
        const __bf16* Gl = sT[gc];
        __bf16* Gp = sT[gc ^ 1];
        const uint32_t m = mask[l - 1];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            const bf16x8 b0 = *reinterpret_cast<const bf16x8*>(Gl + (16 * nt + r) * NRC_LD + 8 * qd);
            f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
            c = mfma16(A.a0, b0, c);
            if (l < 5) {
                const bf16x8 b1 = *reinterpret_cast<const bf16x8*>(Gl + (16 * nt + r) * NRC_LD + 32 + 8 * qd);
                c = mfma16(A.a1, b1, c);
            }
            bf16x4 gv;
#pragma unroll
            for (int i = 0; i < 4; ++i) gv[i] = (__bf16)(((m >> (4 * nt + i)) & 1u) ? c[i] : 0.0f);
            *reinterpret_cast<bf16x4*>(Gp + (16 * nt + r) * NRC_LD + 16 * w + 4 * qd) = gv;
        }
        wave_sync();
        slice_to_global(Gp, gz + (l - 1) * ls, s0, n, w, lane);
        __syncthreads();
        gc ^= 1;
    }
}

// Weight gradient of layer l over the samples of this wave's chunks of 32 (chunks first, first +
// stride, ...): acc[mt][nt] += G^T X over the chunk.  The next chunk's 8 KB are loaded into
// registers while the MFMAs run on the current one (staged through LDS for the transposed reads).
template <int MT>
__device__ __forceinline__ void wgrad_body(const __bf16* __restrict__ X, const __bf16* __restrict__ G, int n,
                                           int first, int stride, __bf16* sG, __bf16* sX, f32x4 (&acc)[MT][4],
                                           int lane) {
    const int r = lane & 15, qd = lane >> 4;
#pragma unroll
    for (int mt = 0; mt < MT; ++mt)
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = f32x4{0.0f, 0.0f, 0.0f, 0.0f};
    // 32 rows of G and X: 256 16-byte pieces each, 4 per lane
    uint4 g[4], x[4];
    auto fetch = [&](int c) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int piece = lane + 64 * k, row = piece >> 3, col = piece & 7;
            const int s = c * 32 + row;
            g[k] = make_uint4(0, 0, 0, 0);
            x[k] = make_uint4(0, 0, 0, 0);
            if (s < n) {
                g[k] = reinterpret_cast<const uint4*>(G + size_t(s) * 64)[col];
                x[k] = reinterpret_cast<const uint4*>(X + size_t(s) * 64)[col];
            }
        }
    };
    if (first * 32 < n) fetch(first);
    for (int c0 = first; c0 * 32 < n; c0 += stride) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int piece = lane + 64 * k, row = piece >> 3, col = piece & 7;
            *reinterpret_cast<uint4*>(sG + row * NRC_LD + col * 8) = g[k];
            *reinterpret_cast<uint4*>(sX + row * NRC_LD + col * 8) = x[k];
        }
        wave_sync();
        if ((c0 + stride) * 32 < n) fetch(c0 + stride);
        bf16x8 b[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int j = 0; j < 8; ++j) b[nt][j] = sX[(8 * qd + j) * NRC_LD + 16 * nt + r];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            bf16x8 a;
#pragma unroll
            for (int j = 0; j < 8; ++j) a[j] = sG[(8 * qd + j) * NRC_LD + 16 * mt + r];
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = mfma16(a, b[nt], acc[mt][nt]);
        }
        wave_sync();
    }
}

// grid (splits, 6), 4 waves per block: block (s, l) sums layer l's gradient over the chunks
// s * 4 + w + k * 4 S of its waves w and writes it to part[(l * S + s) * 4096 ..] (rows of W^T).
__global__ __launch_bounds__(256) void k_nrc_wgrad(const __bf16* __restrict__ act, const __bf16* __restrict__ gz,
                                                   int n, float* __restrict__ part) {
    __shared__ __attribute__((aligned(16))) __bf16 sG[4][32 * NRC_LD];
    __shared__ __attribute__((aligned(16))) __bf16 sX[4][32 * NRC_LD];
    __shared__ float sP[4096];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, l = blockIdx.y;
    const int S = gridDim.x, s = blockIdx.x;
    const int r = lane & 15, qd = lane >> 4;
    const size_t ls = size_t(n) * 64;
    float* dst = part + (size_t(l) * S + s) * 4096;
    // the four waves' partials are added into sP one wave after the other
    auto reduce = [&](auto& acc, auto mt_count) {
        constexpr int MT = decltype(mt_count)::value;
        for (int w = 0; w < 4; ++w) {
            if (wave == w) {
#pragma unroll
                for (int mt = 0; mt < MT; ++mt)
#pragma unroll
                    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
                        for (int i = 0; i < 4; ++i) {
                            float& o = sP[(16 * mt + 4 * qd + i) * 64 + 16 * nt + r];
                            o = w == 0 ? acc[mt][nt][i] : o + acc[mt][nt][i];
                        }
            }
            __syncthreads();
        }
        for (int i = threadIdx.x; i < MT * 16 * 64; i += 256) dst[i] = sP[i];
    };
    if (l < 5) {
        f32x4 acc[4][4];
        wgrad_body<4>(act + l * ls, gz + l * ls, n, s * 4 + wave, S * 4, sG[wave], sX[wave], acc, lane);
        reduce(acc, std::integral_constant<int, 4>{});
    } else {
        f32x4 acc[1][4];
        wgrad_body<1>(act + l * ls, gz + l * ls, n, s * 4 + wave, S * 4, sG[wave], sX[wave], acc, lane);
        reduce(acc, std::integral_constant<int, 1>{});
    }
}

__global__ void k_nrc_adam(float* __restrict__ w32, __bf16* __restrict__ w16, __bf16* __restrict__ w16t,
                           float* __restrict__ m,
                           float* __restrict__ v, float* __restrict__ wema, __bf16* __restrict__ w16inf,
                           const float* __restrict__ part, int P, float lr, float b1, float b2, float eps, float c1,
                           float c2, float ema, float ce, float* __restrict__ grad_out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= NRC_NPARAM) return;
    const int l = i < 5 * 4096 ? i / 4096 : 5, k = i - l * 4096;
    const float* src = part + size_t(l) * P * 4096 + k;
    float g = 0.0f;
    int p = 0;
    for (; p + 8 <= P; p += 8) {
        float t[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) t[j] = src[size_t(p + j) * 4096];
#pragma unroll
        for (int j = 0; j < 8; ++j) g += t[j];
    }
    for (; p < P; ++p) g += src[size_t(p) * 4096];
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
    act.ensure(sizeof(__bf16) * 6 * 64 * size_t(cfg.max_batch));
    gz.ensure(sizeof(__bf16) * 6 * 64 * size_t(cfg.max_batch));
    part.ensure(sizeof(float) * 6 * size_t(64) * 4096);  // at most 64 splits (blocks) per layer
    lossp.ensure(sizeof(float) * size_t((cfg.max_batch + 63) / 64));
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

int Nrc::splits_for(int n) const {
    // blocks per layer: at least 2 chunks of 32 samples per wave, at most 64 (6 x 64 = 384 blocks)
    const int chunks = (n + 31) / 32;
    return std::max(1, std::min(64, chunks / 8));
}

double Nrc::train(const NrcQuery* q, const float4* target, int n, bool want_loss, float* grad_out, hipStream_t s) {
    if (!s) s = stream;
    if (n <= 0) return 0.0;
    if (n > cfg.max_batch) throw std::runtime_error("nrc: training batch larger than max_batch");
    const int chunks = (n + 63) / 64;
    k_nrc_fwd_bwd<<<chunks, FB_WAVES * 64, 0, s>>>(w16.as<__bf16>(), w16t.as<__bf16>(), q, target, n, 1.0f / float(n),
                                                  act.as<__bf16>(), gz.as<__bf16>(), lossp.as<float>());
    check_hip(hipGetLastError(), "k_nrc_fwd_bwd");
    const int splits = splits_for(n);
    k_nrc_wgrad<<<dim3(splits, 6), 256, 0, s>>>(act.as<__bf16>(), gz.as<__bf16>(), n, part.as<float>());
    check_hip(hipGetLastError(), "k_nrc_wgrad");
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
    k_nrc_adam<<<(NRC_NPARAM + 255) / 256, 256, 0, s>>>(w32.as<float>(), w16.as<__bf16>(), w16t.as<__bf16>(),
                                                            m.as<float>(),
                                                            v.as<float>(), wema.as<float>(), w16inf.as<__bf16>(),
                                                            part.as<float>(), splits, cfg.lr, cfg.beta1,
                                                            cfg.beta2, cfg.eps, c1, c2, cfg.ema, ce, gdev);
    check_hip(hipGetLastError(), "k_nrc_adam");
    double loss = 0.0;
    if (grad_out) {
        check_hip(hipMemcpyAsync(grad_out, gdev, sizeof(float) * NRC_NPARAM, hipMemcpyDeviceToHost, s), "grad");
    }
    if (want_loss || grad_out) {
        std::vector<float> lp(chunks);
        check_hip(hipMemcpyAsync(lp.data(), lossp.ptr, sizeof(float) * chunks, hipMemcpyDeviceToHost, s), "loss");
        check_hip(hipStreamSynchronize(s), "nrc train");
        for (float x : lp) loss += double(x);
        loss /= n;
    }
    return loss;
}

}  // namespace nr
