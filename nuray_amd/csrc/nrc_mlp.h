// Neural radiance cache: query record, input encoding and the wave-level BF16 MFMA building
// blocks of the 44 -> 64 x 5 (ReLU) -> 3 MLP.  Shared by the cache kernels and the renderer.
//
// Network (weight layers 0..5, no biases except a constant-1 input feature):
//   layer 0: 64 x 64 (inputs 0..43 = encoding, 44 = 1, 45..63 = 0), ReLU
//   layers 1..4: 64 x 64, ReLU
//   layer 5: 16 x 64 (outputs 0..2 = RGB radiance, 3..15 kept at zero), linear
// Parameters are stored per layer as W^T (row = output unit, 64 inputs per row), layers
// contiguous: 5 * 4096 + 16 * 64 = 21504 values.
//
// Matrix products use v_mfma_f32_16x16x32_bf16 in the transposed form Z^T = W^T X^T, with
// samples on the N side: operand A (16 x 32) = 16 output rows x 32 inputs of W^T, operand B
// (32 x 16) = 32 inputs x 16 samples of X^T.  Lane l supplies row (l & 15) and k = 8 (l >> 4) .. +7
// of A, column (l & 15) and the same k of B, and receives D rows 4 (l >> 4) .. +3 of column
// (l & 15).  With activations stored sample-major ([sample][feature], padded rows), every operand
// fetch is one 16-byte LDS read and every result is one 8-byte LDS write.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nr {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int NRC_IN = 44;            // encoded features
constexpr int NRC_F = 64;             // padded width
constexpr int NRC_LD = 72;            // LDS row stride in bf16 (144 B: rows start on different banks)
constexpr int NRC_OUTP = 16;          // padded output rows of layer 5
constexpr int NRC_LAYERS = 6;
constexpr int NRC_NPARAM = 5 * 64 * 64 + NRC_OUTP * 64;
constexpr int NRC_WROWS = 5 * 64 + NRC_OUTP;           // rows of W^T over all layers
constexpr int NRC_LDS_W = NRC_WROWS * NRC_LD;          // bf16 elements of the weights in LDS
constexpr int NRC_TILE = 64 * NRC_LD;                  // one [64 samples][72] activation tile

__host__ __device__ constexpr int nrc_layer_offset(int l) { return l * 64 * 64; }
__host__ __device__ constexpr int nrc_layer_rows(int l) { return l < 5 ? 64 : NRC_OUTP; }

// One cache query (64 B).  Producers fill a, b, c; the renderer also fills slot and beta.
struct alignas(16) NrcQuery {
    float4 a;  // position xyz, GGX roughness alpha (0 for Lambert)
    float4 b;  // wo and normal in spherical coordinates scaled to [0, 1]: theta/pi, phi/(2pi) + 1/2
    float4 c;  // albedo rgb, path slot (int bits; -1: none)
    float4 d;  // path throughput rgb, unused
};

// Spherical coordinates of a unit vector, scaled to [0, 1] (theta from +y, phi = atan2(z, x)).
__host__ __device__ inline void nrc_spherical(float x, float y, float z, float& th, float& ph) {
    const float yc = fminf(fmaxf(y, -1.0f), 1.0f);
    th = acosf(yc) * 0.31830988618f;
    ph = atan2f(z, x) * 0.15915494309f + 0.5f;
}

// Triangle wave of the reference's position encoder (renderernrc.cpp:50-53): 2|fmod(x, 2) - 1| - 1
// at x / 2^i, i = 0..11, with C fmod (sign of x) and the floating-point absolute value.
// x - 2 trunc(x / 2) equals fmod(x, 2) exactly: x / 2 and the product by 2 are exact, and the
// difference is exact because it is smaller than 2 and a multiple of ulp(x).
__host__ __device__ inline float nrc_tri(float x) {
    const float r = x - 2.0f * truncf(x * 0.5f);
    return 2.0f * fabsf(r - 1.0f) - 1.0f;
}

// The 64 input features of a query (fp32): 0..35 position (axis-major, 12 frequencies each),
// 36..39 wo / normal spherical, 40 exp(-alpha) (renderernrc.cpp:115), 41..43 albedo, 44 = 1.
__host__ __device__ inline float nrc_feature(const NrcQuery& q, int f) {
    if (f < 36) {
        const int ax = f / 12, i = f % 12;
        const float p = ax == 0 ? q.a.x : (ax == 1 ? q.a.y : q.a.z);
        return nrc_tri(p * (1.0f / float(1 << i)));
    }
    switch (f) {
        case 36: return q.b.x;
        case 37: return q.b.y;
        case 38: return q.b.z;
        case 39: return q.b.w;
        case 40: return expf(-q.a.w);
        case 41: return q.c.x;
        case 42: return q.c.y;
        case 43: return q.c.z;
        case 44: return 1.0f;
        default: return 0.0f;
    }
}

#ifdef __HIP_DEVICE_COMPILE__
#define NRC_DEVICE __device__ __forceinline__
#else
#define NRC_DEVICE __device__ inline
#endif

// Ordering of this wave's LDS accesses across its lanes.
NRC_DEVICE void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

NRC_DEVICE f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x2 __attribute__((ext_vector_type(2)));

// Round-to-nearest-even bf16 of two floats as one packed conversion, viewed as int16 pairs.
// (Written pairwise: the compiler lowers a four-wide conversion one element at a time.)
NRC_DEVICE s16x2 bf16_pair(float a, float b) {
    return __builtin_bit_cast(s16x2, __builtin_convertvector(f32x2{a, b}, bf16x2));
}
NRC_DEVICE bf16x4 join_pairs(s16x2 lo, s16x2 hi) {
    return __builtin_bit_cast(bf16x4, u32x2{__builtin_bit_cast(unsigned, lo), __builtin_bit_cast(unsigned, hi)});
}

// ReLU of the bf16-rounded values: a bf16 is negative (or -0) exactly when its bits are negative
// as int16, so the packed integer max with 0 clears those (the same values as rounding max(c, 0)).
NRC_DEVICE bf16x4 relu_bf16(f32x4 c) {
    return join_pairs(__builtin_elementwise_max(bf16_pair(c[0], c[1]), s16x2{0, 0}),
                      __builtin_elementwise_max(bf16_pair(c[2], c[3]), s16x2{0, 0}));
}

// The bf16-rounded c where the (non-negative) activation x is > 0, else 0: min(x, 1) negated is
// an all-ones or all-zeros 16-bit mask.  (Kept four-wide: applied to one 32-bit half of x at a
// time, this compiler selects the packed min on the wrong half.)
NRC_DEVICE bf16x4 masked_bf16(f32x4 c, bf16x4 x) {
    const s16x4 m = -__builtin_elementwise_min(__builtin_bit_cast(s16x4, x), s16x4{1, 1, 1, 1});
    const bf16x4 g = join_pairs(bf16_pair(c[0], c[1]), bf16_pair(c[2], c[3]));
    return __builtin_bit_cast(bf16x4, __builtin_bit_cast(s16x4, g) & m);
}

// Block-cooperative copy of the bf16 weights (global, dense rows of 64) into padded LDS rows.
NRC_DEVICE void nrc_load_weights(const __bf16* __restrict__ w, __bf16* lds) {
    for (int i = threadIdx.x; i < NRC_WROWS * 8; i += blockDim.x) {
        const int r = i >> 3, c = i & 7;
        *reinterpret_cast<uint4*>(lds + r * NRC_LD + c * 8) = *reinterpret_cast<const uint4*>(w + r * 64 + c * 8);
    }
}

// Half-width input tile of the inference kernel: 64 samples x 32 features, 80-byte rows (their
// starting banks differ by 20 dwords, so 16-byte accesses of 16 lanes hit all 64 banks once).
constexpr int NRC_LDH = 40;
constexpr int NRC_TILE_HALF = 64 * NRC_LDH;

// Lane `lane` encodes features 32 h .. 32 h + 31 of one query into row `lane` of the half tile X
// (zeros when !valid).
NRC_DEVICE void nrc_encode_half(const NrcQuery& q, bool valid, __bf16* X, int lane, int h) {
    __bf16* row = X + lane * NRC_LDH;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        bf16x8 v;
#pragma unroll
        for (int j = 0; j < 8; ++j) v[j] = (__bf16)(valid ? nrc_feature(q, 32 * h + c * 8 + j) : 0.0f);
        *reinterpret_cast<bf16x8*>(row + c * 8) = v;
    }
}

// A fragment of hidden or output layer l >= 1 for lane (r, qd) and K chunk ks: row `row` of W^T,
// input features 32 ks + 4 qd + {0..3} and 32 ks + 16 + 4 qd + {0..3} -- the order in which the
// previous layer's output tiles 2 ks and 2 ks + 1 hold them in the lane's registers.
NRC_DEVICE bf16x8 nrc_wfrag_perm(const __bf16* Wl, int row, int ks, int qd) {
    const __bf16* p = Wl + row * NRC_LD + 32 * ks + 4 * qd;
    const bf16x4 lo = *reinterpret_cast<const bf16x4*>(p);
    const bf16x4 hi = *reinterpret_cast<const bf16x4*>(p + 16);
    return __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
}

// Full forward pass of 64 samples whose encoded inputs are the layer-0 B operands b (lane (r, qd):
// features 32 ks + 8 qd .. + 7 of sample 16 nt + r in b[nt][ks]).  The activations stay in registers: output tile mt of a 16x16x32 product leaves features
// 16 mt + 4 qd .. + 3 of sample 16 nt + r in lane (r, qd), so the lane's 8 K values of the next
// product's B operand for chunk ks are tiles 2 ks and 2 ks + 1 side by side -- the K order of
// nrc_wfrag_perm, which reads the weights in that order.  Returns the RGB output of sample
// 16 nt + (lane & 15) in out[nt] for lanes with lane < 16.
NRC_DEVICE void nrc_forward(const __bf16* W, bf16x8 (&b)[4][2], float4 (&out)[4], int lane) {
    const int r = lane & 15, qd = lane >> 4;
#pragma unroll 1
    for (int l = 0; l < 5; ++l) {
        const __bf16* Wl = W + l * 64 * NRC_LD;
        f32x4 acc[4][4];
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            bf16x8 a0, a1;
            if (l == 0) {  // the encoded inputs are in natural order
                a0 = *reinterpret_cast<const bf16x8*>(Wl + (16 * mt + r) * NRC_LD + 8 * qd);
                a1 = *reinterpret_cast<const bf16x8*>(Wl + (16 * mt + r) * NRC_LD + 32 + 8 * qd);
            } else {
                a0 = nrc_wfrag_perm(Wl, 16 * mt + r, 0, qd);
                a1 = nrc_wfrag_perm(Wl, 16 * mt + r, 1, qd);
            }
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
                c = mfma16(a0, b[nt][0], c);
                acc[mt][nt] = mfma16(a1, b[nt][1], c);
            }
        }
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
#pragma unroll
            for (int ks = 0; ks < 2; ++ks)
                b[nt][ks] = __builtin_shufflevector(relu_bf16(acc[2 * ks][nt]), relu_bf16(acc[2 * ks + 1][nt]), 0, 1, 2,
                                                    3, 4, 5, 6, 7);
    }
    const __bf16* W5 = W + 5 * 64 * NRC_LD;
    const bf16x8 a0 = nrc_wfrag_perm(W5, r, 0, qd), a1 = nrc_wfrag_perm(W5, r, 1, qd);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
        f32x4 c = {0.0f, 0.0f, 0.0f, 0.0f};
        c = mfma16(a0, b[nt][0], c);
        c = mfma16(a1, b[nt][1], c);
        out[nt] = make_float4(c[0], c[1], c[2], 0.0f);
    }
}

}  // namespace nr
