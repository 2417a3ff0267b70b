// Device-side records and inline helpers of the MI355X ptnee path.
//
// Layout in HBM (all arrays 64-byte aligned by hipMalloc):
//   Node4    128 B  the SAH tree collapsed 4 wide (boxes of up to four descendants, refs)
//   TriGeom   48 B  one per triangle, in DFS leaf order: p0, e1 = p1-p0, e2 = p2-p0, loader id
//   TriShade  64 B  same order: n0, n1, n2, uv0, uv1, uv2, material row
//   MatRec    48 B  one per material: {type, r, g, b, alpha, ior, texture, 0, Ks r, g, b, Ks texture}
//   LightRec 112 B  one per emissive triangle, loader order, with the area CDF beside it
//
// Every float step that decides a hit follows src/hierarchy/bvhnode.cpp and
// src/scene/triangle.cpp in order and precision; this header is compiled with
// -ffp-contract=off so no multiply-add is fused.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nr {

enum MatType { MAT_LAMBERT = 0, MAT_LIGHT = 1, MAT_GGXREFL = 2, MAT_GGX = 3, MAT_BLINNPHONG = 4, MAT_GLASS = 5 };

// Leaf references inside Node4::ref: bit 31 set, bits 27..30 = primitive count (1..5),
// bits 0..26 = first primitive in TriGeom order.  Internal references are Node4 indices.
constexpr uint32_t LEAF_BIT = 0x80000000u;
constexpr int FLAT_LEAVES = 64;  // scenes with at most this many leaves also get a flat leaf list
__host__ __device__ inline bool ref_is_leaf(int32_t r) { return (uint32_t(r) & LEAF_BIT) != 0; }
__host__ __device__ inline int ref_first(int32_t r) { return int(uint32_t(r) & 0x07FFFFFFu); }
__host__ __device__ inline int ref_count(int32_t r) { return int((uint32_t(r) >> 27) & 0xFu); }
__host__ __device__ inline int32_t make_leaf(int first, int count) {
    return int32_t(LEAF_BIT | (uint32_t(count) << 27) | uint32_t(first));
}

// The traversal kernels walk a 4-wide form of the binary SAH tree: each node holds up to four
// descendants of a binary node (opened largest surface area first) with their exact boxes, as
// per-axis float4s (child k in component k).  The leaves are the binary tree's own leaves, so every
// leaf box is tested with its exact bounds, and every other box contains the leaf boxes below it.
// Internal references are Node4 indices.  An unused slot holds a point box at 1e30 and an empty leaf
// (no triangles), so even a ray that reaches it changes nothing.
struct alignas(128) Node4 {
    float4 lox, loy, loz;  // child k: min x, y, z in component k
    float4 hix, hiy, hiz;  // child k: max x, y, z
    int4 ref;              // child references
    int4 pad;
};
struct alignas(16) TriGeom {
    float4 g0;  // p0.xyz e1.x
    float4 g1;  // e1.yz e2.xy
    float4 g2;  // e2.z, loader index (bits), 0, 0
};
struct alignas(16) TriShade {
    float4 s0;  // n0.xyz n1.x
    float4 s1;  // n1.yz n2.xy
    float4 s2;  // n2.z uv0.xy uv1.x
    float4 s3;  // uv1.y uv2.xy material (bits)
};
struct alignas(16) MatRec {
    float4 a;  // type, r, g, b
    float4 b;  // alpha (Blinn-Phong: Ns), ior, texture, 0
    float4 c;  // Blinn-Phong Ks r, g, b, Ks texture
};
struct alignas(16) LightRec {
    float4 p0, p1, p2;  // xyz used
    float4 n0, n1, n2;  // xyz used
    float4 ke;          // emission
};
struct TexDesc {
    int32_t w, h, fmt, pad;  // fmt 0: RGBA8 (sRGB-like pow 2.2 decode through a LUT), 1: RGB32F
    uint64_t offset;         // byte offset into the texture pool
};

// Everything a kernel needs to read the scene.
struct SceneView {
    const Node4* wide;  // the SAH tree, 4 wide
    const TriGeom* tris;
    const TriShade* shade;
    const MatRec* mats;
    const LightRec* lights;
    const float* light_cdf;
    const TexDesc* tex;
    const uint8_t* texpool;
    const float* lut;  // 256-entry RGBA8 decode table
    float root_min[3], root_max[3];
    int32_t n_lights;
    float light_p;
    int32_t env_tex;   // texture index of the environment map or -1
    float env_gain;
    int32_t short_stack; // stack entries of the first persistent trace launch (see trace.hip)
    int32_t wide_root;       // root as a Node4 index, or a leaf reference
    int32_t wide_stack_cap;  // full-depth stack entries of the 4-wide walk: 3 per level + 1
    int32_t lane_slots;      // LDS stack slots per lane of the per-lane walk (trace_ray): entries + spares
    // small scenes (n_flat > 0 leaves, at most FLAT_LEAVES): the leaves' exact boxes in DFS order,
    // two float4 per leaf {min xyz, leaf reference bits}, {max xyz, 0}, for the per-lane walks
    const float4* flat;
    int32_t n_flat;
    // environment sampler: one cell per texel square of the map (env_w x env_h cells, 0 if none),
    // {alias probability, alias cell (int bits), pdf of the cell, pdf of the alias cell}; the
    // solid-angle density of a direction in cell k is pdf_k * env_norm / sin(theta)
    const float4* env_tab;
    int32_t env_w, env_h;
    float env_norm;
};

// Queued ray: origin + tmax / shadow length, direction + payload.
struct alignas(16) Ray {
    float4 o;  // xyz, w = segment length for shadow rays (unused for closest hit)
    float4 d;  // xyz, w = path slot (bits)
};
struct alignas(16) HitRec {
    float t, b1, b2;
    int32_t prim;  // TriGeom index, -1 for a miss
};
// Vertex records kept by a render for the adjoint's replay: the closest hit of the vertex's ray as
// traced (prim -1: a miss); for a hit, bit 29 of prim is set when the vertex traced its NEE shadow
// ray and bit 30 when that ray found the light visible (primitive indices stay below 2^27).
constexpr int32_t VREC_TRACED = 1 << 29;
constexpr int32_t VREC_VISIBLE = 1 << 30;

// ------------------------------------------------------------------ exact hit arithmetic
struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 f3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ F3 sub(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 add(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 mul(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ F3 mulv(F3 a, F3 b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ F3 divs(F3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ F3 neg(F3 a) { return {-a.x, -a.y, -a.z}; }
// vec3::dot and vec3::cross, src/math/vec3.cpp:54-61 (left-to-right, unfused).
__device__ __forceinline__ float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 cross(F3 a, F3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float norm(F3 a) { return sqrtf(dot(a, a)); }
// vec3::normalized, vec3.cpp:70-73.
__device__ __forceinline__ F3 normalized(F3 a) { return divs(a, norm(a) + 1e-8f); }

// BVHNode::intersect, bvhnode.cpp:35-48.  dd = d + 1e-9f per axis (computed once per ray).
// std::max(a, b) is (a < b) ? b : a and std::min(a, b) is (b < a) ? b : a: NaN in the second
// operand leaves the first one.
__device__ __forceinline__ bool slab(float mnx, float mny, float mnz, float mxx, float mxy, float mxz,
                                     F3 o, F3 dd) {
    float t_in = -1e18f, t_out = 1e18f;
    {
        float t1 = (mnx - o.x) / dd.x, t2 = (mxx - o.x) / dd.x;
        if (t1 > t2) { float s = t1; t1 = t2; t2 = s; }
        t_in = (t_in < t1) ? t1 : t_in;
        t_out = (t2 < t_out) ? t2 : t_out;
    }
    {
        float t1 = (mny - o.y) / dd.y, t2 = (mxy - o.y) / dd.y;
        if (t1 > t2) { float s = t1; t1 = t2; t2 = s; }
        t_in = (t_in < t1) ? t1 : t_in;
        t_out = (t2 < t_out) ? t2 : t_out;
    }
    {
        float t1 = (mnz - o.z) / dd.z, t2 = (mxz - o.z) / dd.z;
        if (t1 > t2) { float s = t1; t1 = t2; t2 = s; }
        t_in = (t_in < t1) ? t1 : t_in;
        t_out = (t2 < t_out) ? t2 : t_out;
    }
    return t_in <= t_out && t_out >= 0.0f;
}

// Triangle::intersection, triangle.cpp:27-33; q rounds (double)(s1.e1) + 1e-6 to float.
__device__ __forceinline__ void tri_test(const TriGeom& g, F3 o, F3 d, float& t, float& b1, float& b2) {
    F3 p0 = f3(g.g0.x, g.g0.y, g.g0.z);
    F3 e1 = f3(g.g0.w, g.g1.x, g.g1.y);
    F3 e2 = f3(g.g1.z, g.g1.w, g.g2.x);
    F3 s = sub(o, p0), s1 = cross(d, e2), s2 = cross(s, e1);
    float tn = dot(s2, e2), b1n = dot(s1, s), b2n = dot(s2, d);
    float q = float(double(dot(s1, e1)) + 1e-6);
    t = tn / q;
    b1 = b1n / q;
    b2 = b2n / q;
}
// Leaf acceptance, bvhnode.cpp:57 (the running best is handled by the caller).
__device__ __forceinline__ bool tri_valid(float t, float b1, float b2) {
    return t > 0.0f && t < 1e18f && b1 >= 0.0f && b2 >= 0.0f && b1 + b2 <= 1.0f;
}

// ------------------------------------------------------------------ filtered exact tests
// The same decisions as slab() / tri_test()+tri_valid(), with most divisions replaced by a
// multiply and an error band.  Outside the band the cheap result provably equals the exact
// one; inside it the exact division path runs.  Hit / miss outcomes and hit values are
// therefore bit-identical to the unfiltered functions.
//
// Slab: per ray r = fl(1 / dd) with dd = d + 1e-9f (correctly rounded) and c = fl(o * r).  For
// every bound b the filter uses t' = fl(b * r - c) (one fused multiply-add) in place of the
// reference's t = fl(fl(b - o) / dd).  Writing r = (1 + d2) / dd, c = o r (1 + d3),
// t' = (b r - c)(1 + d1) with |d1|, |d2|, |d3| <= u = 2^-24:
//   |t' - t| <= 4.03 u |t| + 1.02 u |c|,
// i.e. a relative error plus an absolute one bounded per ray by A = max_k |c_k|.  With finite
// quotients (dd != 0: `finite`) the reference's swap / std::max / std::min chain reduces to
// t_in = max(-1e18, min pairs), t_out = min(1e18, max pairs); max and min move by at most the
// largest error of the elements that attain them, so |t_in' - t_in| <= 4.2 u |t_in'| + 1.2 u A,
// and likewise for t_out.  Hence, with a20 = max(SLAB_BAND_ABS A, 1e-30), SLAB_BAND_ABS = 2^-22:
//   |t_out'| > a20                                             => t_out >= 0 <=> t_out' >= 0, and
//   |t_out' - t_in'| > SLAB_BAND_REL (|t_in'| + |t_out'|) + 2 a20 => t_in <= t_out <=> t_out' > t_in'
// (SLAB_BAND_REL = 2^-21 = 8 u covers 4.2 u per side plus the rounding of the difference and of the
// band itself; 2 a20 = 8 u A covers 2 x 1.2 u A; the 1e-30 floor keeps every value the bounds rely
// on in the normal range).  Everything else takes the exact path.  A host check restates the
// filter and the reference's test and compares every proven decision on 2e7 random and
// adversarial ray / box pairs (tests/cpp/slab_filter_check.cpp: no difference at these constants
// nor at 2^-22 / 2^-24; 814 at 2^-23 / 2^-25, below the bound).  On the 720p frame a lane of the
// wave is inside a band in 6.8 % of the closest-hit node steps (14.5 % with 2^-20 / 2^-20) and in
// 45 % of the shadow ones: shadow rays aimed at a light cross its flat leaf box, where t_in' and
// t_out' come from the same axis.
// Both comparisons use a2 = 2 a20 (the first one is then stricter than it needs to be), and
// a2 = +inf when some dd component is zero, which sends every box of such a ray to the exact
// path: one per-ray threshold and no branches in the filter.  The reference's -1e18 / 1e18 start
// values are applied as med3(az, bz, -+1e18) on the z pair (one instruction each instead of a
// min / max and a clamp).  That equals max(min(az, bz), -1e18) and min(max(az, bz), 1e18) unless
// both z values lie beyond the same bound; then t_out' < -1e18 (both below) or t_in' >= t_out'
// (both above), so the filter reports a miss, as the reference does, or hands the box to the
// exact test.
constexpr float SLAB_BAND_REL = 4.76837158203125e-07f;   // 2^-21
constexpr float SLAB_BAND_ABS = 2.384185791015625e-07f;  // 2^-22
struct SlabRay {
    F3 o, d, r, c;
    float a2;  // 2 a20, or +inf when some component of dd is zero
};
__device__ __forceinline__ SlabRay slab_ray(F3 o, F3 d) {
    SlabRay s;
    s.o = o;
    s.d = d;
    const F3 dd = f3(d.x + 1e-9f, d.y + 1e-9f, d.z + 1e-9f);
    s.r = f3(1.0f / dd.x, 1.0f / dd.y, 1.0f / dd.z);
    s.c = f3(o.x * s.r.x, o.y * s.r.y, o.z * s.r.z);
    const float a20 = fmaxf(fmaxf(fmaxf(fabsf(s.c.x), fabsf(s.c.y)), fabsf(s.c.z)) * SLAB_BAND_ABS, 1e-30f);
    const bool finite = dd.x != 0.0f && dd.y != 0.0f && dd.z != 0.0f;
    s.a2 = finite ? 2.0f * a20 : __builtin_inff();
    return s;
}
// The filtered test without its exact fallback: returns the fast decision and sets `clear` when
// the decision is proven; otherwise the caller runs slab().
__device__ __forceinline__ bool slab_filter(float mnx, float mny, float mnz, float mxx, float mxy, float mxz,
                                            const SlabRay& s, bool& clear) {
    const float ax = fmaf(mnx, s.r.x, -s.c.x), bx = fmaf(mxx, s.r.x, -s.c.x);
    const float ay = fmaf(mny, s.r.y, -s.c.y), by = fmaf(mxy, s.r.y, -s.c.y);
    const float az = fmaf(mnz, s.r.z, -s.c.z), bz = fmaf(mxz, s.r.z, -s.c.z);
    const float t_in = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), __builtin_amdgcn_fmed3f(az, bz, -1e18f));
    const float t_out = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), __builtin_amdgcn_fmed3f(az, bz, 1e18f));
    const float diff = t_out - t_in;
    const float band = fmaf(SLAB_BAND_REL, fabsf(t_in) + fabsf(t_out), s.a2);
    clear = (int(fabsf(t_out) > s.a2) & int(fabsf(diff) > band)) != 0;  // both tests, no branch
    return t_out >= 0.0f && diff > 0.0f;
}
__device__ __forceinline__ bool slab_fast(float mnx, float mny, float mnz, float mxx, float mxy, float mxz,
                                          const SlabRay& s) {
    const float ax = fmaf(mnx, s.r.x, -s.c.x), bx = fmaf(mxx, s.r.x, -s.c.x);
    const float ay = fmaf(mny, s.r.y, -s.c.y), by = fmaf(mxy, s.r.y, -s.c.y);
    const float az = fmaf(mnz, s.r.z, -s.c.z), bz = fmaf(mxz, s.r.z, -s.c.z);
    const float t_in = fmaxf(fmaxf(fminf(ax, bx), fminf(ay, by)), __builtin_amdgcn_fmed3f(az, bz, -1e18f));
    const float t_out = fminf(fminf(fmaxf(ax, bx), fmaxf(ay, by)), __builtin_amdgcn_fmed3f(az, bz, 1e18f));
    const float diff = t_out - t_in;
    const float band = fmaf(SLAB_BAND_REL, fabsf(t_in) + fabsf(t_out), s.a2);
    const bool clear = (int(fabsf(t_out) > s.a2) & int(fabsf(diff) > band)) != 0;
    if (__builtin_expect(clear, 1)) return t_out >= 0.0f && diff > 0.0f;
    return slab(mnx, mny, mnz, mxx, mxy, mxz, s.o, f3(s.d.x + 1e-9f, s.d.y + 1e-9f, s.d.z + 1e-9f));
}

// Triangle: q is formed exactly as in tri_test (so |q| >= ~2.5e-15: a float plus 1e-6 in double
// is never closer to zero), rq = v_rcp_f32(q) (1 ulp), and t', b1', b2' = numerator * rq lie within
// ~2.5 ulp of the exact quotients.  A triangle is rejected without division when any of
//   t' <= -1e-30, b1' <= -1e-30, b2' <= -1e-30  (sign of a quotient that cannot underflow),
//   t' >= best * (1 + 2^-18) + 1e-30,  b1' + b2' >= 1 + 2^-18
// holds -- each implies the exact acceptance test fails.  Everything else is decided on the
// exact quotients.  Returns true iff tri_valid(t, b1, b2) and t <= best, with exact t, b1, b2.
__device__ __forceinline__ bool tri_fast(const TriGeom& g, F3 o, F3 d, float best, float& t, float& b1, float& b2) {
    const F3 p0 = f3(g.g0.x, g.g0.y, g.g0.z);
    const F3 e1 = f3(g.g0.w, g.g1.x, g.g1.y);
    const F3 e2 = f3(g.g1.z, g.g1.w, g.g2.x);
    const F3 s = sub(o, p0), s1 = cross(d, e2), s2 = cross(s, e1);
    const float tn = dot(s2, e2), b1n = dot(s1, s), b2n = dot(s2, d);
    const float q = float(double(dot(s1, e1)) + 1e-6);
    const float rq = __builtin_amdgcn_rcpf(q);
    const float tq = tn * rq, b1q = b1n * rq, b2q = b2n * rq;
    if (tq <= -1e-30f || b1q <= -1e-30f || b2q <= -1e-30f) return false;
    if (tq >= best * 1.000003814697265625f + 1e-30f || b1q + b2q >= 1.000003814697265625f) return false;
    t = tn / q;
    b1 = b1n / q;
    b2 = b2n / q;
    return tri_valid(t, b1, b2) && t <= best;
}

// ------------------------------------------------------------------ Philox4x32-10
// Salmon et al., SC'11.  Counter = (pixel, sample, c2, c3), key = (seed, 0x6E755261).
__device__ __forceinline__ uint4 philox(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
        uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
        c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}
__device__ __forceinline__ float u01(uint32_t v) { return float(v >> 8) * (1.0f / 16777216.0f); }
__device__ __forceinline__ float4 rng4(uint32_t seed, uint32_t pixel, uint32_t sample, uint32_t c2, uint32_t c3) {
    uint4 r = philox(make_uint4(pixel, sample, c2, c3), seed, 0x6E755261u);
    return make_float4(u01(r.x), u01(r.y), u01(r.z), u01(r.w));
}

}  // namespace nr
