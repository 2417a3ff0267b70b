// Device-side BVH traversal building blocks (shared by the persistent trace kernels and by the
// per-path replay kernels).  Semantics: see trace.hip -- no distance culling, the reference's
// exact slab and triangle tests (filtered), DFS-order tie rule.
#pragma once
#include "nr_device.h"

namespace nr {

struct TravState {
    int sp = 0;
    int32_t leaf = 0;
    bool has_leaf = false;
    bool overflow = false;
};

// The 4-wide walk (Node4, see nr_device.h): the same leaves, and the same hits, as a DFS of the
// binary SAH tree, with half the node steps.  Pushes are branch-free: the four
// child references are written at their compacted positions in slot order (a missed child's entry is
// overwritten by the next hit one or left dead above the new top), so the stack needs three spare
// slots beyond `cap` live entries.
__device__ __forceinline__ void trav_begin_wide(const SceneView& sv, const SlabRay& sr, int* stk, TravState& s) {
    s = TravState{};
    if (slab_fast(sv.root_min[0], sv.root_min[1], sv.root_min[2], sv.root_max[0], sv.root_max[1], sv.root_max[2],
                  sr)) {
        stk[0] = sv.wide_root;
        s.sp = 1;
    }
}

// Wave-cooperative expansion of the 4-wide tree for the persistent kernels: every lane keeps expanding
// internal nodes while other lanes of the wave are still looking for a leaf.  A lane that already
// holds a leaf goes on expanding internal nodes (its leaf is postponed); when it pops a second leaf
// it puts that back and parks until the wave's leaf round.  The loop ends when no lane is still
// looking, or when at most LEAF_ROUND_SEARCHERS lanes are and some lane holds a leaf (waiting for
// the last few searchers would leave the parked lanes idle; they search on after the leaf round).
// Every call makes progress: while no lane holds a leaf the loop pops an entry per searching lane.
// The set of nodes and leaves visited is the same as in the sequential DFS (only their interleaving
// changes), and the result does not depend on the order.
// The top entry lives in a register (`top`): a pop takes it from there and reads the entry below
// it from LDS at once, so that read is in flight together with the popped node's loads; after the
// push the new top is the last child pushed, or that entry when no child was.
template <bool STATS, int STRIDE, int LEAF_ROUND_SEARCHERS = 16>
__device__ __forceinline__ void trav_expand_wave4(const SceneView& sv, const SlabRay& sr, int* stk, int cap,
                                                  TravState& s, uint32_t& n_nf) {
    bool parked = false;
    int32_t top = stk[max(s.sp - 1, 0) * STRIDE];
    for (;;) {
        const int searching = __popcll(__ballot(!s.has_leaf && s.sp > 0));
        if (searching == 0 || (searching <= LEAF_ROUND_SEARCHERS && __any(s.has_leaf))) return;
        if (s.sp > 0 && !parked) {
            const int32_t e = top;
            const int32_t below = stk[max(s.sp - 2, 0) * STRIDE];  // the top after this pop
            if (ref_is_leaf(e)) {
                parked = s.has_leaf;
                s.leaf = s.has_leaf ? s.leaf : e;
                top = s.has_leaf ? top : below;
                s.sp -= s.has_leaf ? 0 : 1;
                s.has_leaf = true;
            } else {
                --s.sp;
                const Node4* nd = sv.wide + e;
                const float4 lx = nd->lox, ly = nd->loy, lz = nd->loz;
                const float4 hx = nd->hix, hy = nd->hiy, hz = nd->hiz;
                const int4 ch = nd->ref;
                if (STATS) ++n_nf;
                bool c0, c1, c2, c3;
                bool h0 = slab_filter(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, sr, c0);
                bool h1 = slab_filter(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, sr, c1);
                bool h2 = slab_filter(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, sr, c2);
                bool h3 = slab_filter(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, sr, c3);
                if (__builtin_expect(!(c0 & c1 & c2 & c3), 0)) {  // inside an error band: the exact test
                    const F3 dd = f3(sr.d.x + 1e-9f, sr.d.y + 1e-9f, sr.d.z + 1e-9f);
                    if (!c0) h0 = slab(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, sr.o, dd);
                    if (!c1) h1 = slab(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, sr.o, dd);
                    if (!c2) h2 = slab(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, sr.o, dd);
                    if (!c3) h3 = slab(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, sr.o, dd);
                }
                // sp <= cap - 1 here, so the writes (at most sp + 3) stay inside cap + 3 slots even
                // when the push overflows (which then ends the ray with the overflow flag)
                int p = s.sp;
                stk[p * STRIDE] = ch.x;
                p += int(h0);
                stk[p * STRIDE] = ch.y;
                p += int(h1);
                stk[p * STRIDE] = ch.z;
                p += int(h2);
                stk[p * STRIDE] = ch.w;
                p += int(h3);
                top = h3 ? ch.w : h2 ? ch.z : h1 ? ch.y : h0 ? ch.x : below;
                const bool ovf = p > cap;
                s.overflow = s.overflow || ovf;
                s.sp = ovf ? 0 : p;
            }
        }
    }
}

// Per-lane form of the 4-wide walk (trace_ray): pops entries, expanding Node4s, until the lane
// holds a leaf or the stack is empty.  Same compacted branch-free push as trav_expand_wave4 (three
// spare slots beyond `cap`); an overflowing push ends the walk with the overflow flag.
template <bool STATS, int STRIDE>
__device__ __forceinline__ void trav_expand4(const SceneView& sv, const SlabRay& sr, int* stk, int cap, TravState& s,
                                             uint32_t& n_nf) {
    while (s.sp > 0) {
        --s.sp;
        const int32_t e = stk[s.sp * STRIDE];
        if (ref_is_leaf(e)) {
            s.leaf = e;
            s.has_leaf = true;
            return;
        }
        const Node4* nd = sv.wide + e;
        const float4 lx = nd->lox, ly = nd->loy, lz = nd->loz;
        const float4 hx = nd->hix, hy = nd->hiy, hz = nd->hiz;
        const int4 ch = nd->ref;
        if (STATS) ++n_nf;
        const bool h0 = slab_fast(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, sr);
        const bool h1 = slab_fast(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, sr);
        const bool h2 = slab_fast(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, sr);
        const bool h3 = slab_fast(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, sr);
        int p = s.sp;  // <= cap - 1, so the writes (at most p + 3) stay inside cap + 3 slots
        stk[p * STRIDE] = ch.x;
        p += int(h0);
        stk[p * STRIDE] = ch.y;
        p += int(h1);
        stk[p * STRIDE] = ch.z;
        p += int(h2);
        stk[p * STRIDE] = ch.w;
        p += int(h3);
        if (p > cap) {
            s.overflow = true;
            s.sp = 0;
            return;
        }
        s.sp = p;
    }
}

// Leaf loop, bvhnode.cpp:50-66, merged into the running winner with the DFS tie rule (closest
// hit), or the shadow rule (double)t + 2e-3 <= len (any hit: returns true when blocked).
template <bool ANY, bool STATS>
__device__ __forceinline__ bool trav_leaf(const SceneView& sv, F3 o, F3 d, double len, TravState& s, float& bt,
                                          float& bb1, float& bb2, int& bprim, int& blf, uint32_t& n_tt) {
    s.has_leaf = false;
    const int first = ref_first(s.leaf), cnt = ref_count(s.leaf);
    if (STATS) n_tt += cnt;
    if (cnt == 0) return false;
    // the next triangle's loads are in flight while the current one is tested
    TriGeom nxt = sv.tris[first];
    for (int k = 0; k < cnt; ++k) {
        const TriGeom g = nxt;
        if (k + 1 < cnt) nxt = sv.tris[first + k + 1];
        float t, b1, b2;
        if (!tri_fast(g, o, d, bt, t, b1, b2)) continue;
        if (ANY) {
            if (double(t) + 2e-3 <= len) {
                bt = t; bb1 = b1; bb2 = b2; bprim = first + k;
                return true;
            }
        } else if (t < bt || first > blf) {  // t <= bt here: equal t -> later leaf wins
            bt = t; bb1 = b1; bb2 = b2; bprim = first + k; blf = first;
        }
    }
    return false;
}

// Wave-cooperative leaf round of the persistent kernels (trav_leaf's role there).  The triangles of
// every lane's pending leaf are laid out one after another in lane order and tested 64 at a time,
// one per lane, so lanes without a leaf of their own test triangles of the others' leaves (the
// per-lane loop runs as many iterations as the wave's longest leaf with a third of the lanes busy).
// Closest hit: each owner takes its leaf's smallest t, the first triangle among equal t (the loop's
// strict '<'), and merges it into its running winner with the DFS tie rule.  Any hit: the owner's
// first triangle with (double)t + 2e-3 <= len blocks the ray (returns true).  Bit-identical to
// trav_leaf: tri_fast accepts a triangle against a winner w iff the triangle's own t is <= w (the
// rcp filter's margin never rejects such a t), so testing every triangle against the winner held
// at the start of the round and keeping the leaf's minimum is the loop's outcome; the any-hit
// winner never changes before the ray is blocked.
// All 64 lanes call it (wave-uniform control flow); `mark` is 64 ints of LDS owned by the wave.
template <bool ANY, bool STATS>
__device__ __forceinline__ bool trav_leaf_wave(const SceneView& sv, F3 o, F3 d, float len, bool own, TravState& s,
                                               float& bt, float& bb1, float& bb2, int& bprim, int& blf,
                                               uint32_t& n_tt, int* mark, int lane) {
    if (__ballot(own) == 0ull) return false;
    const int cnt = own ? ref_count(s.leaf) : 0;
    const int first = own ? ref_first(s.leaf) : 0;
    if (own) s.has_leaf = false;
    if (STATS) n_tt += cnt;
    // exclusive prefix of the 4-bit counts over the lanes, and their total (wave-uniform)
    const unsigned long long below = (1ull << lane) - 1ull;
    int excl = 0, total = 0;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const unsigned long long m = __ballot((cnt >> b) & 1);
        excl += __popcll(m & below) << b;
        total += __popcll(m) << b;
    }
    float lt = __builtin_inff(), lb1 = 0.0f, lb2 = 0.0f;  // the owner's leaf winner so far
    int lk = -1;
    for (int base = 0; base < total; base += 64) {
        // owner of position base + lane: the last owner whose run starts at or before it (marked in
        // LDS at the run's start), else the owner whose run began in an earlier pass
        const int st = excl - base;
        __hip_atomic_store(&mark[lane], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        if (cnt > 0 && st >= 0 && st < 64) __hip_atomic_store(&mark[st], lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        const int v = __hip_atomic_load(&mark[lane], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        const unsigned long long upto = __ballot(v >= 0) & (~0ull >> (63 - lane));
        const unsigned long long carried = __ballot(cnt > 0 && st < 0);
        const int vs = __shfl(v, upto ? 63 - __builtin_clzll(upto) : 0);
        const int ow = upto ? vs : (carried ? 63 - __builtin_clzll(carried) : 0);
        // the owner's ray, winner and triangle offset; one triangle per position
        const int p = base + lane;
        const F3 po = f3(__shfl(o.x, ow), __shfl(o.y, ow), __shfl(o.z, ow));
        const F3 pd = f3(__shfl(d.x, ow), __shfl(d.y, ow), __shfl(d.z, ow));
        const float pbt = __shfl(bt, ow);
        const float plen = ANY ? __shfl(len, ow) : 0.0f;  // (shuffles only here, with every lane active)
        const int tri = __shfl(first - excl, ow) + p;
        float t = __builtin_inff(), b1 = 0.0f, b2 = 0.0f;
        bool blocked = false;
        if (p < total) {
            float tt, u1, u2;
            if (tri_fast(sv.tris[tri], po, pd, pbt, tt, u1, u2)) {
                t = tt, b1 = u1, b2 = u2;
                if (ANY) blocked = double(tt) + 2e-3 <= double(plen);
            }
        }
        // the owner's positions in this pass: q0 .. q0 + nq - 1, in triangle order
        const int q0 = max(st, 0), nq = max(min(st + cnt, 64) - q0, 0);
        if (ANY) {
            const unsigned long long seg = nq == 0 ? 0ull : nq == 64 ? ~0ull : ((1ull << nq) - 1ull) << q0;
            const unsigned long long hm = __ballot(blocked) & seg;
            const int hq = hm ? __builtin_ctzll(hm) : 0;
            const float ht = __shfl(t, hq), h1 = __shfl(b1, hq), h2 = __shfl(b2, hq);
            if (hm && lk < 0) lt = ht, lb1 = h1, lb2 = h2, lk = base + hq - excl;
        } else {
            float pt = __builtin_inff();
            int pq = q0;
            for (int j = 0; __ballot(j < nq) != 0ull; ++j) {
                const float tj = __shfl(t, q0 + j);
                if (j < nq && tj < pt) pt = tj, pq = q0 + j;
            }
            const float w1 = __shfl(b1, pq), w2 = __shfl(b2, pq);
            if (pt < lt) lt = pt, lb1 = w1, lb2 = w2, lk = base + pq - excl;  // earlier passes hold lower k
        }
    }
    if (ANY) {
        if (lk < 0) return false;
        bt = lt; bb1 = lb1; bb2 = lb2; bprim = first + lk;
        return true;
    }
    if (lk >= 0 && (lt < bt || first > blf)) {  // lt <= bt: equal t -> later leaf wins
        bt = lt; bb1 = lb1; bb2 = lb2; bprim = first + lk; blf = first;
    }
    return false;
}

// Small scenes (SceneView::n_flat > 0): every lane of a wave tests the same leaf boxes in the same
// order -- a wave-uniform loop whose box loads have one address per wave -- and records the leaves
// its ray passes in a 64-bit mask; then each lane tests the triangles of its own leaves only.  The
// tree walk instead leaves most of a wave idle on scenes this small (the lanes' walks differ in
// length).  Exact: the tree walk reaches a leaf iff the leaf's own box passes the slab test (every
// ancestor box contains it), and the winner does not depend on the order in which leaves are tested.
template <bool ANY>
__device__ __forceinline__ HitRec trace_flat(const SceneView& sv, F3 o, F3 d, float len) {
    const SlabRay sr = slab_ray(o, d);
    uint64_t pass = 0;
    for (int k = 0; k < sv.n_flat; ++k) {
        const float4 lo = sv.flat[2 * k], hi = sv.flat[2 * k + 1];
        if (slab_fast(lo.x, lo.y, lo.z, hi.x, hi.y, hi.z, sr)) pass |= 1ull << k;
    }
    TravState s;
    float bt = ANY ? fminf(len, 1e18f) : 1e18f, bb1 = 0.0f, bb2 = 0.0f;
    int bprim = -1, blf = -1;
    uint32_t ntt = 0;
    bool done = false;
    while (pass != 0 && !done) {
        const int k = __builtin_ctzll(pass);
        pass &= pass - 1;
        s.leaf = __float_as_int(sv.flat[2 * k].w);
        done = trav_leaf<ANY, false>(sv, o, d, double(len), s, bt, bb1, bb2, bprim, blf, ntt);
    }
    if (ANY) return HitRec{done ? bt : 1e18f, bb1, bb2, done ? bprim : -1};
    return HitRec{bt, bb1, bb2, bprim};
}

// Whole traversal of one ray by one lane (no refill) over the 4-wide tree: closest hit, or shadow
// test with segment length len (returns hit.prim >= 0 iff blocked).  overflow is set if the stack
// ran out.  The stack holds SceneView::lane_slots slots STRIDE ints apart.
template <bool ANY, int STRIDE>
__device__ __forceinline__ HitRec trace_ray(const SceneView& sv, F3 o, F3 d, float len, int* stk, bool& overflow) {
    if (sv.n_flat > 0) return trace_flat<ANY>(sv, o, d, len);
    const SlabRay sr = slab_ray(o, d);
    TravState s;
    trav_begin_wide(sv, sr, stk, s);
    float bt = ANY ? fminf(len, 1e18f) : 1e18f, bb1 = 0.0f, bb2 = 0.0f;
    int bprim = -1, blf = -1;
    uint32_t npf = 0, ntt = 0;
    bool done = false;
    for (;;) {
        trav_expand4<false, STRIDE>(sv, sr, stk, sv.wide_stack_cap, s, npf);
        if (!s.has_leaf) break;
        if (trav_leaf<ANY, false>(sv, o, d, double(len), s, bt, bb1, bb2, bprim, blf, ntt)) {
            done = true;
            break;
        }
    }
    overflow = overflow || s.overflow;
    if (ANY) return HitRec{done ? bt : 1e18f, bb1, bb2, done ? bprim : -1};
    return HitRec{bt, bb1, bb2, bprim};
}

}  // namespace nr
