// BVH traversal for gfx950.
//
// Semantics (bit-exact with the reference traversal, src/hierarchy/bvh.cpp:99-137):
//   * a node's children are visited iff the reference slab test (bvhnode.cpp:35-48) accepts
//     their boxes; there is no distance culling, exactly like BVH::_traversal, so the set of
//     candidate triangles is the reference's;
//   * the winner is the smallest t; among equal t the later leaf in DFS (left-first) order wins
//     (bvh.cpp:119-126 returns the right result unless left.t < right.t); inside one leaf the
//     first triangle wins (strict '<', bvhnode.cpp:57).  Leaves are stored in DFS order, so the
//     "later leaf" is the one with the larger first-primitive index, which makes the winner
//     independent of the order this kernel visits nodes in.
//
// Execution: persistent waves; lane 0 of each wave claims 64 rays with one atomic, every lane
// traverses one ray with a private stack kept in LDS at stack[slot * blockDim + tid] (one
// 4-byte bank per lane, conflict-free).  The stack depth is the BVH depth + 2, checked on
// every push; an overflow is reported, never written past.
#include "trace.h"
#include "scene.h"

namespace nr {

namespace {

constexpr int BLOCK = 256;

template <bool ANY, bool STATS>
__global__ __launch_bounds__(BLOCK) void k_trace(SceneView sv, const Ray* __restrict__ rays,
                                                  const int* __restrict__ n_ptr, HitRec* __restrict__ hits,
                                                  uint8_t* __restrict__ occ, const float4* __restrict__ contrib,
                                                  PathState* __restrict__ state, int* __restrict__ work,
                                                  int* __restrict__ err, TraceStats* __restrict__ stats) {
    extern __shared__ int stk[];
    const int tid = threadIdx.x;
    const int bdim = blockDim.x;
    const int lane = tid & 63;
    const int cap = sv.stack_cap;
    const int n = *n_ptr;
    for (;;) {
        int base = 0;
        if (lane == 0) base = atomicAdd(work, 64);
        base = __shfl(base, 0, 64);
        if (base >= n) break;
        const int i = base + lane;
        if (i >= n) continue;

        const Ray r = rays[i];
        const F3 o = f3(r.o.x, r.o.y, r.o.z);
        const F3 d = f3(r.d.x, r.d.y, r.d.z);
        const SlabRay sr = slab_ray(o, d);
        const double len = double(r.o.w);
        // closest hit: running winner; any hit: `best` caps the triangle filter at the segment
        float bt = ANY ? fminf(r.o.w, 1e18f) : 1e18f, bb1 = 0.0f, bb2 = 0.0f;
        int bprim = -1, blf = -1;
        bool done = false, overflow = false;
        uint32_t n_pf = 0, n_tt = 0;

        // The stack holds internal-node and leaf references.  Internal nodes are expanded
        // until this lane holds a leaf (or runs dry); the wave then tests its leaves together.
        int sp = 0;
        int32_t leaf = 0;
        bool has_leaf = false;
        if (slab_fast(sv.root_min[0], sv.root_min[1], sv.root_min[2], sv.root_max[0], sv.root_max[1],
                      sv.root_max[2], sr)) {
            if (ref_is_leaf(sv.root_ref)) leaf = sv.root_ref, has_leaf = true;
            else stk[tid] = sv.root_ref, sp = 1;
        }
        for (;;) {
            while (sp > 0 && !has_leaf) {
                --sp;
                const int32_t e = stk[sp * bdim + tid];
                if (ref_is_leaf(e)) {
                    leaf = e;
                    has_leaf = true;
                    break;
                }
                const NodePair* np = sv.nodes + e;
                const float4 a = np->a, b = np->b, c = np->c;
                const int4 ch = np->d;
                if (STATS) ++n_pf;
                const bool hl = slab_fast(a.x, a.y, a.z, a.w, b.x, b.y, sr);
                const bool hr = slab_fast(b.z, b.w, c.x, c.y, c.z, c.w, sr);
                // a hit leaf child is taken at once; everything else is pushed
                int32_t push0 = 0, push1 = 0;
                int np_ = 0;
                if (hl) {
                    if (ref_is_leaf(ch.x)) leaf = ch.x, has_leaf = true;
                    else push0 = ch.x, np_ = 1;
                }
                if (hr) {
                    if (ref_is_leaf(ch.y) && !has_leaf) leaf = ch.y, has_leaf = true;
                    else if (np_ == 0) push0 = ch.y, np_ = 1;
                    else push1 = ch.y, np_ = 2;
                }
                if (sp + np_ > cap) {
                    overflow = true;
                    sp = 0;
                    has_leaf = false;
                    break;
                }
                if (np_ >= 1) stk[sp * bdim + tid] = push0;
                if (np_ == 2) stk[(sp + 1) * bdim + tid] = push1;
                sp += np_;
            }
            if (!has_leaf) break;
            has_leaf = false;
            // Leaf loop, bvhnode.cpp:50-66, merged into the running winner with the DFS tie rule.
            const int first = ref_first(leaf), cnt = ref_count(leaf);
            if (STATS) n_tt += cnt;
            for (int k = 0; k < cnt; ++k) {
                const TriGeom g = sv.tris[first + k];
                float t, b1, b2;
                if (!tri_fast(g, o, d, bt, t, b1, b2)) continue;
                if (ANY) {
                    if (double(t) + 2e-3 <= len) {
                        bt = t; bb1 = b1; bb2 = b2; bprim = first + k;
                        done = true;
                        break;
                    }
                } else if (t < bt || first > blf) {  // t <= bt here: equal t -> later leaf wins
                    bt = t; bb1 = b1; bb2 = b2; bprim = first + k; blf = first;
                }
            }
            if (ANY && done) break;
        }
        if (overflow) atomicOr(err, 1);
        if (STATS) {
            atomicAdd(&stats->rays, 1ull);
            atomicAdd(&stats->pair_fetch, (unsigned long long)n_pf);
            atomicAdd(&stats->tri_tests, (unsigned long long)n_tt);
        }
        if (!ANY) {
            hits[i] = HitRec{bt, bb1, bb2, bprim};
        } else {
            if (occ) occ[i] = done ? 1 : 0;
            if (hits) hits[i] = HitRec{done ? bt : 1e18f, bb1, bb2, done ? bprim : -1};
            if (contrib && !done) {
                const float4 cv = contrib[i];
                const int slot = __float_as_int(r.d.w);
                float4 L = state[slot].L_depth;
                L.x += cv.x;
                L.y += cv.y;
                L.z += cv.z;
                state[slot].L_depth = L;
            }
        }
    }
}

template <bool ANY, bool STATS>
int grid_for(size_t shmem) {
    static int cached_shmem = -1, cached_blocks = 0;
    if (cached_shmem != int(shmem)) {
        int per_cu = 0;
        check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<ANY, STATS>, BLOCK, shmem),
                  "occupancy");
        if (per_cu < 1) per_cu = 1;
        cached_blocks = per_cu * device_cu_count();
        cached_shmem = int(shmem);
    }
    return cached_blocks;
}

template <bool ANY>
void launch(const SceneView& sv, const Ray* rays, const int* n, HitRec* hits, uint8_t* occ, const float4* contrib,
            PathState* state, int* work, int* err, TraceStats* stats, hipStream_t s) {
    const size_t shmem = size_t(sv.stack_cap) * BLOCK * sizeof(int);
    if (shmem > 160 * 1024) throw std::runtime_error("BVH too deep for the LDS traversal stack");
    if (stats) {
        hipLaunchKernelGGL((k_trace<ANY, true>), dim3(grid_for<ANY, true>(shmem)), dim3(BLOCK), shmem, s, sv, rays, n,
                           hits, occ, contrib, state, work, err, stats);
    } else {
        hipLaunchKernelGGL((k_trace<ANY, false>), dim3(grid_for<ANY, false>(shmem)), dim3(BLOCK), shmem, s, sv, rays,
                           n, hits, occ, contrib, state, work, err, stats);
    }
    check_hip(hipGetLastError(), "k_trace launch");
}

}  // namespace

int device_cu_count() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        check_hip(hipGetDevice(&dev), "hipGetDevice");
        check_hip(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "CU count");
    }
    return cus;
}

void launch_closest(const SceneView& sv, const Ray* rays, const int* n, HitRec* hits, int* work, int* err,
                    TraceStats* stats, hipStream_t s) {
    launch<false>(sv, rays, n, hits, nullptr, nullptr, nullptr, work, err, stats, s);
}

void launch_shadow(const SceneView& sv, const Ray* rays, const int* n, uint8_t* occ, HitRec* hits,
                   const float4* contrib, PathState* state, int* work, int* err, TraceStats* stats, hipStream_t s) {
    launch<true>(sv, rays, n, hits, occ, contrib, state, work, err, stats, s);
}

}  // namespace nr
