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
// Execution: persistent waves over the 4-wide form of the tree (Node4: the binary tree's own
// leaves and boxes, half the node steps).  Every lane traverses one ray with a private stack kept
// in LDS at stack[slot * BLOCK + tid] (one 4-byte bank per lane, conflict-free).  Internal nodes
// are expanded wave-cooperatively (trav_expand_wave4) until (nearly) every lane holds a leaf, then
// the wave tests the held leaves' triangles together, one triangle per lane, idle lanes included
// (trav_leaf_wave); a lane whose ray is done goes idle, and once refill_min lanes are
// idle one atomic claims new rays for all of them (ballot + prefix count), so long rays do not hold
// a whole wave hostage.  The first launch gives each lane SceneView::short_stack live entries (plus
// three spare slots for the branch-free push).  The rare ray
// that needs more is abandoned, listed, and traced again from its start by a second launch whose
// stack holds 3 entries per Node4 level + 2 (checked on every push; an overflow there is reported,
// never written past).
#include "trace.h"
#include "nr_traverse.h"
#include "scene.h"

namespace nr {

namespace {

constexpr int BLOCK = 256;

// Lanes whose ray has finished are refilled when at least this many are idle.  A wave claims
// rays from the global counter up to CLAIM at a time and hands them out from that private range;
// for small batches the claim shrinks to the batch's share per wave, so that every wave starts
// rays at once instead of a few waves working through long private ranges (the launch then
// takes about one ray's latency, not several).  Large launches wait for more idle lanes (fewer,
// fuller refills): on the 720p x 64 spp terrain frame (camera rays queued pixel by pixel) 40 idle
// lanes measured best for both launches (36 and 44: -0.5 %, 24 and 48: -1.5 to -3 %), while the
// 1080p x 1 spp frames, whose launches hold at most 2 M rays, are 1.5 % faster with 24.
constexpr int REFILL_MIN = 24, REFILL_MIN_LARGE = 40, REFILL_LARGE_RAYS = 1 << 22;
// Leaf-round thresholds of the wave-cooperative expansion.  With the lane-parallel leaf round a
// round costs about ceil(triangles / 64) triangle tests, so it pays to start one earlier than with
// the per-lane loop (which favoured 24 / 8) for shadow rays: on the 720p terrain frame 12-20 are
// within 0.5 % of each other.  With camera rays queued pixel by pixel and 40-lane refills, 24
// closest-hit searchers measured 0.4 % ahead of 32 and 1.3 % ahead of 40.
constexpr int CLOSEST_SEARCHERS = 24;
constexpr int SHADOW_SEARCHERS = 16;
constexpr int CLAIM = 128;
// Live stack entries of the first launch default to SceneView::short_stack = 16:
// (16 + 3) x 256 x 4 B = 19 KB of LDS per block.  The kernel is compiled for 6 waves per SIMD
// (6 blocks per CU, 114 KB of LDS): the 4-wide node step needs 76-79 VGPRs, and capping it at 64 for
// 8 waves spills registers to scratch memory, which costs more than the extra waves bring.

// rays are rays[idx[i]] for i < *n_ptr when idx is given (the retry launch, RETRY = true: a
// separate kernel symbol so that profiles keep the two launches apart), else rays[i].
template <bool ANY, bool STATS, bool RETRY>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(6))) void k_trace(SceneView sv, const Ray* __restrict__ rays,
                                                  const int* __restrict__ n_ptr, const int* __restrict__ idx,
                                                  HitRec* __restrict__ hits, uint8_t* __restrict__ occ,
                                                  const float4* __restrict__ contrib, float4* __restrict__ radiance,
                                                  int* __restrict__ work, int* __restrict__ err,
                                                  int* __restrict__ retry, int* __restrict__ n_retry, int cap,
                                                  TraceStats* __restrict__ stats, HitRec* __restrict__ vrec) {
    extern __shared__ int stk_all[];
    __shared__ int leaf_mark[BLOCK];  // the leaf round's run starts, 64 per wave
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    int* const stk = stk_all + tid;
    const unsigned long long below = (1ull << lane) - 1ull;
    const int n = *n_ptr;
    if (n <= 0) return;  // nothing queued (the retry launch, most of the time): no claims at all
    const int n_waves = int(gridDim.x) * (BLOCK / 64);
    const int claim = max(1, min(CLAIM, (n + n_waves - 1) / n_waves));
    const int refill_min = n >= REFILL_LARGE_RAYS ? REFILL_MIN_LARGE : REFILL_MIN;

    // per-lane ray state; i < 0 means the lane is idle
    int i = -1;
    F3 o = f3(0, 0, 0), d = f3(0, 0, 0);
    SlabRay sr;
    TravState ts;
    float wr = 0.0f;  // ray payload words: segment length / vertex depth, path slot
    float pw = 0.0f;
    float bt = 1e18f, bb1 = 0.0f, bb2 = 0.0f;
    int bprim = -1, blf = -1;
    bool done = false;
    uint32_t n_pf = 0, n_tt = 0;
    bool drained = false;  // wave-uniform: the global counter is exhausted
    int rb = 0, re = 0;    // wave-uniform: claimed, not yet assigned ray indices [rb, re)

    for (;;) {
        // ---- refill idle lanes from the wave's claimed range; claim more when it is empty
        const unsigned long long idle_mask = __ballot(i < 0);
        const int n_idle = __popcll(idle_mask);
        if ((n_idle >= refill_min || n_idle == 64) && !(drained && rb == re)) {
            if (rb == re) {
                int base = 0;
                if (lane == 0) base = atomicAdd(work, claim);
                base = __shfl(base, 0, 64);
                if (base >= n) drained = true;
                else rb = base, re = min(base + claim, n);
            }
            const int take = min(n_idle, re - rb);
            if (i < 0) {
                const int rank = __popcll(idle_mask & below);
                if (rank < take) {
                    i = rb + rank;
                    const Ray r = rays[idx ? idx[i] : i];
                    o = f3(r.o.x, r.o.y, r.o.z);
                    d = f3(r.d.x, r.d.y, r.d.z);
                    sr = slab_ray(o, d);
                    wr = r.o.w;
                    pw = r.d.w;
                    bt = ANY ? fminf(wr, 1e18f) : 1e18f;
                    bb1 = bb2 = 0.0f;
                    bprim = blf = -1;
                    done = false;
                    n_pf = n_tt = 0;
                    trav_begin_wide(sv, sr, stk, ts);
                }
            }
            rb += take;
        }
        if (__ballot(i >= 0) == 0ull) {
            if (drained && rb == re) break;
            continue;
        }
        if (i >= 0)
            trav_expand_wave4<STATS, BLOCK, ANY ? SHADOW_SEARCHERS : CLOSEST_SEARCHERS>(sv, sr, stk, cap, ts, n_pf);
        // the leaf round: every lane of the wave takes part, idle ones included
        if (trav_leaf_wave<ANY, STATS>(sv, o, d, wr, i >= 0 && ts.has_leaf, ts, bt, bb1, bb2, bprim, blf, n_tt,
                                       leaf_mark + (tid & ~63), lane))
            done = true;
        if (i >= 0) {
            // ---- finished: write the result, go idle
            if ((ts.sp == 0 && !ts.has_leaf) || done || ts.overflow) {
                if (ts.overflow && retry) {  // traced again from the start with the full stack
                    retry[atomicAdd(n_retry, 1)] = idx ? idx[i] : i;
                    if (STATS) atomicAdd(&stats->retried, 1ull);
                    i = -1;
                    continue;
                }
                if (ts.overflow) atomicOr(err, 1);
                if (idx) i = idx[i];
                if (STATS) {
                    atomicAdd(&stats->rays, 1ull);
                    atomicAdd(&stats->node_fetch, (unsigned long long)n_pf);
                    atomicAdd(&stats->tri_tests, (unsigned long long)n_tt);
                }
                if (!ANY) {
                    hits[i] = HitRec{bt, bb1, bb2, bprim};
                } else {
                    if (occ) occ[i] = done ? 1 : 0;
                    if (hits) hits[i] = HitRec{done ? bt : 1e18f, bb1, bb2, done ? bprim : -1};
                    if (contrib && !done) {
                        const float4 cv = contrib[i];
                        const int slot = __float_as_int(pw);
                        float4 L = radiance[slot];
                        L.x += cv.x;
                        L.y += cv.y;
                        L.z += cv.z;
                        radiance[slot] = L;
                        // the vertex record of the ray's path vertex (index in w, -1: none)
                        const int ri = __float_as_int(cv.w);
                        if (vrec && ri >= 0) vrec[ri].prim |= VREC_VISIBLE;
                    }
                }
                i = -1;
            }
        }
    }
}

constexpr int SORT_BLOCK = 256, SORT_BINS = 64;

template <bool OCT>
__device__ __forceinline__ int dir_bin(float4 d) {
    if (OCT) return (d.x < 0.0f ? 4 : 0) | (d.y < 0.0f ? 2 : 0) | (d.z < 0.0f ? 1 : 0);
    auto q = [](float x) { return min(3, max(0, int((x + 1.0f) * 2.0f))); };
    return q(d.x) * 16 + q(d.y) * 4 + q(d.z);
}

// One chunk of CHUNK queue positions per block: bin counts in LDS, an exclusive scan, then
// every position is written to its bin's next free place (the order inside a bin is whatever the
// LDS atomics give: the traversal's results do not depend on the order rays are traced in).
template <int CHUNK, bool OCT>
__global__ __launch_bounds__(SORT_BLOCK) void k_sort_dir(const Ray* __restrict__ rays, const int* __restrict__ n_ptr,
                                                        int* __restrict__ order) {
    __shared__ int bins[SORT_BINS];
    const int n = *n_ptr;
    const int c0 = blockIdx.x * CHUNK;
    if (c0 >= n) return;
    const int m = min(CHUNK, n - c0);
    if (threadIdx.x < SORT_BINS) bins[threadIdx.x] = 0;
    __syncthreads();
    constexpr int PER = CHUNK / SORT_BLOCK;
    int key[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int j = int(threadIdx.x) + k * SORT_BLOCK;
        key[k] = j < m ? dir_bin<OCT>(rays[c0 + j].d) : -1;
        if (key[k] >= 0) atomicAdd(&bins[key[k]], 1);
    }
    __syncthreads();
    if (threadIdx.x < 64) {  // exclusive scan of the 64 bin counts by wave 0
        const int v = bins[threadIdx.x];
        int incl = v;
        for (int off = 1; off < 64; off <<= 1) {
            const int u = __shfl_up(incl, off, 64);
            if (int(threadIdx.x) >= off) incl += u;
        }
        bins[threadIdx.x] = incl - v;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k)
        if (key[k] >= 0) order[c0 + atomicAdd(&bins[key[k]], 1)] = c0 + int(threadIdx.x) + k * SORT_BLOCK;
}

template <bool ANY, bool STATS>
int grid_for(size_t shmem) {
    static int cached_shmem = -1, cached_blocks = 0;
    if (cached_shmem != int(shmem)) {
        int per_cu = 0;
        check_hip(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_trace<ANY, STATS, false>, BLOCK, shmem),
                  "occupancy");
        if (per_cu < 1) per_cu = 1;
        cached_blocks = per_cu * device_cu_count();
        cached_shmem = int(shmem);
    }
    return cached_blocks;
}

template <bool ANY>
void launch(const SceneView& sv, const Ray* rays, const int* n, HitRec* hits, uint8_t* occ, const float4* contrib,
            float4* radiance, int* work, int* err, TraceStats* stats, hipStream_t s, int* retry, HitRec* vrec,
            const int* order) {
    const int full = sv.wide_stack_cap;
    const size_t shmem_full = size_t(full + 3) * BLOCK * sizeof(int);  // + the spare push slots
    if (shmem_full + BLOCK * sizeof(int) > device_max_lds())  // + the leaf round's marks
        throw std::runtime_error("BVH too deep for the LDS traversal stack");
    const int cap = retry ? std::min(full, std::max(1, sv.short_stack)) : full;
    int* const rl = cap < full ? retry : nullptr;
    const size_t shmem = size_t(cap + 3) * BLOCK * sizeof(int);
    if (stats) {
        hipLaunchKernelGGL((k_trace<ANY, true, false>), dim3(grid_for<ANY, true>(shmem)), dim3(BLOCK), shmem, s, sv, rays, n,
                           order, hits, occ, contrib, radiance, work, err, rl, work + 1, cap, stats, vrec);
    } else {
        hipLaunchKernelGGL((k_trace<ANY, false, false>), dim3(grid_for<ANY, false>(shmem)), dim3(BLOCK), shmem, s, sv, rays,
                           n, order, hits, occ, contrib, radiance, work, err, rl, work + 1, cap, stats, vrec);
    }
    check_hip(hipGetLastError(), "k_trace launch");
    if (!rl) return;
    // the abandoned rays (count in work[1], claims in work[2]): one block per CU is plenty
    const dim3 g(device_cu_count());
    if (stats) {
        hipLaunchKernelGGL((k_trace<ANY, true, true>), g, dim3(BLOCK), shmem_full, s, sv, rays, work + 1, rl, hits, occ,
                           contrib, radiance, work + 2, err, nullptr, nullptr, full, stats, vrec);
    } else {
        hipLaunchKernelGGL((k_trace<ANY, false, true>), g, dim3(BLOCK), shmem_full, s, sv, rays, work + 1, rl, hits, occ,
                           contrib, radiance, work + 2, err, nullptr, nullptr, full, stats, vrec);
    }
    check_hip(hipGetLastError(), "k_trace retry launch");
}

}  // namespace

size_t device_max_lds() {
    static int lds = 0;
    if (!lds) {
        int dev = 0;
        check_hip(hipGetDevice(&dev), "hipGetDevice");
        check_hip(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev), "LDS per block");
    }
    return size_t(lds);
}

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
                    TraceStats* stats, hipStream_t s, int* retry, const int* order) {
    launch<false>(sv, rays, n, hits, nullptr, nullptr, nullptr, work, err, stats, s, retry, nullptr, order);
}

void launch_shadow(const SceneView& sv, const Ray* rays, const int* n, uint8_t* occ, HitRec* hits,
                   const float4* contrib, float4* radiance, int* work, int* err, TraceStats* stats, hipStream_t s,
                   int* retry, HitRec* vrec, const int* order) {
    launch<true>(sv, rays, n, hits, occ, contrib, radiance, work, err, stats, s, retry, vrec, order);
}

void sort_by_direction(const Ray* rays, const int* n, int max_n, int mode, int* order, hipStream_t s) {
    if (max_n <= 0) return;
    if (mode == 1)
        hipLaunchKernelGGL((k_sort_dir<1024, true>), dim3((max_n + 1023) / 1024), dim3(SORT_BLOCK), 0, s, rays, n, order);
    else if (mode == 3)
        hipLaunchKernelGGL((k_sort_dir<256, false>), dim3((max_n + 255) / 256), dim3(SORT_BLOCK), 0, s, rays, n, order);
    else if (mode == 4)
        hipLaunchKernelGGL((k_sort_dir<1024, false>), dim3((max_n + 1023) / 1024), dim3(SORT_BLOCK), 0, s, rays, n, order);
    else
        hipLaunchKernelGGL((k_sort_dir<4096, false>), dim3((max_n + 4095) / 4096), dim3(SORT_BLOCK), 0, s, rays, n, order);
    check_hip(hipGetLastError(), "k_sort_dir");
}

}  // namespace nr
