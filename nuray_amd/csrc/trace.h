// BVH traversal kernels (closest hit and shadow any-hit) over the flattened SAH BVH.
#pragma once
#include "nr_device.h"

namespace nr {

// Path state of a queued extension ray lives beside it: Ray.o.w = vertex depth (int bits),
// Ray.d.w = path slot (int bits), and a float4 (throughput xyz, MIS factor w) at the same queue
// position.  Only the radiance accumulator is indexed by path slot (float4, xyz used).

struct TraceStats {
    unsigned long long rays, node_fetch, tri_tests, retried;
};

// `work` points at three ints that must be zero at launch (ray claims, retry count, retry claims).
// With `retry` (room for *n ints) the persistent launch keeps a short LDS stack (more waves per
// CU); a ray whose stack would overflow it is abandoned, listed in `retry`, and traced again from
// the start by a second launch with the full-depth stack.  Without `retry` every ray gets the
// full-depth stack at once.  A ray that overflows the full-depth stack sets *err.
//
// Closest hit for rays[0 .. *n): writes hits[i].
// With `order`, rays are taken in the order rays[order[0]], rays[order[1]], ... (a permutation of
// 0 .. *n - 1); results still go to the rays' own indices.
void launch_closest(const SceneView& sv, const Ray* rays, const int* n, HitRec* hits, int* work, int* err,
                    TraceStats* stats, hipStream_t s, int* retry = nullptr, const int* order = nullptr);
// Shadow test for rays[0 .. *n) with segment length rays[i].o.w: occluded iff some accepted
// triangle has (double)t + 2e-3 <= (double)len (rendererptls.cpp:70-71).  Writes occ[i] (0/1)
// and the first occluder found (prim / t) when `hits` is non-null; when `contrib` is non-null
// a visible ray adds contrib[i].xyz to radiance[slot], slot = bits of rays[i].d.w, and, with vrec,
// sets VREC_VISIBLE in the vertex record vrec[bits of contrib[i].w] (when that index is >= 0).
void launch_shadow(const SceneView& sv, const Ray* rays, const int* n, uint8_t* occ, HitRec* hits,
                   const float4* contrib, float4* radiance, int* work, int* err, TraceStats* stats, hipStream_t s,
                   int* retry = nullptr, HitRec* vrec = nullptr, const int* order = nullptr);

// order[0 .. *n): queue positions grouped by ray direction inside chunks of consecutive
// positions, so that the rays a wave traces together point the same way.  mode 1: the 8 octants
// of rays[i].d within chunks of 1024; mode 2: 64 bins (each component quantised to 4) within
// chunks of 4096; mode 3: the 64 bins within chunks of 256; mode 4: within chunks of 1024.
// `max_n` bounds *n (sizes the grid).
void sort_by_direction(const Ray* rays, const int* n, int max_n, int mode, int* order, hipStream_t s);

int device_cu_count();
// Largest dynamic LDS allocation one workgroup may request on the current device, in bytes.
size_t device_max_lds();

}  // namespace nr
