// BVH traversal kernels (closest hit and shadow any-hit) over the flattened SAH BVH.
#pragma once
#include "nr_device.h"

namespace nr {

// Per camera-sample path state, indexed by path slot.
struct alignas(16) PathState {
    float4 beta_fac;  // throughput xyz, MIS factor carried to the next emitter hit (w)
    float4 L_depth;   // radiance xyz, vertex depth (w, int bits)
};

struct TraceStats {
    unsigned long long rays, pair_fetch, tri_tests, pad;
};

// Closest hit for rays[0 .. *n): writes hits[i].  `work` must be zero at launch.
void launch_closest(const SceneView& sv, const Ray* rays, const int* n, HitRec* hits, int* work, int* err,
                    TraceStats* stats, hipStream_t s);
// Shadow test for rays[0 .. *n) with segment length rays[i].o.w: occluded iff some accepted
// triangle has (double)t + 2e-3 <= (double)len (rendererptls.cpp:70-71).  Writes occ[i] (0/1)
// and the first occluder found (prim / t) when `hits` is non-null; when `contrib` is non-null
// a visible ray adds contrib[i].xyz to state[slot].L, slot = bits of rays[i].d.w.
void launch_shadow(const SceneView& sv, const Ray* rays, const int* n, uint8_t* occ, HitRec* hits,
                   const float4* contrib, PathState* state, int* work, int* err, TraceStats* stats, hipStream_t s);

int device_cu_count();

}  // namespace nr
