// Ray-batch intersection that allocates nothing (see noalloc.hip).
#pragma once
#include <cstddef>
#include <cstdint>

namespace nr {

class Scene;

// Rays per nr_intersect_noalloc call, and the bytes of the per-GPU scratch array that holds their
// ray records (32 B), hit records (16 B) and retry-list entries (4 B) plus 256 B of counters.
constexpr int NOALLOC_MAX_RAYS = 1 << 18;
constexpr size_t NOALLOC_SCRATCH_BYTES = 256 + size_t(NOALLOC_MAX_RAYS) * (32 + 16 + 4);

// Looks up the scratch array's device address on `device` (which must be current); idempotent.
void noalloc_setup(int device);
// Forgets the looked-up addresses (nr_shutdown).
void noalloc_forget();
// nr_intersect on caller-owned device arrays, through the scratch array: no allocation.
void intersect_noalloc(Scene& sc, const float* o, const float* d, uint32_t n, int any_hit, const float* tmax,
                       int32_t* tri, float* t, float* b1, float* b2);

}  // namespace nr
