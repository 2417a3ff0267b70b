// Ray-batch intersection that allocates nothing (nr_intersect_noalloc).
//
// For callers that cannot allocate device memory -- code under HIP graph capture, tight serving
// loops, harnesses that forbid hipMalloc.  The caller passes device arrays for the rays and the
// results; every buffer the traversal needs beyond them (the packed ray records, the hit records,
// the list of rays handed to the full-stack retry launch, the launch counters) lives in one
// file-scope device array, `scratch`, of NOALLOC_SCRATCH_BYTES (13 MiB + 256 B, room for
// NOALLOC_MAX_RAYS rays), which every GPU the library runs on holds until the process ends.
//
// In host code `scratch` names a host-side stand-in, not device memory: its device address is
// looked up once per device (nr_init, or the first call on a device that nr_init did not set up)
// with hipGetSymbolAddress while that device is current, kept per device, and that address is all
// host code ever passes on.  Each GPU has its own copy at its own address.
//
// Every call on one GPU shares that GPU's array and the array keeps the last call's contents; the
// counters are rewritten by a kernel at the start of each call.  Calls that overlap on one GPU
// (two host threads, or two scenes' streams) race on it.
#include "noalloc.h"
#include <stdexcept>
#include <string>
#include "scene.h"
#include "trace.h"

namespace nr {

namespace {

// scratch layout: counters | ray records | hit records | retry list, each 256-byte aligned
constexpr size_t OFF_RAYS = 256;
constexpr size_t OFF_HITS = OFF_RAYS + size_t(NOALLOC_MAX_RAYS) * sizeof(Ray);
constexpr size_t OFF_RETRY = OFF_HITS + size_t(NOALLOC_MAX_RAYS) * sizeof(HitRec);
static_assert(OFF_RETRY + size_t(NOALLOC_MAX_RAYS) * sizeof(int) == NOALLOC_SCRATCH_BYTES, "scratch layout");
static_assert(OFF_HITS % 256 == 0 && OFF_RETRY % 256 == 0, "scratch alignment");

constexpr int MAX_DEVICES = 64;
void* g_scratch[MAX_DEVICES] = {};  // device address of `scratch` per device, once looked up

// counters: [0] ray count, [1..3] the trace launch's work counters (ray claims, retry count, retry
// claims), [4] stack overflow, [5] a NaN segment length
__global__ void k_noalloc_begin(int* cnt, int n) {
    if (threadIdx.x < 8) cnt[threadIdx.x] = threadIdx.x == 0 ? n : 0;
}

__global__ void k_noalloc_pack(const float* __restrict__ o, const float* __restrict__ d,
                               const float* __restrict__ tmax, int n, Ray* __restrict__ rays, int* cnt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float w = 0.0f;
    if (tmax) {
        w = tmax[i];
        if (!(w == w)) {  // reported after the launch; the ray is traced as a zero-length segment
            atomicOr(cnt + 5, 1);
            w = 0.0f;
        }
    }
    rays[i].o = make_float4(o[3 * i], o[3 * i + 1], o[3 * i + 2], w);
    rays[i].d = make_float4(d[3 * i], d[3 * i + 1], d[3 * i + 2], 0.0f);
}

// hit records (TriGeom order) -> loader-order triangle ids and t, b1, b2
__global__ void k_noalloc_unpack(const TriGeom* __restrict__ tris, const HitRec* __restrict__ hits, int n,
                                 int32_t* __restrict__ tri, float* __restrict__ t, float* __restrict__ b1,
                                 float* __restrict__ b2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const HitRec h = hits[i];
    tri[i] = h.prim >= 0 ? __float_as_int(tris[h.prim].g2.y) : -1;
    t[i] = h.t;
    b1[i] = h.b1;
    b2[i] = h.b2;
}

}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

void noalloc_setup(int device) {
    if (device < 0 || device >= MAX_DEVICES) throw std::runtime_error("device index out of range");
    if (g_scratch[device]) return;
    void* ptr = nullptr;
    check_hip(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!ptr) throw std::runtime_error("hipGetSymbolAddress(scratch) returned a null address");
    g_scratch[device] = ptr;
}

void noalloc_forget() {
    for (void*& p : g_scratch) p = nullptr;
}

void intersect_noalloc(Scene& sc, const float* o, const float* d, uint32_t n, int any_hit, const float* tmax,
                       int32_t* tri, float* t, float* b1, float* b2) {
    if (!sc.built) throw std::runtime_error("BVH not built");
    if (n > uint32_t(NOALLOC_MAX_RAYS))
        throw std::runtime_error("nr_intersect_noalloc: " + std::to_string(n) + " rays exceed the maximum of " +
                                 std::to_string(NOALLOC_MAX_RAYS) + " per call");
    if (any_hit && !tmax) throw std::runtime_error("any_hit needs tmax");
    if (n == 0) return;
    if (!o || !d || !tri || !t || !b1 || !b2) throw std::runtime_error("null ray or result array");
    int dev = 0;
    check_hip(hipGetDevice(&dev), "hipGetDevice");
    if (dev != sc.device) throw std::runtime_error("the scene lives on another device than the current one");
    noalloc_setup(dev);
    unsigned char* base = static_cast<unsigned char*>(g_scratch[dev]);
    int* cnt = reinterpret_cast<int*>(base);
    Ray* rays = reinterpret_cast<Ray*>(base + OFF_RAYS);
    HitRec* hits = reinterpret_cast<HitRec*>(base + OFF_HITS);
    int* retry = reinterpret_cast<int*>(base + OFF_RETRY);
    hipStream_t s = sc.stream;
    const int grid = int((n + 255) / 256);
    hipLaunchKernelGGL(k_noalloc_begin, dim3(1), dim3(64), 0, s, cnt, int(n));
    hipLaunchKernelGGL(k_noalloc_pack, dim3(grid), dim3(256), 0, s, o, d, any_hit ? tmax : nullptr, int(n), rays,
                       cnt);
    if (any_hit)
        launch_shadow(sc.view, rays, cnt, nullptr, hits, nullptr, nullptr, cnt + 1, cnt + 4, nullptr, s, retry);
    else
        launch_closest(sc.view, rays, cnt, hits, cnt + 1, cnt + 4, nullptr, s, retry);
    hipLaunchKernelGGL(k_noalloc_unpack, dim3(grid), dim3(256), 0, s, sc.view.tris, hits, int(n), tri, t, b1, b2);
    check_hip(hipGetLastError(), "nr_intersect_noalloc launch");
    int flags[2] = {0, 0};
    check_hip(hipMemcpyAsync(flags, cnt + 4, sizeof(flags), hipMemcpyDeviceToHost, s), "flags D2H");
    check_hip(hipStreamSynchronize(s), "nr_intersect_noalloc sync");
    if (flags[1]) throw std::runtime_error("tmax must not be NaN");
    if (flags[0]) throw std::runtime_error("traversal stack overflow");
}

}  // namespace nr
