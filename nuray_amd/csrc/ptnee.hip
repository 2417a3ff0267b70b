// Wavefront ptnee integrator for gfx950 (RendererPTLS::trace / ::render,
// src/renderer/rendererptls.cpp:13-236), made iterative:
//
//   raygen  -> camera rays for kb samples of every pixel (slot = s * npix + pixel)
//   repeat: closest-hit trace (persistent) -> shade -> shadow any-hit trace (persistent, adds
//           the NEE term of visible rays straight into the path's radiance)
//   finalize-> per-pixel FP64 sums of max(0, L) and of its square (the only FP64 on the path)
//
// shade() handles one path vertex: miss -> environment, emitter -> Le * fac (MIS weight of the
// previous vertex's NEE direction, rendererptls.cpp:33-36), back face -> stop, NEE with the
// reference's MIS weight and shadow-ray test, Russian roulette, BSDF sampling.  New shadow
// and extension rays are appended with one atomic per block and queue: ballot the lanes that
// emit, prefix-count them, and let one thread reserve the block's range.
//
// Random numbers come from Philox4x32-10 keyed by (seed, pixel, sample, vertex, slot), so any
// split of pixels or samples over batches or devices gives the same per-sample results.
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>
#include "nr_envis.h"
#include "nr_material.h"
#include "nrc.h"
#include "adjoint.h"
#include "render.h"

namespace nr {

namespace {
// Counters of the wavefront, each on a 128-byte line of its own: device-scope atomics on one line
// are serialised (about 11 ns each, 88 M/s on MI355X), while atomics on different lines proceed in
// parallel, so the shade step's two queue appends, and the ray claims of two trace launches that
// run side by side, do not queue behind each other.  C_WORK0 / C_WORK1 hold three ints: ray claims,
// retry count, retry claims of a trace launch.  The shadow queue's count and its launch's claims
// alternate by bounce parity (C_SHADOW / C_SHADOW_B, C_WORK1 / C_WORK1_B): bounce k's shadow launch
// may still run while bounce k + 1 resets its own.
constexpr int CSTRIDE = 32;  // ints per line
enum { C_RAY0 = 0, C_RAY1 = 1, C_SHADOW = 2, C_SHADOW_B = 3, C_ERR = 4, C_QUERY = 5, C_REC = 6, C_WORK0 = 7,
       C_WORK1 = 8, C_WORK1_B = 9, N_COUNTER_LINES = 10 };
__host__ __device__ constexpr int at(int line) { return line * CSTRIDE; }
__host__ __device__ constexpr int shadow_slot(int it) { return at((it & 1) ? C_SHADOW_B : C_SHADOW); }
__host__ __device__ constexpr int work1_slot(int it) { return at((it & 1) ? C_WORK1_B : C_WORK1); }
}  // namespace

CamView make_camera(const float pos[3], const float euler[3], float fov, float aspect) {
    CamView c;
    float cs[3], sn[3];
    for (int i = 0; i < 3; ++i) {
        cs[i] = float(std::cos(double(euler[i] / 180) * 3.14159));
        sn[i] = float(std::sin(double(euler[i] / 180) * 3.14159));
    }
    const float c1 = cs[0], s1 = sn[0], c2 = cs[1], s2 = sn[1], c3 = cs[2], s3 = sn[2];
    const float m12 = c3 * s1 * s2 - c1 * s3, m13 = c2 * s1, m22 = c2 * c3, m23 = -s2;
    const float m32 = s1 * s3 + c1 * c3 * s2, m33 = c1 * c2;
    const float up[3] = {m12, m22, m32}, gaze[3] = {-m13, -m23, -m33};
    for (int i = 0; i < 3; ++i) { c.pos[i] = pos[i]; c.up[i] = up[i]; c.gaze[i] = gaze[i]; }
    float h[3] = {gaze[1] * up[2] - gaze[2] * up[1], gaze[2] * up[0] - gaze[0] * up[2],
                  gaze[0] * up[1] - gaze[1] * up[0]};
    const float nrm = std::sqrt(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]) + 1e-8f;
    for (int i = 0; i < 3; ++i) c.hand[i] = h[i] / nrm;
    c.fh = float(2 * std::tan(double(fov) * 3.14159 / 180 / 2));
    c.fw = c.fh * aspect;
    return c;
}

RenderTarget::~RenderTarget() {
    if (pinned) (void)hipHostFree(pinned);
}

void RenderTarget::ensure(int w, int h, size_t c) {
    if (w != W || h != H) {
        sum.ensure(sizeof(double) * 3 * size_t(w) * h);
        sum2.ensure(sizeof(double) * 3 * size_t(w) * h);
        W = w; H = h;
        spp_total = 0;
    }
    ensure_queues(c);
}

void RenderTarget::ensure_queues(size_t c) {
    if (c > cap) {
        for (int k = 0; k < 2; ++k) {
            rays[k].ensure(sizeof(Ray) * c);
            beta[k].ensure(sizeof(float4) * c);
        }
        hits.ensure(sizeof(HitRec) * c);
        shadow.ensure(sizeof(Ray) * c);
        contrib.ensure(sizeof(float4) * c);
        radiance.ensure(sizeof(float4) * c);
        retry.ensure(sizeof(int) * c);
        retry_sh.ensure(sizeof(int) * c);
        cap = c;
    }
    counters.ensure(sizeof(int) * N_COUNTER_LINES * CSTRIDE);
    stats.ensure(sizeof(TraceStats) * 2);
    if (!pinned)
        check_hip(hipHostMalloc((void**)&pinned, sizeof(int) * N_COUNTER_LINES * CSTRIDE, hipHostMallocDefault),
                  "hipHostMalloc");
}

namespace {

constexpr int BLK = 256;
constexpr int LDS_LIGHTS = 2048;  // light CDFs up to this size are searched in LDS

__global__ void k_zero_doubles(double* a, double* b, size_t n) {
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x) {
        a[i] = 0.0;
        b[i] = 0.0;
    }
}

__global__ void k_clear_counters(int* c) {
    if (threadIdx.x < N_COUNTER_LINES * 3) c[at(threadIdx.x / 3) + threadIdx.x % 3] = 0;
}

// Before bounce `it`: the ray counter (raygen) or the next queue's counter, and the counters of
// this bounce's closest-hit and shadow launches and of its shadow queue.
__global__ void k_set_counters(int* c, int ray_slot, int ray_value, int clear_slot, int it) {
    if (threadIdx.x == 0) {
        if (ray_slot >= 0) c[at(ray_slot)] = ray_value;
        if (clear_slot >= 0) c[at(clear_slot)] = 0;
        c[shadow_slot(it)] = 0;
        for (int k = 0; k < 3; ++k) {
            c[at(C_WORK0) + k] = 0;
            c[work1_slot(it) + k] = 0;
        }
    }
}

// Camera ray of one sample: rendererptls.cpp:172 jitter (x + U, y + U), Camera::generateRay
// (camera.cpp:24-32).
__device__ __forceinline__ F3 camera_dir(const CamView& cam, int W, int H, int pixel, uint32_t seed, uint32_t sample) {
    const int x = pixel % W, y = pixel / W;
    const float4 u = rng4(seed, uint32_t(pixel), sample, 0u, 0u);
    const float ix = float(x) + u.x, iy = float(y) + u.y;
    const float cx = ix / float(W) * 2.0f - 1.0f;
    const float cy = iy / float(H) * 2.0f - 1.0f;
    const F3 gaze = f3(cam.gaze[0], cam.gaze[1], cam.gaze[2]);
    const F3 hand = f3(cam.hand[0], cam.hand[1], cam.hand[2]);
    const F3 up = f3(cam.up[0], cam.up[1], cam.up[2]);
    return normalized(sub(add(gaze, mul(mul(mul(hand, cam.fw), cx), 0.5f)), mul(mul(mul(up, cam.fh), cy), 0.5f)));
}

// Camera rays of a batch of kb = P / npix samples per pixel.  Queue position q holds sample
// q % kb of pixel q / kb: the rays a wave traces, and those the whole GPU has in flight at any
// moment, come from a few neighbouring pixels, so the nodes and triangles they reach stay in the
// L2s (and the compacting appends of later bounces keep that order).  The path's slot, which
// indexes its radiance and vertex records, stays sample offset * npix + pixel.  The throughput of
// every camera ray is 1 (the first shade step is given no throughput array).
__global__ void k_raygen(CamView cam, int W, int H, int npix, int P, uint32_t seed, uint32_t sample0, Ray* rays,
                         float4* radiance, float4* radiance_env) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= P) return;
    const int kb = P / npix;
    const int pixel = q / kb, k = q - pixel * kb;
    const int slot = k * npix + pixel;
    const F3 d = camera_dir(cam, W, H, pixel, seed, sample0 + uint32_t(k));
    rays[q].o = make_float4(cam.pos[0], cam.pos[1], cam.pos[2], __int_as_float(0));
    rays[q].d = make_float4(d.x, d.y, d.z, __int_as_float(slot));
    radiance[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);  // slots [0, P): every one is some q
    if (radiance_env) radiance_env[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// Sample index of training path i of a frame: its own Philox counters, disjoint from the render
// samples (bit 31 set).
__device__ __forceinline__ uint32_t train_sample(uint32_t frame_sample0, int T, int i) {
    return 0x80000000u | (frame_sample0 * uint32_t(T) + uint32_t(i));
}

// Radiance-cache render: slots [0, P) are the frame's camera samples as in k_raygen, slots
// [P, P + T) training paths from pixels drawn uniformly; every path starts with an empty
// spread state.
__global__ void k_raygen_nrc(CamView cam, int W, int H, int npix, int P, int T, uint32_t seed, uint32_t sample0,
                             Ray* rays, float4* beta, float4* radiance, float4* state, float4* radiance_env) {
    const int slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= P + T) return;
    int pixel;
    uint32_t sample;
    if (slot < P) {
        pixel = slot % npix;
        sample = sample0 + uint32_t(slot / npix);
    } else {
        sample = train_sample(sample0, T, slot - P);
        const uint4 h = philox(make_uint4(uint32_t(slot - P), sample0, 0x6E7263u, 0u), seed, 0x74726E67u);
        pixel = int(h.x % uint32_t(npix));
    }
    const F3 d = camera_dir(cam, W, H, pixel, seed, sample);
    rays[slot].o = make_float4(cam.pos[0], cam.pos[1], cam.pos[2], __int_as_float(0));
    rays[slot].d = make_float4(d.x, d.y, d.z, __int_as_float(slot));
    beta[slot] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    radiance[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (radiance_env) radiance_env[slot] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    state[slot] = make_float4(0.0f, 0.0f, 1.0f, __int_as_float(pixel));
}

// Radiance-cache part of the shade step.  Path state beside each queued ray: {sum of sqrt spread
// increments, sqrt(c a0), pdf of the direction that led here, pixel (int bits)}.
struct NrcShade {
    const float4* sin;
    float4* sout;
    NrcQuery* queries;   // cache queries: the path ends here and adds beta * cache(x, wo)
    NrcQuery* recs;      // training vertices (training paths only)
    float4* rec_lb;      // radiance of the path before the recorded vertex
    int* n_queries;
    int* n_recs;
    int query_cap, rec_cap;
    int P_render;        // slots >= P_render are training paths
    int T;
    float c_spread;
    int train_maxlen;
    int use_cache;
    int min_depth;
};

struct ShadeArgs {
    SceneView sv;
    int npix;
    uint32_t seed, sample0;
    float prr;
    int mis, maxdepth, rrdepth;
    int nee;  // 1: ptnee; 0: pt (rendererpt.cpp:13-61: no light sampling, so emitters keep fac = 1)
    const Ray* rin;
    const float4* bin;  // throughput and MIS factor per queued ray; null: (1, 1, 1, 1) for every ray
    const HitRec* hits;
    const int* n_in;
    float4* radiance;
    Ray* rout;
    float4* bout;
    int* n_out;
    Ray* shadow;
    float4* contrib;
    int* n_shadow;
    // sky sampling (ENV kernels): BSDF-escape weight per path slot, and the slot offset of the
    // radiance region that sky shadow rays add to (one writer per slot per shadow launch)
    float* skyw;
    int env_base;
    // vertex records for the adjoint (plain kernels only): vrec[slot * vrec_cap + depth], or null
    HitRec* vrec;
    int vrec_cap;
};

constexpr int NWAVE = BLK / 64;

// Stream compaction for the two output queues of one block-wide step, with a single atomic per
// queue per block: every wave ballots its emitting lanes, the per-wave counts meet in LDS, one
// lane reserves the block's range, and each lane's slot is the block base + the counts of the
// waves before it + the emitting lanes below it in its own wave.  (Device-scope atomics on one
// address serialise outside the per-XCD L2s, so they are kept to one per 256 lanes.)
struct BlockAppend {
    int cnt[2][2][NWAVE];  // [parity][queue][wave]
    int base[2][2];        // [parity][queue]
};
__device__ __forceinline__ void block_append2(BlockAppend& sa, int parity, bool e0, bool e1, int* c0, int* c1,
                                              int& p0, int& p1) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    const unsigned long long m0 = __ballot(e0), m1 = __ballot(e1);
    if (lane == 0) {
        sa.cnt[parity][0][wave] = __popcll(m0);
        sa.cnt[parity][1][wave] = __popcll(m1);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int t0 = 0, t1 = 0;
        for (int w = 0; w < NWAVE; ++w) { t0 += sa.cnt[parity][0][w]; t1 += sa.cnt[parity][1][w]; }
        sa.base[parity][0] = t0 ? atomicAdd(c0, t0) : 0;
        sa.base[parity][1] = t1 ? atomicAdd(c1, t1) : 0;
    }
    __syncthreads();
    int o0 = sa.base[parity][0], o1 = sa.base[parity][1];
    for (int w = 0; w < wave; ++w) { o0 += sa.cnt[parity][0][w]; o1 += sa.cnt[parity][1][w]; }
    p0 = o0 + __popcll(m0 & below);
    p1 = o1 + __popcll(m1 & below);
}

// The same for four queues (radiance-cache shading: shadow, extension, cache query, training record).
struct BlockAppend4 {
    int cnt[2][4][NWAVE];
    int base[2][4];
};
__device__ __forceinline__ void block_append4(BlockAppend4& sa, int parity, const bool (&e)[4], int* const (&c)[4],
                                              int (&pos)[4]) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    unsigned long long m[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) m[k] = __ballot(e[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 4; ++k) sa.cnt[parity][k][wave] = __popcll(m[k]);
    __syncthreads();
    if (threadIdx.x < 4) {
        const int k = threadIdx.x;
        int t = 0;
        for (int w = 0; w < NWAVE; ++w) t += sa.cnt[parity][k][w];
        sa.base[parity][k] = t ? atomicAdd(c[k], t) : 0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        int o = sa.base[parity][k];
        for (int w = 0; w < wave; ++w) o += sa.cnt[parity][k][w];
        pos[k] = o + __popcll(m[k] & below);
    }
}

// LightSampler::sampleLight over an LDS copy of the CDF (lower_bound, lightsampler.cpp:55-56).
__device__ __forceinline__ int pick_light_lds(const float* cdf, int n, float u) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (cdf[mid] < u) lo = mid + 1;
        else hi = mid;
    }
    return lo < n ? lo : n - 1;
}

template <unsigned F, bool NRC, bool ENV = false>
__device__ __forceinline__ void shade_body(const ShadeArgs& a, const NrcShade& nrc) {
    __shared__ float s_cdf[LDS_LIGHTS];
    const SceneView& sv = a.sv;
    const bool cdf_lds = sv.n_lights <= LDS_LIGHTS;
    if (cdf_lds)
        for (int k = threadIdx.x; k < sv.n_lights; k += blockDim.x) s_cdf[k] = sv.light_cdf[k];
    __syncthreads();
    __shared__ BlockAppend s_app;
    __shared__ BlockAppend4 s_app4;
    const int n = *a.n_in;
    const int stride = gridDim.x * blockDim.x;
    int parity = 0;
    for (int base = blockIdx.x * blockDim.x; base < n; base += stride, parity ^= 1) {
        const int j = base + int(threadIdx.x);
        const bool active = j < n;
        bool emit_sh = false, emit_next = false, emit_q = false, emit_rec = false, emit_env = false;
        int rec_idx = -1;  // this vertex's record (plain kernels with records)
        HitRec rec_h{};
        Ray sh_ray, next_ray, env_ray;
        float4 sh_c = make_float4(0, 0, 0, 0), next_b = make_float4(0, 0, 0, 0), env_c = make_float4(0, 0, 0, 0);
        float4 ns = make_float4(0, 0, 0, 0), rec_l = make_float4(0, 0, 0, 0);
        NrcQuery qrec;
        if (active) {
            const Ray r = a.rin[j];
            const HitRec h = a.hits[j];
            const float4 bf = a.bin ? a.bin[j] : make_float4(1.0f, 1.0f, 1.0f, 1.0f);  // null: camera rays
            const int slot = __float_as_int(r.d.w);
            const int depth = __float_as_int(r.o.w);
            if (!NRC && !ENV && a.vrec && depth < a.vrec_cap) {
                rec_idx = slot * a.vrec_cap + depth;
                rec_h = h;
            }
            const F3 beta = f3(bf.x, bf.y, bf.z);
            const float fac = bf.w;
            const F3 o = f3(r.o.x, r.o.y, r.o.z), d = f3(r.d.x, r.d.y, r.d.z);
            if (NRC) ns = nrc.sin[j];
            F3 Le = f3(0.0f, 0.0f, 0.0f);  // radiance this vertex adds directly
            if (h.prim < 0) {
                Le = mulv(beta, env_radiance(sv, d));
                if (ENV && depth > 0) Le = mul(Le, a.skyw[slot]);   // BSDF-escape MIS weight
            } else {
                const TriShade ts = sv.shade[h.prim];
                const float b1 = h.b1, b2 = h.b2, w0 = 1.0f - b1 - b2;
                const F3 n0 = f3(ts.s0.x, ts.s0.y, ts.s0.z), n1 = f3(ts.s0.w, ts.s1.x, ts.s1.y);
                const F3 n2 = f3(ts.s1.z, ts.s1.w, ts.s2.x);
                const F3 nrm = normalized(add(add(mul(n0, w0), mul(n1, b1)), mul(n2, b2)));  // triangle.cpp:35-38
                const Mat m = load_mat(sv, __float_as_int(ts.s3.w));
                if (m.type == MAT_LIGHT) {
                    Le = mulv(beta, mul(m.kd, fac));  // rendererptls.cpp:33-36
                } else if (!(dot(nrm, d) > 0.0f && !mat_transmissive(m))) {
                    const float tu = w0 * ts.s2.y + b1 * ts.s2.w + b2 * ts.s3.y;  // triangle.cpp:40-43
                    const float tv = w0 * ts.s2.z + b1 * ts.s3.x + b2 * ts.s3.z;
                    const F3 hp = add(o, mul(d, h.t));
                    const F3 wo = neg(d);
                    const int pixel = NRC ? __float_as_int(ns.w) : slot % a.npix;
                    const bool train = NRC && slot >= nrc.P_render;
                    const uint32_t sample = train ? train_sample(a.sample0, nrc.T, slot - nrc.P_render)
                                                  : a.sample0 + uint32_t(slot / a.npix);
                    bool stop = false;  // the cache answers for this vertex and everything after it
                    if (NRC) {
                        // Path spread (Mueller et al. 2021): a0 = |x0 - x1|^2 / (4 pi cos1) at the primary
                        // vertex, then sqrt(|x_i-1 - x_i|^2 / (p_i-1 |cos_i|)) summed; the render path
                        // stops when the sum exceeds sqrt(c a0).
                        const float cs = fabsf(dot(nrm, d));
                        const float t2 = h.t * h.t;
                        if (depth == 0) {
                            ns.x = 0.0f;
                            ns.y = sqrtf(nrc.c_spread * t2 / (4.0f * 3.14159265f * fmaxf(cs, 1e-6f)));
                        } else {
                            ns.x += sqrtf(t2 / fmaxf(ns.z * cs, 1e-20f));
                        }
                        bool query = false;
                        if (!train) query = nrc.use_cache && depth >= nrc.min_depth && ns.x > ns.y;
                        else query = depth + 1 >= nrc.train_maxlen;
                        if (query || train) {
                            float th, ph, nth, nph;
                            nrc_spherical(wo.x, wo.y, wo.z, th, ph);
                            nrc_spherical(nrm.x, nrm.y, nrm.z, nth, nph);
                            const F3 alb = mat_albedo<F>(sv, m, tu, tv);
                            qrec.a = make_float4(hp.x, hp.y, hp.z, m.type == MAT_LAMBERT ? 0.0f : m.alpha);
                            qrec.b = make_float4(th, ph, nth, nph);
                            qrec.c = make_float4(alb.x, alb.y, alb.z, __int_as_float(slot));
                            qrec.d = make_float4(beta.x, beta.y, beta.z, 0.0f);
                        }
                        if (train && !query) {  // the tail vertex answers from the cache: no target
                            emit_rec = true;
                            rec_l = a.radiance[slot];
                            if (ENV) {  // sky shadow rays of earlier vertices landed in the second region
                                const float4 E = a.radiance[a.env_base + slot];
                                rec_l.x += E.x;
                                rec_l.y += E.y;
                                rec_l.z += E.z;
                            }
                        }
                        if (query) {
                            emit_q = true;
                            stop = true;
                        }
                    }
                    const float4 ua = rng4(a.seed, uint32_t(pixel), sample, uint32_t(1 + depth), 0u);
                    // NEE direction (rendererptls.cpp:50-69): light pick, point on it, wl
                    const bool do_nee = !stop && a.nee && mat_needs_nee(m) && sv.n_lights > 0;
                    F3 wl = f3(0, 0, 0), lv = f3(0, 0, 0), ke = f3(0, 0, 0);
                    float lcos = 0.0f, ncos = 0.0f;
                    if (do_nee) {
                        const int li = cdf_lds ? pick_light_lds(s_cdf, sv.n_lights, ua.x) : pick_light(sv, ua.x);
                        const LightRec lr = sv.lights[li];
                        float r1 = ua.y, r2 = ua.z;
                        if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }
                        const float lw = 1.0f - r1 - r2;
                        const F3 lp = add(add(mul(f3(lr.p0.x, lr.p0.y, lr.p0.z), lw), mul(f3(lr.p1.x, lr.p1.y, lr.p1.z), r1)),
                                          mul(f3(lr.p2.x, lr.p2.y, lr.p2.z), r2));
                        const F3 ln = normalized(add(add(mul(f3(lr.n0.x, lr.n0.y, lr.n0.z), lw),
                                                         mul(f3(lr.n1.x, lr.n1.y, lr.n1.z), r1)),
                                                     mul(f3(lr.n2.x, lr.n2.y, lr.n2.z), r2)));
                        lv = sub(lp, hp);
                        wl = normalized(lv);
                        lcos = dot(ln, neg(wl));
                        ncos = dot(nrm, normalized(sub(lp, hp)));
                        ke = f3(lr.ke.x, lr.ke.y, lr.ke.z);
                    }
                    // Russian roulette and the continuation direction (rendererptls.cpp:80-88)
                    bool can_next = !stop;  // the BSDF strategy exists at this vertex (before roulette)
                    if (a.maxdepth > 0 && depth + 1 >= a.maxdepth) can_next = false;
                    bool do_next = can_next;
                    float prr = 1.0f;
                    if (depth >= a.rrdepth) {
                        do_next = do_next && !(ua.w > a.prr);
                        prr = a.prr;
                    }
                    F3 wi = f3(0, 0, 0);
                    if (do_next) {
                        const float4 ub = rng4(a.seed, uint32_t(pixel), sample, uint32_t(1 + depth), 1u);
                        wi = bsdf_sample<F>(m, ub.x, ub.y, ub.z, wo, nrm);
                    }
                    // sky sample (its own random dimension, so renders without it draw the same numbers)
                    const bool env_vertex = ENV && can_next && sv.env_w > 0 && env_sampling_vertex(m);
                    bool do_env = false;
                    F3 we = f3(0, 0, 0);
                    float pe = 0.0f;
                    if (env_vertex) {
                        const float4 ue = rng4(a.seed, uint32_t(pixel), sample, uint32_t(1 + depth), 2u);
                        do_env = env_sample(sv, ue, we, pe);
                    }
                    // f and pdf for every direction through one copy of the material code
                    F3 f_l = f3(0, 0, 0), f_i = f3(0, 0, 0), f_e = f3(0, 0, 0);
                    float p_l = 1.0f, p_i = 1.0f, p_e = 1.0f;
#pragma unroll 1
                    for (int k = 0; k < (ENV ? 3 : 2); ++k) {
                        const bool act = k == 0 ? do_nee : (k == 1 ? do_next : do_env);
                        if (!act) continue;
                        const F3 dir = k == 0 ? wl : (k == 1 ? wi : we);
                        const F3 fr = bsdf_eval<F>(sv, m, wo, nrm, dir, tu, tv);
                        const float pr = bsdf_pdf<F>(m, wo, nrm, dir);
                        if (k == 0) f_l = fr, p_l = pr;
                        else if (k == 1) f_i = fr, p_i = pr;
                        else f_e = fr, p_e = pr;
                    }
                    if (ENV && do_env) {  // g(we) L / (p_env + q), g = q / p_rep * f |cos|
                        const float qe = bsdf_sample_density(m, wo, nrm, we);
                        if (qe > 0.0f) {
                            const float w = qe / p_e * fabsf(dot(we, nrm)) / (pe + qe);
                            const F3 c = mul(mulv(mulv(beta, f_e), env_radiance(sv, we)), w);
                            if (c.x != 0.0f || c.y != 0.0f || c.z != 0.0f) {
                                emit_env = true;
                                const F3 eo = add(hp, mul(we, 1e-3f));
                                env_ray.o = make_float4(eo.x, eo.y, eo.z, 1e18f);   // blocked by any hit
                                env_ray.d = make_float4(we.x, we.y, we.z, __int_as_float(a.env_base + slot));
                                env_c = make_float4(c.x, c.y, c.z, 0.0f);
                            }
                        }
                    }
                    float fac_next = 1.0f;
                    if (do_nee) {  // rendererptls.cpp:61-77
                        const float lpdf = sv.light_p;
                        const float lop = lpdf * dot(lv, lv) / ((1e-6f < lcos) ? lcos : 1e-6f);
                        float mis = lop / (lop + p_l);
                        if (a.mis == 0) mis = 0.5f;
                        fac_next = 1.0f - mis;
                        const F3 Ll = mul(divs(ke, dot(lv, lv)), (0.0f < lcos ? lcos : 0.0f) / lpdf);
                        const F3 c = mul(mulv(mulv(beta, Ll), f_l), (0.0f < ncos ? ncos : 0.0f) * mis);
                        if (c.x != 0.0f || c.y != 0.0f || c.z != 0.0f) {
                            emit_sh = true;
                            const F3 so = add(hp, mul(wl, 1e-3f));
                            sh_ray.o = make_float4(so.x, so.y, so.z, norm(lv));
                            sh_ray.d = make_float4(wl.x, wl.y, wl.z, __int_as_float(slot));
                            sh_c = make_float4(c.x, c.y, c.z, 0.0f);
                        }
                    }
                    if (do_next) {  // rendererptls.cpp:86-91
                        const float c = fabsf(dot(wi, nrm));
                        const F3 nb = f3(beta.x * c * f_i.x / p_i / prr, beta.y * c * f_i.y / p_i / prr,
                                         beta.z * c * f_i.z / p_i / prr);
                        if (nb.x != 0.0f || nb.y != 0.0f || nb.z != 0.0f) {
                            emit_next = true;
                            const F3 no = add(hp, mul(wi, 1e-3f));
                            next_ray.o = make_float4(no.x, no.y, no.z, __int_as_float(depth + 1));
                            next_ray.d = make_float4(wi.x, wi.y, wi.z, __int_as_float(slot));
                            next_b = make_float4(nb.x, nb.y, nb.z, fac_next);
                            if (NRC) ns.z = p_i;
                            if (ENV) {
                                float sw = 1.0f;
                                if (env_vertex) {
                                    const float qi = bsdf_sample_density(m, wo, nrm, wi);
                                    const float pi_env = env_pdf(sv, wi);
                                    sw = qi + pi_env > 0.0f ? qi / (qi + pi_env) : 1.0f;
                                }
                                a.skyw[slot] = sw;
                            }
                        }
                    }
                }
            }
            if (Le.x != 0.0f || Le.y != 0.0f || Le.z != 0.0f) {
                float4 L = a.radiance[slot];
                L.x += Le.x;
                L.y += Le.y;
                L.z += Le.z;
                a.radiance[slot] = L;
            }
        }
        if (NRC) {
            const bool e[4] = {emit_sh, emit_next, emit_q, emit_rec};
            int* const c[4] = {a.n_shadow, a.n_out, nrc.n_queries, nrc.n_recs};
            int pos[4];
            block_append4(s_app4, parity, e, c, pos);
            if (emit_sh) {
                a.shadow[pos[0]] = sh_ray;
                a.contrib[pos[0]] = sh_c;
            }
            if (emit_next) {
                a.rout[pos[1]] = next_ray;
                a.bout[pos[1]] = next_b;
                nrc.sout[pos[1]] = ns;
            }
            if (emit_q && pos[2] < nrc.query_cap) nrc.queries[pos[2]] = qrec;
            if (emit_rec && pos[3] < nrc.rec_cap) {
                nrc.recs[pos[3]] = qrec;
                nrc.rec_lb[pos[3]] = rec_l;
            }
            if (ENV) {  // sky shadow rays share the shadow queue
                int pe, unused;
                block_append2(s_app, parity, emit_env, false, a.n_shadow, a.n_shadow, pe, unused);
                if (emit_env) {
                    a.shadow[pe] = env_ray;
                    a.contrib[pe] = env_c;
                }
            }
        } else if (ENV) {  // light and sky shadow rays share the shadow queue
            const bool e[4] = {emit_sh, emit_next, emit_env, false};
            int* const c[4] = {a.n_shadow, a.n_out, a.n_shadow, a.n_shadow};
            int pos[4];
            block_append4(s_app4, parity, e, c, pos);
            if (emit_sh) {
                a.shadow[pos[0]] = sh_ray;
                a.contrib[pos[0]] = sh_c;
            }
            if (emit_next) {
                a.rout[pos[1]] = next_ray;
                a.bout[pos[1]] = next_b;
            }
            if (emit_env) {
                a.shadow[pos[2]] = env_ray;
                a.contrib[pos[2]] = env_c;
            }
        } else {
            if (rec_idx >= 0)
                a.vrec[rec_idx] = HitRec{rec_h.t, rec_h.b1, rec_h.b2,
                                         rec_h.prim >= 0 && emit_sh ? (rec_h.prim | VREC_TRACED) : rec_h.prim};
            sh_c.w = __int_as_float(rec_idx);  // the shadow launch marks the record when visible
            int ps, pn;
            block_append2(s_app, parity, emit_sh, emit_next, a.n_shadow, a.n_out, ps, pn);
            if (emit_sh) {
                a.shadow[ps] = sh_ray;
                a.contrib[ps] = sh_c;
            }
            if (emit_next) {
                a.rout[pn] = next_ray;
                a.bout[pn] = next_b;
            }
        }
    }
}

// Lambert / GGX-reflect / light scenes without textures: 96 VGPRs, 5 waves per SIMD, no spills.
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(5, 8))) void k_shade_basic(ShadeArgs a) {
    shade_body<0u, false>(a, NrcShade{});
}
__global__ __launch_bounds__(BLK) void k_shade_full(ShadeArgs a) { shade_body<F_ALL, false>(a, NrcShade{}); }
// With sky sampling.
__global__ __launch_bounds__(BLK) void k_shade_basic_env(ShadeArgs a) { shade_body<0u, false, true>(a, NrcShade{}); }
__global__ __launch_bounds__(BLK) void k_shade_full_env(ShadeArgs a) { shade_body<F_ALL, false, true>(a, NrcShade{}); }
// Radiance-cache rendering (spread state, cache queries, training records).
__global__ __launch_bounds__(BLK) void k_shade_nrc_basic(ShadeArgs a, NrcShade n) { shade_body<0u, true>(a, n); }
__global__ __launch_bounds__(BLK) void k_shade_nrc_full(ShadeArgs a, NrcShade n) { shade_body<F_ALL, true>(a, n); }
__global__ __launch_bounds__(BLK) void k_shade_nrc_basic_env(ShadeArgs a, NrcShade n) { shade_body<0u, true, true>(a, n); }
__global__ __launch_bounds__(BLK) void k_shade_nrc_full_env(ShadeArgs a, NrcShade n) { shade_body<F_ALL, true, true>(a, n); }

// Adds the sky shadow rays' region into the main per-slot radiance.
__global__ void k_fold_env(float4* __restrict__ radiance, const float4* __restrict__ radiance_env, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 E = radiance_env[i];
    float4 L = radiance[i];
    L.x += E.x;
    L.y += E.y;
    L.z += E.z;
    radiance[i] = L;
}

// Per-pixel reduction of the batch's kb samples: sums of max(0, L) and its square in FP64
// (rendererptls.cpp:173-175: max(0, trace(...)) per component, which also maps NaN to 0).
__global__ void k_finalize(const float4* __restrict__ radiance, const float4* __restrict__ radiance_env, int npix,
                           int kb, double* __restrict__ sum, double* __restrict__ sum2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npix) return;
    double s[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
    for (int k = 0; k < kb; ++k) {
        float4 L = radiance[size_t(k) * npix + p];
        if (radiance_env) {  // sky shadow-ray contributions kept apart during the wavefront loop
            const float4 E = radiance_env[size_t(k) * npix + p];
            L.x += E.x;
            L.y += E.y;
            L.z += E.z;
        }
        const float v[3] = {0.0f < L.x ? L.x : 0.0f, 0.0f < L.y ? L.y : 0.0f, 0.0f < L.z ? L.z : 0.0f};
        for (int c = 0; c < 3; ++c) {
            s[c] += double(v[c]);
            s2[c] += double(v[c]) * double(v[c]);
        }
    }
    for (int c = 0; c < 3; ++c) {
        sum[3 * size_t(p) + c] += s[c];
        sum2[3 * size_t(p) + c] += s2[c];
    }
}

// Mean, then the reference's display transform (rendererptls.cpp:175-181): pow(1/2.2), *255,
// clamp to [0, 255], truncate.
__global__ void k_tonemap(const double* __restrict__ sum, size_t n3, double inv_spp, float* __restrict__ rgb,
                          uint8_t* __restrict__ rgb8) {
    const size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x;
    if (i >= n3) return;
    const float m = float(sum[i] * inv_spp);
    rgb[i] = m;
    float g = powf(m, float(1.0 / 2.2)) * 255.0f;
    g = (0.0f < g) ? g : 0.0f;
    g = (g < 255.0f) ? g : 255.0f;
    rgb8[i] = uint8_t(int(g));
}

}  // namespace

namespace {

void mark_event(std::vector<hipEvent_t>* v, hipStream_t s) {
    if (!v) return;
    hipEvent_t e;
    check_hip(hipEventCreate(&e), "hipEventCreate");
    check_hip(hipEventRecord(e, s), "hipEventRecord");
    v->push_back(e);
}

// One wavefront batch of P paths -- samples sample0, sample0 + 1, ... of every pixel, slot = sample
// offset * npix + pixel: raygen, then a closest-hit, a shade and a shadow launch per bounce until no
// path is left (or maxdepth).  Each path's radiance ends in t.radiance[slot] (the sky rays' share in
// rad_env with sky sampling).  A bounce's shadow launch runs on sc.stream2, beside the next
// bounce's closest-hit launch (the two share no data: shadow rays read the shadow queue and add to
// the radiance, closest hits read the next ray queue); the next shade waits for it, since it
// refills the shadow queue and adds to the same radiance.  Small batches gain most (their trace
// launches are bound by their slowest rays, and the shadow launch fills the idle CUs: 1080p x 1
// spp frames 10-14 % faster), the 720p x 64 spp frame 2.3 %.  Timed or counted runs (the ev_*
// vectors or ts non-null) keep every launch on sc.stream instead, so that each kernel's time and
// counts are its own.  Everything has joined sc.stream on return.  The ev_* vectors (null:
// untimed) collect start / stop events around the launches.  Returns 1 if cancelled between
// bounces.
int wavefront_batch(Scene& sc, RenderTarget& t, const CamView& cam, const RenderParams& p, uint32_t sample0, int P,
                    bool env, float4* rad_env, TraceStats* ts, volatile int* cancel, std::vector<hipEvent_t>* ev_c,
                    std::vector<hipEvent_t>* ev_h, std::vector<hipEvent_t>* ev_s, int& iterations,
                    HitRec* vrec = nullptr) {
    if (env) vrec = nullptr;  // records come from the plain shade kernels only
    const hipStream_t s = sc.stream;
    const int npix = p.W * p.H;
    int* cnt = t.counters.as<int>();
    const int grid_shade = device_cu_count() * 8;
    const hipStream_t s2 = (ev_c || ts) ? s : sc.stream2;
    if (sc.trace_order) t.order.ensure(sizeof(int) * 3 * t.cap);  // ray queue (cap), shadow queue (2 cap)
    if (P % npix != 0) throw std::runtime_error("wavefront batch of a partial sample per pixel");
    hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, s, cnt, int(C_RAY0), P, -1, 0);
    hipLaunchKernelGGL(k_raygen, dim3((P + BLK - 1) / BLK), dim3(BLK), 0, s, cam, p.W, p.H, npix, P, p.seed,
                       sample0, t.rays[0].as<Ray>(), t.radiance.as<float4>(), rad_env);
    auto join = [&]() { check_hip(hipStreamWaitEvent(s, sc.shadow_done, 0), "join shadow stream"); };
    for (int it = 0;; ++it) {
        const int cur = it & 1, nxt = cur ^ 1;
        int* const sh_cnt = cnt + shadow_slot(it);
        hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, s, cnt, -1, 0, int(C_RAY0 + nxt), it);
        // optional direction grouping of secondary and shadow rays (primary rays: pixel order)
        int* const ord = sc.trace_order ? t.order.as<int>() : nullptr;
        mark_event(ev_c, s);
        const bool sort_ext = ord && it > 0 && sc.trace_order <= 2;  // modes 3 and 4 sort shadow rays only
        if (sort_ext) sort_by_direction(t.rays[cur].as<Ray>(), cnt + at(C_RAY0 + cur), P, sc.trace_order, ord, s);
        launch_closest(sc.view, t.rays[cur].as<Ray>(), cnt + at(C_RAY0 + cur), t.hits.as<HitRec>(), cnt + at(C_WORK0),
                       cnt + at(C_ERR), ts, s, t.retry.as<int>(), sort_ext ? ord : nullptr);
        mark_event(ev_c, s);
        ShadeArgs a{sc.view, npix, p.seed, sample0, p.prr, p.mis, p.maxdepth, p.rrdepth,
                    p.integrator == 1 ? 0 : 1,
                    t.rays[cur].as<Ray>(), it == 0 ? nullptr : t.beta[cur].as<float4>(), t.hits.as<HitRec>(),
                    cnt + at(C_RAY0 + cur), t.radiance.as<float4>(), t.rays[nxt].as<Ray>(), t.beta[nxt].as<float4>(),
                    cnt + at(C_RAY0 + nxt), t.shadow.as<Ray>(), t.contrib.as<float4>(), sh_cnt,
                    env ? t.skyw.as<float>() : nullptr, env ? P : 0, vrec, VREC_CAP};
        if (it > 0) join();  // the previous bounce's shadow launch has read the shadow queue
        mark_event(ev_h, s);
        if (env) {
            if (sc.features == 0) hipLaunchKernelGGL(k_shade_basic_env, dim3(grid_shade), dim3(BLK), 0, s, a);
            else hipLaunchKernelGGL(k_shade_full_env, dim3(grid_shade), dim3(BLK), 0, s, a);
        } else if (sc.features == 0) {
            hipLaunchKernelGGL(k_shade_basic, dim3(grid_shade), dim3(BLK), 0, s, a);
        } else {
            hipLaunchKernelGGL(k_shade_full, dim3(grid_shade), dim3(BLK), 0, s, a);
        }
        mark_event(ev_h, s);
        check_hip(hipEventRecord(sc.shadow_ready, s), "shade done");
        check_hip(hipStreamWaitEvent(s2, sc.shadow_ready, 0), "shadow after shade");
        mark_event(ev_s, s2);
        if (ord) sort_by_direction(t.shadow.as<Ray>(), sh_cnt, env ? 2 * P : P, sc.trace_order, ord + t.cap, s2);
        launch_shadow(sc.view, t.shadow.as<Ray>(), sh_cnt, nullptr, nullptr, t.contrib.as<float4>(),
                      t.radiance.as<float4>(), cnt + work1_slot(it), cnt + at(C_ERR), ts ? ts + 1 : nullptr, s2,
                      t.retry_sh.as<int>(), vrec, ord ? ord + t.cap : nullptr);
        mark_event(ev_s, s2);
        check_hip(hipEventRecord(sc.shadow_done, s2), "shadow done");
        ++iterations;
        if (p.maxdepth > 0) {
            if (it + 1 >= p.maxdepth) break;
        } else {
            check_hip(hipMemcpyAsync(t.pinned, cnt + at(C_RAY0 + nxt), sizeof(int), hipMemcpyDeviceToHost, s), "cnt");
            check_hip(hipStreamSynchronize(s), "sync");
            if (*t.pinned == 0) break;
            if (cancel && *cancel) {
                join();
                return 1;
            }
        }
    }
    join();
    return 0;
}

}  // namespace

// Per-sample radiance of samples sample0 ... of every pixel (P paths, a whole number of samples per
// pixel) into t.radiance, slot = sample offset * npix + pixel, by the wavefront; nothing is
// accumulated.  The adjoint recomputes the radiance of its samples with it on scenes too large for
// the flat leaf list, where the wavefront is several times faster than one lane per path.
bool wavefront_radiance(Scene& sc, const CamView& cam, const RenderParams& p, uint32_t sample0, int P,
                        bool records) {
    if (!sc.target) sc.target = new RenderTarget();
    RenderTarget& t = *sc.target;
    t.ensure_queues(size_t(P));
    t.adj_valid = false;  // t.radiance no longer holds the last render's samples
    t.adj_vrec_valid = false;
    records = records && !p.env_is && size_t(P) * VREC_CAP * sizeof(HitRec) <= VREC_BUDGET;
    if (records) t.adj_vrec.ensure(size_t(P) * VREC_CAP * sizeof(HitRec));
    int* cnt = t.counters.as<int>();
    hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(64), 0, sc.stream, cnt);
    int iterations = 0;
    wavefront_batch(sc, t, cam, p, sample0, P, false, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                    iterations, records ? t.adj_vrec.as<HitRec>() : nullptr);
    check_hip(hipGetLastError(), "wavefront launch");
    check_hip(hipMemcpyAsync(t.pinned, cnt, sizeof(int) * at(C_REC + 1), hipMemcpyDeviceToHost, sc.stream), "counters");
    check_hip(hipStreamSynchronize(sc.stream), "wavefront sync");
    if (t.pinned[at(C_ERR)]) throw std::runtime_error("traversal stack overflow");
    return records;
}

int render_ptnee(Scene& sc, const CamView& cam, const RenderParams& p, ProgressFn cb, void* user,
                 volatile int* cancel, RenderStats* st, bool trace_stats, TraceStats* tstats, KernelTimes* kt) {
    if (!sc.built) throw std::runtime_error("scene not built (call nr_bvh_build_sah)");
    if (p.W <= 0 || p.H <= 0 || p.spp <= 0) throw std::runtime_error("bad image size or spp");
    const int npix = p.W * p.H;
    // Default batch: up to 64M paths in flight (~160 B of queues and state each, ~10 GB of the
    // 288 GB HBM), so a 720p frame at 64 spp runs as one wavefront and the short tail
    // iterations of Russian roulette are shared by every sample of the frame.
    int kb = p.batch_spp > 0 ? p.batch_spp : std::max(1, int((size_t(64) << 20) / size_t(npix)));
    kb = std::min(kb, p.spp);
    if (!sc.target) sc.target = new RenderTarget();
    RenderTarget& t = *sc.target;
    t.ensure(p.W, p.H, size_t(kb) * npix);
    // sky sampling: a second radiance region and shadow-queue half for the sky rays, the
    // BSDF-escape weights; without an environment sampler (no map, or a black one) it is a no-op
    const bool env = p.env_is && sc.view.env_w > 0;
    if (p.env_is && p.integrator != 0) throw std::runtime_error("environment sampling needs the ptnee integrator");
    if (p.path_kernel == 1 && p.env_is) throw std::runtime_error("the path kernel does not sample the sky (env_is)");
    // automatic choice (the image is the same either way): one lane per path for small scenes (a
    // flat leaf list) and for renders whose samples x triangles stay below AUTO_PATHS_WORK
    const bool paths = p.path_kernel == 1 ||
                       (p.path_kernel == 2 && !p.env_is &&
                        (sc.view.n_flat > 0 || double(npix) * p.spp * sc.n_tris() <= AUTO_PATHS_WORK));
    if (env) {
        const size_t c = size_t(kb) * npix;
        t.radiance.ensure(sizeof(float4) * 2 * c);
        t.shadow.ensure(sizeof(Ray) * 2 * c);
        t.contrib.ensure(sizeof(float4) * 2 * c);
        t.skyw.ensure(sizeof(float) * c);
        t.retry_sh.ensure(sizeof(int) * 2 * c);
    }
    t.adj_valid = kb >= p.spp && !p.accumulate && !p.env_is;
    t.adj_p = p;
    t.adj_cam = cam;
    t.adj_gen = sc.generation;
    // a one-batch render also keeps vertex records for the adjoint's replay
    const bool keep_vrec = t.adj_valid && size_t(kb) * npix * VREC_CAP * sizeof(HitRec) <= VREC_BUDGET;
    if (keep_vrec) t.adj_vrec.ensure(size_t(kb) * npix * VREC_CAP * sizeof(HitRec));
    t.adj_vrec_valid = keep_vrec;
    hipStream_t s = sc.stream;
    int* cnt = t.counters.as<int>();
    TraceStats* ts = trace_stats ? t.stats.as<TraceStats>() : nullptr;
    hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(64), 0, s, cnt);
    if (ts) check_hip(hipMemsetAsync(ts, 0, 2 * sizeof(TraceStats), s), "memset stats");
    if (!p.accumulate) {
        hipLaunchKernelGGL(k_zero_doubles, dim3(1024), dim3(BLK), 0, s, t.sum.as<double>(), t.sum2.as<double>(),
                           size_t(npix) * 3);
        t.spp_total = 0;
    }
    auto t0 = std::chrono::steady_clock::now();
    RenderStats rs;
    int status = 0;
    std::vector<hipEvent_t> ev_c, ev_s, ev_h;  // start/stop pairs per launch
    for (int b = 0; b < p.spp; b += kb) {
        if (cancel && *cancel) { status = 1; break; }
        const int k = std::min(kb, p.spp - b);
        const int P = k * npix;
        float4* const rad_env = env ? t.radiance.as<float4>() + P : nullptr;
        if (paths) {  // one lane per path: the same per-sample radiance without per-bounce launches
            mark_event(kt ? &ev_c : nullptr, s);
            render_paths(sc, cam, p, uint32_t(p.spp_offset + b), P, t.radiance.as<float4>(), cnt + at(C_ERR),
                         keep_vrec ? t.adj_vrec.as<HitRec>() : nullptr, VREC_CAP);
            mark_event(kt ? &ev_c : nullptr, s);
            rs.iterations++;
        } else {
            status = wavefront_batch(sc, t, cam, p, uint32_t(p.spp_offset + b), P, env, rad_env, ts, cancel,
                                     kt ? &ev_c : nullptr, kt ? &ev_h : nullptr, kt ? &ev_s : nullptr, rs.iterations,
                                     keep_vrec ? t.adj_vrec.as<HitRec>() : nullptr);
        }
        if (status) break;
        hipLaunchKernelGGL(k_finalize, dim3((npix + BLK - 1) / BLK), dim3(BLK), 0, s, t.radiance.as<float4>(), rad_env,
                           npix, k, t.sum.as<double>(), t.sum2.as<double>());
        check_hip(hipGetLastError(), "ptnee launch");
        t.spp_total += k;
        rs.samples += (long long)P;
        if (cb) {
            check_hip(hipStreamSynchronize(s), "sync");
            cb(float(b + k) / float(p.spp), user);
        }
    }
    check_hip(hipMemcpyAsync(t.pinned, cnt, sizeof(int) * at(C_REC + 1), hipMemcpyDeviceToHost, s), "counters");
    check_hip(hipStreamSynchronize(s), "render sync");
    if (status) t.adj_valid = false;
    if (t.pinned[at(C_ERR)]) throw std::runtime_error("traversal stack overflow");
    rs.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (kt) {
        *kt = KernelTimes{};
        auto total = [](std::vector<hipEvent_t>& v, double& ms, int& launches) {
            for (size_t i = 0; i + 1 < v.size(); i += 2) {
                float x = 0;
                check_hip(hipEventElapsedTime(&x, v[i], v[i + 1]), "elapsed");
                ms += x;
                launches++;
            }
        };
        total(ev_c, kt->closest_ms, kt->closest_launches);
        total(ev_s, kt->shadow_ms, kt->shadow_launches);
        total(ev_h, kt->shade_ms, kt->shade_launches);
    }
    for (auto* v : {&ev_c, &ev_s, &ev_h})
        for (hipEvent_t e : *v) (void)hipEventDestroy(e);
    if (ts && tstats) check_hip(hipMemcpy(tstats, ts, 2 * sizeof(TraceStats), hipMemcpyDeviceToHost), "stats");
    if (st) *st = rs;
    return status;
}

int render_nrc(Scene& sc, Nrc& nrc, const CamView& cam, const RenderParams& p, const NrcRenderParams& np,
               NrcFrameStats* fs, RenderStats* st) {
    if (!sc.built) throw std::runtime_error("scene not built (call nr_bvh_build_sah)");
    if (p.W <= 0 || p.H <= 0 || p.spp <= 0) throw std::runtime_error("bad image size or spp");
    if (np.train_paths < 0 || np.train_maxlen < 1) throw std::runtime_error("bad radiance-cache parameters");
    if (p.path_kernel == 1) throw std::runtime_error("the radiance cache renders with the wavefront (path_kernel = 0)");
    if (p.env_is && p.integrator != 0) throw std::runtime_error("environment sampling needs the ptnee integrator");
    const bool env = p.env_is && sc.view.env_w > 0;
    const int npix = p.W * p.H;
    const size_t P = size_t(p.spp) * npix, T = size_t(np.train_paths);
    if (P + T > (size_t(64) << 20)) throw std::runtime_error("radiance-cache frame larger than one wavefront batch");
    if (!sc.target) sc.target = new RenderTarget();
    RenderTarget& t = *sc.target;
    const size_t cap = P + T, rec_cap = std::max<size_t>(1, T * size_t(np.train_maxlen));
    t.ensure(p.W, p.H, cap);
    for (int k = 0; k < 2; ++k) t.nstate[k].ensure(sizeof(float4) * cap);
    t.nrc_queries.ensure(sizeof(NrcQuery) * cap);
    t.nrc_recs.ensure(sizeof(NrcQuery) * rec_cap);
    t.nrc_lb.ensure(sizeof(float4) * rec_cap);
    t.nrc_tgt.ensure(sizeof(float4) * rec_cap);
    if (env) {  // sky shadow rays: second radiance region and shadow-queue half, BSDF-escape weights
        t.radiance.ensure(sizeof(float4) * 2 * cap);
        t.shadow.ensure(sizeof(Ray) * 2 * cap);
        t.contrib.ensure(sizeof(float4) * 2 * cap);
        t.skyw.ensure(sizeof(float) * cap);
        t.retry_sh.ensure(sizeof(int) * 2 * cap);
    }
    float4* const rad_env = env ? t.radiance.as<float4>() + cap : nullptr;
    t.adj_valid = false;
    hipStream_t s = sc.stream;
    int* cnt = t.counters.as<int>();
    hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(64), 0, s, cnt);
    if (!p.accumulate) {
        hipLaunchKernelGGL(k_zero_doubles, dim3(1024), dim3(BLK), 0, s, t.sum.as<double>(), t.sum2.as<double>(),
                           size_t(npix) * 3);
        t.spp_total = 0;
    }
    const int grid_shade = device_cu_count() * 8;
    auto t0 = std::chrono::steady_clock::now();
    RenderStats rs;
    const hipStream_t s2 = sc.stream2;  // the shadow launches beside the next closest-hit ones
    auto join = [&]() { check_hip(hipStreamWaitEvent(s, sc.shadow_done, 0), "join shadow stream"); };
    hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, s, cnt, int(C_RAY0), int(cap), -1, 0);
    hipLaunchKernelGGL(k_raygen_nrc, dim3((cap + BLK - 1) / BLK), dim3(BLK), 0, s, cam, p.W, p.H, npix, int(P), int(T),
                       p.seed, uint32_t(p.spp_offset), t.rays[0].as<Ray>(), t.beta[0].as<float4>(),
                       t.radiance.as<float4>(), t.nstate[0].as<float4>(), rad_env);
    for (int it = 0;; ++it) {
        const int cur = it & 1, nxt = cur ^ 1;
        int* const sh_cnt = cnt + shadow_slot(it);
        hipLaunchKernelGGL(k_set_counters, dim3(1), dim3(64), 0, s, cnt, -1, 0, int(C_RAY0 + nxt), it);
        launch_closest(sc.view, t.rays[cur].as<Ray>(), cnt + at(C_RAY0 + cur), t.hits.as<HitRec>(), cnt + at(C_WORK0),
                       cnt + at(C_ERR), nullptr, s, t.retry.as<int>());
        ShadeArgs a{sc.view, npix, p.seed, uint32_t(p.spp_offset), p.prr, p.mis, p.maxdepth, p.rrdepth,
                    p.integrator == 1 ? 0 : 1,
                    t.rays[cur].as<Ray>(), t.beta[cur].as<float4>(), t.hits.as<HitRec>(), cnt + at(C_RAY0 + cur),
                    t.radiance.as<float4>(), t.rays[nxt].as<Ray>(), t.beta[nxt].as<float4>(), cnt + at(C_RAY0 + nxt),
                    t.shadow.as<Ray>(), t.contrib.as<float4>(), sh_cnt, env ? t.skyw.as<float>() : nullptr,
                    env ? int(cap) : 0};
        NrcShade ns{t.nstate[cur].as<float4>(), t.nstate[nxt].as<float4>(), t.nrc_queries.as<NrcQuery>(),
                    t.nrc_recs.as<NrcQuery>(), t.nrc_lb.as<float4>(), cnt + at(C_QUERY), cnt + at(C_REC), int(cap),
                    int(rec_cap), int(P), int(T), np.c_spread, np.train_maxlen, np.use_cache,
                    std::max(1, np.min_depth)};
        if (it > 0) join();  // the previous bounce's shadow launch has read the shadow queue
        if (env) {
            if (sc.features == 0) hipLaunchKernelGGL(k_shade_nrc_basic_env, dim3(grid_shade), dim3(BLK), 0, s, a, ns);
            else hipLaunchKernelGGL(k_shade_nrc_full_env, dim3(grid_shade), dim3(BLK), 0, s, a, ns);
        } else if (sc.features == 0) {
            hipLaunchKernelGGL(k_shade_nrc_basic, dim3(grid_shade), dim3(BLK), 0, s, a, ns);
        } else {
            hipLaunchKernelGGL(k_shade_nrc_full, dim3(grid_shade), dim3(BLK), 0, s, a, ns);
        }
        check_hip(hipEventRecord(sc.shadow_ready, s), "shade done");
        check_hip(hipStreamWaitEvent(s2, sc.shadow_ready, 0), "shadow after shade");
        launch_shadow(sc.view, t.shadow.as<Ray>(), sh_cnt, nullptr, nullptr, t.contrib.as<float4>(),
                      t.radiance.as<float4>(), cnt + work1_slot(it), cnt + at(C_ERR), nullptr, s2, t.retry_sh.as<int>());
        check_hip(hipEventRecord(sc.shadow_done, s2), "shadow done");
        rs.iterations++;
        if (p.maxdepth > 0) {
            if (it + 1 >= p.maxdepth) break;
        } else {
            check_hip(hipMemcpyAsync(t.pinned, cnt + at(C_RAY0 + nxt), sizeof(int), hipMemcpyDeviceToHost, s), "cnt");
            check_hip(hipStreamSynchronize(s), "sync");
            if (*t.pinned == 0) break;
        }
    }
    join();
    if (env)
        hipLaunchKernelGGL(k_fold_env, dim3((cap + BLK - 1) / BLK), dim3(BLK), 0, s, t.radiance.as<float4>(), rad_env,
                           int(cap));
    // the cache answers every stopped path (render paths and training-path tails)
    nrc.infer_accumulate(t.nrc_queries.as<NrcQuery>(), cnt + at(C_QUERY), int(cap), t.radiance.as<float4>(), s);
    hipLaunchKernelGGL(k_finalize, dim3((npix + BLK - 1) / BLK), dim3(BLK), 0, s, t.radiance.as<float4>(), nullptr, npix,
                       p.spp, t.sum.as<double>(), t.sum2.as<double>());
    check_hip(hipGetLastError(), "nrc frame launch");
    t.spp_total += p.spp;
    rs.samples = (long long)P;
    check_hip(hipMemcpyAsync(t.pinned, cnt, sizeof(int) * at(C_REC + 1), hipMemcpyDeviceToHost, s), "counters");
    check_hip(hipStreamSynchronize(s), "nrc frame sync");
    if (t.pinned[at(C_ERR)]) throw std::runtime_error("traversal stack overflow");
    NrcFrameStats f;
    f.queries = t.pinned[at(C_QUERY)];
    f.records = std::min<long long>(t.pinned[at(C_REC)], (long long)rec_cap);
    t.nrc_last_records = int(std::min<long long>(f.records, nrc.cfg.max_batch));
    if (t.nrc_last_records > 0)  // targets (L_path - L_before) / beta of the frame's records
        nrc_targets(t.nrc_recs.as<NrcQuery>(), t.nrc_lb.as<float4>(), t.radiance.as<float4>(), t.nrc_last_records,
                    t.nrc_tgt.as<float4>(), s);
    if (np.train && f.records > 0) {
        const int n = t.nrc_last_records;
        // vertices are recorded bounce by bounce: interleave them so every batch mixes all depths
        const int steps = std::max(1, std::min(np.train_steps, n));
        t.nrc_batch.ensure(sizeof(NrcQuery) * size_t(n));
        t.nrc_batch_tgt.ensure(sizeof(float4) * size_t(n));
        nrc_interleave(t.nrc_recs.as<NrcQuery>(), t.nrc_tgt.as<float4>(), n, steps, t.nrc_batch.as<NrcQuery>(),
                       t.nrc_batch_tgt.as<float4>(), s);
        double loss = 0.0;
        for (int k = 0, b0 = 0; k < steps; ++k) {
            const int cnt = n / steps + (k < n % steps ? 1 : 0);
            const bool last = k + 1 == steps;
            loss = nrc.train(t.nrc_batch.as<NrcQuery>() + b0, t.nrc_batch_tgt.as<float4>() + b0, cnt,
                             last && fs != nullptr, nullptr, s);
            b0 += cnt;
        }
        f.train_loss = loss;  // of the last step
        f.trained = n;
    }
    check_hip(hipStreamSynchronize(s), "nrc train sync");
    rs.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (fs) *fs = f;
    if (st) *st = rs;
    return 0;
}

int nrc_frame_records(Scene& sc, float* recs, float* targets, int cap) {
    if (!sc.target) return 0;
    RenderTarget& t = *sc.target;
    const int n = std::min(cap, t.nrc_last_records);
    if (n > 0) {
        check_hip(hipMemcpyAsync(recs, t.nrc_recs.ptr, sizeof(NrcQuery) * size_t(n), hipMemcpyDeviceToHost, sc.stream),
                  "records D2H");
        check_hip(hipMemcpyAsync(targets, t.nrc_tgt.ptr, sizeof(float4) * size_t(n), hipMemcpyDeviceToHost, sc.stream),
                  "targets D2H");
        check_hip(hipStreamSynchronize(sc.stream), "records sync");
    }
    return t.nrc_last_records;
}

void render_fetch(Scene& sc, float* rgb, double* sum, double* sum2, uint8_t* rgb8) {
    if (!sc.target) throw std::runtime_error("nothing rendered");
    RenderTarget& t = *sc.target;
    const size_t n3 = size_t(t.W) * t.H * 3;
    hipStream_t s = sc.stream;
    if (sum) check_hip(hipMemcpyAsync(sum, t.sum.ptr, n3 * sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
    if (sum2) check_hip(hipMemcpyAsync(sum2, t.sum2.ptr, n3 * sizeof(double), hipMemcpyDeviceToHost, s), "D2H");
    if (rgb || rgb8) {
        DevBuf drgb, d8;
        drgb.ensure(n3 * sizeof(float));
        d8.ensure(n3);
        hipLaunchKernelGGL(k_tonemap, dim3((n3 + BLK - 1) / BLK), dim3(BLK), 0, s, t.sum.as<double>(), n3,
                           t.spp_total ? 1.0 / double(t.spp_total) : 0.0, drgb.as<float>(), d8.as<uint8_t>());
        if (rgb) check_hip(hipMemcpyAsync(rgb, drgb.ptr, n3 * sizeof(float), hipMemcpyDeviceToHost, s), "D2H");
        if (rgb8) check_hip(hipMemcpyAsync(rgb8, d8.ptr, n3, hipMemcpyDeviceToHost, s), "D2H");
        check_hip(hipStreamSynchronize(s), "fetch sync");
    }
    check_hip(hipStreamSynchronize(s), "fetch sync");
}

}  // namespace nr
