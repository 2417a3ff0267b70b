// Differentiable ptnee: gradients of an image loss with respect to per-material albedo (RGB)
// and GGX roughness, by path-replay backpropagation (Vicini, Speierer, Jakob, SIGGRAPH 2021).
// There is no reference counterpart; the estimator being differentiated is the reference's
// RendererPTLS::trace (rendererptls.cpp:13-93) with its quirks.
//
// What is differentiated: the expectation of the per-sample estimator.  Albedo enters each
// sample multilinearly, so its per-sample derivative is exact.  For GGX roughness the sampled
// direction is held fixed (detached) and the log-derivative of the sampler's density is added:
// the reference sampler draws theta_h = 0.9999 * atan(alpha * sqrt(u / (1 - u)))
// (matggxrefl.cpp:24-29), whose density in theta0 = theta_h / 0.9999 is
// 2 alpha^2 tan(theta0) sec^2(theta0) / (alpha^2 + tan^2 theta0)^2, so
//   d log p / d alpha = 2 / alpha - 4 alpha / (alpha^2 + tan^2 theta0)
// (the reflection and 0.9999 Jacobians do not depend on alpha).  The MIS weight of NEE and the
// factor `fac` carried to the next emitter hit depend on alpha through the reference's pdf
// (matggxrefl.cpp:54-64) and are differentiated as well.
//
// Replay: the backward kernel re-derives each path from the same Philox counters and the same
// exact traversal, and peels each vertex's own contribution off the sample's radiance L, so the
// radiance still to come through the next bounce is known at every vertex without storing any
// path.  L is the forward render's per-sample radiance when that render ran as one batch with
// the same samples; otherwise each chunk of samples is rendered forward first (recompute-only
// mode, any number of samples: one lane per path on small scenes, the wavefront on the others).
// Lanes take their samples from a shared counter, one atomic per wave and round.  Material
// gradients accumulate per block in FP32 in LDS, then into FP64; texel gradients (albedo
// textures) go straight to FP64 atomics at the four bilinear corners.
#include <cstring>
#include "adjoint.h"
#include "nr_material.h"
#include "nr_traverse.h"
#include "render.h"

namespace nr {

namespace {

constexpr int ABLK = 256;
constexpr int LDS_PARAMS = 4096;  // per-block FP32 gradient accumulators (materials x 4)

// ---- derivatives of the reference's GGX-reflect model (matggxrefl.cpp:36-64)
struct GgxParts {
    F3 f;          // bxdf value
    F3 df_dkd;     // d f_c / d Kd_c
    F3 df_da;      // d f_c / d alpha
    float pdf;     // reference pdf
    float dpdf_da; // d pdf / d alpha
};
__device__ __forceinline__ void ggx_d_and_deriv(float a, float c, float& D, float& dD) {
    const float k = c * c * (a * a - 1.0f) + 1.0f;
    D = a * a / 3.14159f / (k * k);
    dD = 2.0f * a / (3.14159f * k * k * k) * (k - 2.0f * a * a * c * c);
}
__device__ __forceinline__ void ggx_g_and_deriv(float a, float c, float& G, float& dG) {
    const float kap = a * 0.5f + 0.5f;
    const float S = sqrtf(kap * kap + (1.0f - kap * kap * c * c));
    G = 2.0f * c / (c + S);
    const float dS = kap * (1.0f - c * c) / (2.0f * S);
    dG = -2.0f * c / ((c + S) * (c + S)) * dS;
}
__device__ __forceinline__ GgxParts ggx_parts(const Mat& m, F3 kd, F3 wo, F3 n, F3 wi) {
    GgxParts r;
    r.f = r.df_dkd = r.df_da = f3(0, 0, 0);
    r.pdf = bsdf_pdf<0u>(m, wo, n, wi);
    r.dpdf_da = 0.0f;
    const float ci = dot(n, wi), co = dot(n, wo);
    if (!(double(ci) <= 1e-4 || co <= 1e-4f)) {
        const F3 wh = normalized(add(wo, wi));
        const float ch = dot(n, wh);
        float D, dD;
        ggx_d_and_deriv(m.alpha, ch, D, dD);
        r.dpdf_da = dD * ch / 4.0f;
    }
    if (dot(wi, n) * dot(wo, n) < 0.0f) return r;
    const F3 wh = normalized(add(wo, wi));
    float D, dD, Gl, dGl, Gv, dGv;
    ggx_d_and_deriv(m.alpha, dot(n, wh), D, dD);
    ggx_g_and_deriv(m.alpha, ci, Gl, dGl);
    ggx_g_and_deriv(m.alpha, co, Gv, dGv);
    const float fr = powf(1.0f - dot(wh, wi), 5.0f);
    const float den = 4.0f * (co + 1e-6f) * (ci + 1e-6f);
    const float s = D * Gl * Gv / 4.0f / (co + 1e-6f) / (ci + 1e-6f);
    const float ds = (dD * Gl * Gv + D * dGl * Gv + D * Gl * dGv) / den;
    const F3 F = f3(kd.x + (1.0f - kd.x) * fr, kd.y + (1.0f - kd.y) * fr, kd.z + (1.0f - kd.z) * fr);
    r.f = mul(F, s);
    r.df_dkd = f3(s * (1.0f - fr), s * (1.0f - fr), s * (1.0f - fr));
    r.df_da = mul(F, ds);
    return r;
}

// Per-material parts for the two models the adjoint supports; the albedo (Kd or F0) comes from
// the material row or its texture.
template <unsigned F>
__device__ __forceinline__ GgxParts mat_parts(const SceneView& sv, const Mat& m, F3 wo, F3 n, F3 wi, float u, float v) {
    if (m.type == MAT_GGXREFL) return ggx_parts(m, mat_albedo<F>(sv, m, u, v), wo, n, wi);
    GgxParts r;
    r.f = bsdf_eval<F>(sv, m, wo, n, wi, u, v);
    r.pdf = bsdf_pdf<F>(m, wo, n, wi);
    if (m.type == MAT_GGX || m.type == MAT_BLINNPHONG || m.type == MAT_GLASS) {  // replayed, not differentiated
        r.df_dkd = r.df_da = f3(0, 0, 0);
        r.dpdf_da = 0.0f;
        return r;
    }
    const bool same = !(dot(wi, n) * dot(wo, n) < 0.0f);
    const float inv = same ? 1.0f / 3.14159f : 0.0f;  // Lambert: f = Kd / pi
    r.df_dkd = f3(inv, inv, inv);
    r.df_da = f3(0, 0, 0);
    r.dpdf_da = 0.0f;
    return r;
}

// f and pdf at one vertex: with derivatives for the gradient pass; for the forward pass exactly
// the wavefront shade kernel's bsdf_eval / bsdf_pdf, so that a forward replay reproduces the
// wavefront's per-sample radiance bit for bit.
template <unsigned F, bool GRAD>
__device__ __forceinline__ GgxParts vertex_parts(const SceneView& sv, const Mat& m, F3 wo, F3 n, F3 wi, float u,
                                                 float v) {
    if (GRAD) return mat_parts<F>(sv, m, wo, n, wi, u, v);
    GgxParts r;
    r.f = bsdf_eval<F>(sv, m, wo, n, wi, u, v);
    r.pdf = bsdf_pdf<F>(m, wo, n, wi);
    r.df_dkd = r.df_da = f3(0.0f, 0.0f, 0.0f);
    r.dpdf_da = 0.0f;
    return r;
}

// d log p(wi) / d alpha of the reference GGX sampler at a fixed direction.
__device__ __forceinline__ float ggx_score(float a, F3 wo, F3 n, F3 wi) {
    const F3 h = normalized(add(wo, wi));
    const float ch = fminf(fmaxf(dot(n, h), -1.0f), 1.0f);
    const float th0 = acosf(ch) / 0.9999f;
    if (!(th0 < 1.5707963f)) return 0.0f;
    const float t = tanf(th0);
    return 2.0f / a - 4.0f * a / (a * a + t * t);
}

struct AdjArgs {
    SceneView sv;
    CamView cam;
    int W, H, npix, P;
    uint32_t seed, sample0;
    float prr;
    int mis, maxdepth, rrdepth;
    const float4* radiance;  // per-sample radiance of the forward pass (slot layout), or null: recompute it
    int slot0, slot1;        // samples [slot0, slot1) of this launch
    const float* dldi;       // dLoss / dImage, W*H*3
    float inv_spp;
    int n_params;            // materials * 4
    double* grad;            // n_params
    int* err;
    int tex_id;              // texture whose texel gradient is wanted (-1: none)
    double* gtex;            // w * h * 3 texel gradient (linear texel values)
    int nee;                 // 1: ptnee, 0: pt (no light sampling)
    int rad_off;             // radiance[] holds samples from this index on
    unsigned* claim;         // samples claimed so far in this launch (zero at launch); a launch covers
                             // fewer than 2^31 samples and each wave claims once more at most after
                             // the last one is handed out, so the count stays below 2^32
    HitRec* vrec;            // per-vertex records of the first vrec_cap vertices of each sample, at
    int vrec_cap;            // vrec[(slot - rad_off) * vrec_cap + depth], or null
};

// Vertex records (HitRec with the VREC_* flags): the forward path kernel and the wavefront write
// them; the replay of the same samples reads them instead of tracing both rays again, and traces
// the shadow ray itself where the record has none.

// Per-wave sample source: a private range [rb, re) of claim offsets, refilled from the launch's
// counter.  Device-scope atomics on one address serialise (about 11 ns each), so one atomic per
// round of every wave bounds a launch of short steps (the replay reads its vertices instead of
// tracing them: 1.27 -> 0.68 ms per 256^2 x 32 spp replay with 128-sample refills).  Guided
// refills (a quarter of the samples left per wave, at most 128) suit long paths (the forward
// render): the ranges shrink towards the end of the launch, so no wave is left holding many
// samples while others idle.  A refill is never smaller than what the asking lanes need.
struct SampleSource {
    unsigned rb = 0, re = 0, chunk = 1, waves = 1;
    bool drained = false, guided = false;
};
__device__ __forceinline__ unsigned refill_size(unsigned left, unsigned waves, bool guided) {
    return max(1u, min(128u, guided ? left / (4u * waves) : (left + waves - 1) / waves));
}
__device__ __forceinline__ SampleSource sample_source(const AdjArgs& a, bool guided) {
    SampleSource src;
    src.waves = gridDim.x * (blockDim.x / 64);
    src.guided = guided;
    src.chunk = refill_size(unsigned(a.slot1 - a.slot0), src.waves, guided);
    return src;
}

// Hands the next samples of [a.slot0, a.slot1) to the lanes of the wave that ask for one (need),
// in lane order, from the wave's private range (claiming more when it runs short).  Called by every
// lane of the wave at the same point.  Returns the lane's new sample, or a.slot1 (no more samples)
// when it asked; else keeps `slot`.
__device__ __forceinline__ int claim_samples(const AdjArgs& a, SampleSource& src, bool need, int slot) {
    const int lane = int(threadIdx.x & 63);
    const unsigned long long ask = __ballot(need);
    if (ask == 0ull) return slot;
    const unsigned n_ask = unsigned(__popcll(ask)), rank = unsigned(__popcll(ask & ((1ull << lane) - 1ull)));
    const unsigned have = src.re - src.rb;
    unsigned pick;  // this lane's claim offset (wave-uniform bookkeeping below)
    if (n_ask <= have || src.drained) {
        pick = rank < have ? src.rb + rank : ~0u;
        src.rb += min(n_ask, have);
    } else {
        const unsigned want = max(src.chunk, n_ask - have);
        unsigned base = 0;
        if (lane == 0) base = atomicAdd(a.claim, want);
        base = __shfl(base, 0, 64);
        const unsigned total = unsigned(a.slot1 - a.slot0);
        pick = rank < have ? src.rb + rank : base + (rank - have);
        src.rb = base + (n_ask - have);
        src.re = base + want;
        if (base + want >= total) src.drained = true;
        else if (src.guided) src.chunk = refill_size(total - (base + want), src.waves, true);
    }
    if (!need) return slot;
    return pick < unsigned(a.slot1 - a.slot0) ? a.slot0 + int(pick) : a.slot1;
}

__device__ __forceinline__ void gadd(float* lds, double* glob, bool use_lds, int idx, float v) {
    if (v == 0.0f || !(v == v)) return;
    if (use_lds) atomicAdd(lds + idx, v);
    else atomicAdd(glob + idx, double(v));
}

// Texture::pixelUV / pixelBi (texture.cpp:54-72) run backwards: the albedo gradient goes to the
// four bilinear corner texels with their weights; texels outside the image (read as black) get none.
__device__ __forceinline__ void scatter_texels(const AdjArgs& a, const SceneView& sv, int tex, float u, float v, F3 dalb) {
    if (!a.gtex || tex != a.tex_id) return;
    const TexDesc t = sv.tex[tex];
    const float fu = fmodf(fmodf(u, 1.0f) + 1.0f, 1.0f);
    const float fv = fmodf(fmodf(1.0f - v, 1.0f) + 1.0f, 1.0f);
    const float x = fu * float(t.w), y = fv * float(t.h);
    const int x0 = int(floorf(x)), y0 = int(floorf(y));
    const int x1 = x0 + 1, y1 = y0 + 1;
    const int xs[4] = {x1, x1, x0, x0}, ys[4] = {y1, y0, y1, y0};
    const float ws[4] = {(x - float(x0)) * (y - float(y0)), (x - float(x0)) * (float(y1) - y),
                         (float(x1) - x) * (y - float(y0)), (float(x1) - x) * (float(y1) - y)};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (xs[k] < 0 || ys[k] < 0 || xs[k] >= t.w || ys[k] >= t.h || ws[k] == 0.0f) continue;
        double* g = a.gtex + (size_t(ys[k]) * t.w + xs[k]) * 3;
        if (dalb.x != 0.0f) atomicAdd(g + 0, double(dalb.x * ws[k]));
        if (dalb.y != 0.0f) atomicAdd(g + 1, double(dalb.y * ws[k]));
        if (dalb.z != 0.0f) atomicAdd(g + 2, double(dalb.z * ws[k]));
    }
}

// One sample's path, replayed from its Philox counters one vertex at a time.  GRAD = false: the
// sample's radiance (the forward estimator, accumulated in vertex order as the wavefront does).
// GRAD = true: the parameter gradient of g . max(0, L), with Lt the sample's radiance.  The kernels
// drive path_step from a flat loop in which a lane whose path has ended starts its next sample at
// once (path regeneration), so a wave does not wait at each sample for its longest path.
struct PathState {
    F3 o, d, beta;
    F3 L;         // forward: radiance so far; gradient: radiance not yet accounted for
    F3 g;         // gradient: dLoss / d(sample radiance) per channel (zero in the forward replay)
    float fac, dfac;  // MIS factor for the next emitter hit and d/d alpha of it
    int prev_mat, depth, pixel;
    uint32_t sample;
    int rslot;    // slot - rad_off: the sample's index in the radiance and vertex-record buffers
};

// Camera ray of sample `slot` (identical to k_raygen) and the path's starting state.
__device__ __forceinline__ void path_begin(const AdjArgs& a, int slot, F3 Lt, F3 g, PathState& p) {
    p.pixel = slot % a.npix;
    p.sample = a.sample0 + uint32_t(slot / a.npix);
    p.rslot = slot - a.rad_off;
    const int x = p.pixel % a.W, y = p.pixel / a.W;
    const float4 u0 = rng4(a.seed, uint32_t(p.pixel), p.sample, 0u, 0u);
    const float cx = (float(x) + u0.x) / float(a.W) * 2.0f - 1.0f;
    const float cy = (float(y) + u0.y) / float(a.H) * 2.0f - 1.0f;
    const CamView& cam = a.cam;
    p.d = normalized(sub(add(f3(cam.gaze[0], cam.gaze[1], cam.gaze[2]),
                             mul(mul(mul(f3(cam.hand[0], cam.hand[1], cam.hand[2]), cam.fw), cx), 0.5f)),
                         mul(mul(mul(f3(cam.up[0], cam.up[1], cam.up[2]), cam.fh), cy), 0.5f)));
    p.o = f3(cam.pos[0], cam.pos[1], cam.pos[2]);
    p.beta = f3(1.0f, 1.0f, 1.0f);
    p.L = Lt;
    p.g = g;
    p.fac = 1.0f;
    p.dfac = 0.0f;
    p.prev_mat = -1;
    p.depth = 0;
}

// One vertex of the path; returns false when the path has ended.
template <unsigned F, bool GRAD>
__device__ __forceinline__ bool path_step(const AdjArgs& a, PathState& p, int* stk, float* gacc, bool use_lds,
                                          bool& overflow) {
    const SceneView& sv = a.sv;
    auto add_albedo = [&](const Mat& m, int mid, float tu, float tv, F3 v) {
        if (m.tex < 0) {
            gadd(gacc, a.grad, use_lds, 4 * mid + 0, v.x);
            gadd(gacc, a.grad, use_lds, 4 * mid + 1, v.y);
            gadd(gacc, a.grad, use_lds, 4 * mid + 2, v.z);
        } else if (F & F_TEX) {
            scatter_texels(a, sv, m.tex, tu, tv, v);
        }
    };
    HitRec* const rec = a.vrec && p.depth < a.vrec_cap ? a.vrec + (size_t(p.rslot) * a.vrec_cap + p.depth) : nullptr;
    HitRec h;
    int32_t rec_flags = 0;
    if (GRAD && rec) {
        h = *rec;
        if (h.prim >= 0) {
            rec_flags = h.prim & (VREC_TRACED | VREC_VISIBLE);
            h.prim &= ~(VREC_TRACED | VREC_VISIBLE);
        }
    } else {
        h = trace_ray<false, ABLK>(sv, p.o, p.d, 0.0f, stk, overflow);
        if (!GRAD && rec) *rec = h;
    }
    if (h.prim < 0) {  // environment: no parameter dependence
        if (!GRAD) {
            const F3 e = mulv(p.beta, env_radiance(sv, p.d));
            p.L = add(p.L, e);
        }
        return false;
    }
    const TriShade ts = sv.shade[h.prim];
    const float b1 = h.b1, b2 = h.b2, w0 = 1.0f - b1 - b2;
    const F3 nrm = normalized(add(add(mul(f3(ts.s0.x, ts.s0.y, ts.s0.z), w0), mul(f3(ts.s0.w, ts.s1.x, ts.s1.y), b1)),
                                  mul(f3(ts.s1.z, ts.s1.w, ts.s2.x), b2)));
    const int mid = __float_as_int(ts.s3.w);
    const Mat m = load_mat(sv, mid);
    if (m.type == MAT_LIGHT) {
        if (!GRAD) {
            p.L = add(p.L, mulv(p.beta, mul(m.kd, p.fac)));
        } else if (p.prev_mat >= 0 && p.dfac != 0.0f) {
            // Le * fac: fac depends on the previous vertex's roughness through its NEE MIS weight
            const F3 g = p.g, beta = p.beta;
            const float v = g.x * beta.x * m.kd.x * p.dfac + g.y * beta.y * m.kd.y * p.dfac + g.z * beta.z * m.kd.z * p.dfac;
            gadd(gacc, a.grad, use_lds, 4 * p.prev_mat + 3, v);
        }
        return false;
    }
    if (dot(nrm, p.d) > 0.0f && !mat_transmissive(m)) return false;
    const float tu = w0 * ts.s2.y + b1 * ts.s2.w + b2 * ts.s3.y;
    const float tv = w0 * ts.s2.z + b1 * ts.s3.x + b2 * ts.s3.z;
    const F3 hp = add(p.o, mul(p.d, h.t));
    const F3 wo = neg(p.d);
    const F3 g = p.g, beta = p.beta;
    const float4 ua = rng4(a.seed, uint32_t(p.pixel), p.sample, uint32_t(1 + p.depth), 0u);
    float fac_next = 1.0f, dfac_next = 0.0f;
    if (a.nee && mat_needs_nee(m) && sv.n_lights > 0) {
        const LightRec lr = sv.lights[pick_light(sv, ua.x)];
        float r1 = ua.y, r2 = ua.z;
        if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }
        const float lw = 1.0f - r1 - r2;
        const F3 lp = add(add(mul(f3(lr.p0.x, lr.p0.y, lr.p0.z), lw), mul(f3(lr.p1.x, lr.p1.y, lr.p1.z), r1)),
                          mul(f3(lr.p2.x, lr.p2.y, lr.p2.z), r2));
        const F3 ln = normalized(add(add(mul(f3(lr.n0.x, lr.n0.y, lr.n0.z), lw), mul(f3(lr.n1.x, lr.n1.y, lr.n1.z), r1)),
                                     mul(f3(lr.n2.x, lr.n2.y, lr.n2.z), r2)));
        const F3 lv = sub(lp, hp);
        const F3 wl = normalized(lv);
        const float lcos = dot(ln, neg(wl));
        const float ncos = dot(nrm, normalized(sub(lp, hp)));
        const GgxParts pl = vertex_parts<F, GRAD>(sv, m, wo, nrm, wl, tu, tv);
        const float lpdf = sv.light_p;
        const float lop = lpdf * dot(lv, lv) / ((1e-6f < lcos) ? lcos : 1e-6f);
        float mis = lop / (lop + pl.pdf);
        float dmis = -lop / ((lop + pl.pdf) * (lop + pl.pdf)) * pl.dpdf_da;
        if (a.mis == 0) mis = 0.5f, dmis = 0.0f;
        fac_next = 1.0f - mis;
        dfac_next = -dmis;
        const F3 Ll = mul(divs(f3(lr.ke.x, lr.ke.y, lr.ke.z), dot(lv, lv)), (0.0f < lcos ? lcos : 0.0f) / lpdf);
        const float cn = (0.0f < ncos ? ncos : 0.0f);
        const F3 c = mul(mulv(mulv(beta, Ll), pl.f), cn * mis);
        if (c.x != 0.0f || c.y != 0.0f || c.z != 0.0f) {
            bool visible;
            if (GRAD && (rec_flags & VREC_TRACED)) {
                visible = (rec_flags & VREC_VISIBLE) != 0;
            } else {
                const F3 so = add(hp, mul(wl, 1e-3f));
                visible = trace_ray<true, ABLK>(sv, so, wl, norm(lv), stk, overflow).prim < 0;
                if (!GRAD && rec) rec->prim = h.prim | VREC_TRACED | (visible ? VREC_VISIBLE : 0);
            }
            if (visible) {  // visible: d(beta Ll f cos mis)
                if (!GRAD) {
                    p.L = add(p.L, c);
                } else {
                    const F3 k = mul(mulv(beta, Ll), cn);
                    add_albedo(m, mid, tu, tv,
                               f3(g.x * k.x * pl.df_dkd.x * mis, g.y * k.y * pl.df_dkd.y * mis,
                                  g.z * k.z * pl.df_dkd.z * mis));
                    const float va = g.x * k.x * (pl.df_da.x * mis + pl.f.x * dmis) +
                                     g.y * k.y * (pl.df_da.y * mis + pl.f.y * dmis) +
                                     g.z * k.z * (pl.df_da.z * mis + pl.f.z * dmis);
                    gadd(gacc, a.grad, use_lds, 4 * mid + 3, va);
                    p.L = sub(p.L, c);
                }
            }
        }
    }
    bool cont = true;
    float prr = 1.0f;
    if (p.depth >= a.rrdepth) {
        cont = !(ua.w > a.prr);
        prr = a.prr;
    }
    if (a.maxdepth > 0 && p.depth + 1 >= a.maxdepth) cont = false;
    if (!cont) return false;
    const float4 ub = rng4(a.seed, uint32_t(p.pixel), p.sample, uint32_t(1 + p.depth), 1u);
    const F3 wi = bsdf_sample<F>(m, ub.x, ub.y, ub.z, wo, nrm);
    const GgxParts pi = vertex_parts<F, GRAD>(sv, m, wo, nrm, wi, tu, tv);
    const float cw = fabsf(dot(wi, nrm));
    const F3 nb = f3(beta.x * cw * pi.f.x / pi.pdf / prr, beta.y * cw * pi.f.y / pi.pdf / prr,
                     beta.z * cw * pi.f.z / pi.pdf / prr);
    if (GRAD) {
        // everything still in p.L (the radiance not yet accounted for) arrives through this bounce:
        // d log T = d log f - d log pdf + score
        const F3 Lrem = p.L;
        const float score = m.type == MAT_GGXREFL ? ggx_score(m.alpha, wo, nrm, wi) : 0.0f;
        const float dlp = pi.pdf > 0.0f ? pi.dpdf_da / pi.pdf : 0.0f;
        float va = 0.0f;
        F3 dal = f3(0.0f, 0.0f, 0.0f);
        if (pi.f.x != 0.0f) {
            dal.x = g.x * Lrem.x * pi.df_dkd.x / pi.f.x;
            va += g.x * Lrem.x * (pi.df_da.x / pi.f.x - dlp + score);
        }
        if (pi.f.y != 0.0f) {
            dal.y = g.y * Lrem.y * pi.df_dkd.y / pi.f.y;
            va += g.y * Lrem.y * (pi.df_da.y / pi.f.y - dlp + score);
        }
        if (pi.f.z != 0.0f) {
            dal.z = g.z * Lrem.z * pi.df_dkd.z / pi.f.z;
            va += g.z * Lrem.z * (pi.df_da.z / pi.f.z - dlp + score);
        }
        if (m.type == MAT_LAMBERT || m.type == MAT_GGXREFL) add_albedo(m, mid, tu, tv, dal);
        if (m.type == MAT_GGXREFL) gadd(gacc, a.grad, use_lds, 4 * mid + 3, va);
    }
    if (nb.x == 0.0f && nb.y == 0.0f && nb.z == 0.0f) return false;
    p.beta = nb;
    p.fac = fac_next;
    p.dfac = dfac_next;
    p.prev_mat = mid;
    p.o = add(hp, mul(wi, 1e-3f));
    p.d = wi;
    ++p.depth;
    return true;
}

// Persistent grid over samples [slot0, slot1) with their radiance in a.radiance[slot - a.rad_off]
// (the last single-batch render, or a forward pass over the chunk just before).
template <unsigned F>
__global__ __launch_bounds__(ABLK) void k_adjoint(AdjArgs a) {
    extern __shared__ char smem[];
    int* stk = reinterpret_cast<int*>(smem) + threadIdx.x;
    float* gacc = reinterpret_cast<float*>(smem + size_t(a.sv.lane_slots) * ABLK * sizeof(int));
    const bool use_lds = a.n_params <= LDS_PARAMS;
    if (use_lds)
        for (int k = threadIdx.x; k < a.n_params; k += blockDim.x) gacc[k] = 0.0f;
    __syncthreads();
    bool overflow = false;
    PathState p;
    // Each round every lane that has no path asks for the next sample (claim_samples); a sample
    // whose weight is zero -- max(0, L) per channel at the camera (rendererptls.cpp:173) has zero
    // slope where L <= 0 or NaN -- ends at once and the lane asks again next round.
    int slot = a.slot0;
    bool need = true, active = false;
    SampleSource src = sample_source(a, false);
    for (;;) {
        slot = claim_samples(a, src, need, slot);
        if (need) {
            need = false;
            if (slot < a.slot1) {
                const int pixel = slot % a.npix;
                const float4 r = a.radiance[slot - a.rad_off];
                const F3 g = f3(r.x > 0.0f ? a.dldi[3 * pixel] * a.inv_spp : 0.0f,
                                r.y > 0.0f ? a.dldi[3 * pixel + 1] * a.inv_spp : 0.0f,
                                r.z > 0.0f ? a.dldi[3 * pixel + 2] * a.inv_spp : 0.0f);
                active = g.x != 0.0f || g.y != 0.0f || g.z != 0.0f;
                need = !active;
                if (active) path_begin(a, slot, f3(r.x, r.y, r.z), g, p);
            }
        }
        if (!__any(active || need)) break;
        if (active && !path_step<F, true>(a, p, stk, gacc, use_lds, overflow)) {
            active = false;
            need = true;
        }
    }
    if (overflow) atomicOr(a.err, 1);
    __syncthreads();
    if (use_lds)
        for (int k = threadIdx.x; k < a.n_params; k += blockDim.x)
            if (gacc[k] != 0.0f) atomicAdd(a.grad + k, double(gacc[k]));
}

// Forward-only render, one lane per path: each lane replays its samples' paths to the end (the
// radiance is the wavefront's, bit for bit: same Philox numbers, same traversal, same BSDF code,
// same order of additions) and writes it to radiance[slot - a.rad_off].  For small frames of long
// paths (inverse rendering on a closed room) this avoids the wavefront's per-bounce launches and
// tails.
template <unsigned F>
__global__ __launch_bounds__(ABLK) __attribute__((amdgpu_waves_per_eu(4))) void k_render_paths(AdjArgs a, float4* __restrict__ radiance) {
    extern __shared__ char smem[];
    int* stk = reinterpret_cast<int*>(smem) + threadIdx.x;
    bool overflow = false;
    PathState p;
    int slot = a.slot0;
    bool need = true, active = false;
    SampleSource src = sample_source(a, true);
    for (;;) {
        slot = claim_samples(a, src, need, slot);
        if (need) {
            need = false;
            active = slot < a.slot1;
            if (active) path_begin(a, slot, f3(0.0f, 0.0f, 0.0f), f3(0.0f, 0.0f, 0.0f), p);
        }
        if (!__any(active)) break;
        if (active && !path_step<F, false>(a, p, stk, nullptr, false, overflow)) {
            radiance[slot - a.rad_off] = make_float4(p.L.x, p.L.y, p.L.z, 0.0f);
            active = false;
            need = true;
        }
    }
    if (overflow) atomicOr(a.err, 1);
}

// L2 loss (mean over pixels and channels) of the rendered mean against the target, and its
// gradient with respect to the mean image.
__global__ void k_loss_l2(const double* __restrict__ sum, double inv_spp, const float* __restrict__ target, size_t n3,
                          float* __restrict__ dldi, double* __restrict__ loss) {
    double part = 0.0;
    const double inv_n = 1.0 / double(n3);
    for (size_t i = blockIdx.x * size_t(blockDim.x) + threadIdx.x; i < n3; i += size_t(gridDim.x) * blockDim.x) {
        const double r = sum[i] * inv_spp - double(target[i]);
        part += r * r;
        dldi[i] = float(2.0 * r * inv_n);
    }
    // block reduction in LDS, one FP64 atomic per block
    __shared__ double red[256];
    red[threadIdx.x] = part;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0] * inv_n);
}

__global__ void k_zero_int(unsigned* a) {
    if (threadIdx.x == 0) *a = 0;
}

__global__ void k_zero(double* a, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = 0.0;
}

}  // namespace

void set_material_params(Scene& sc, const float* theta, int n_mats) {
    if (n_mats != int(sc.mats.size() / 8)) throw std::runtime_error("parameter rows must match the material count");
    bool changed = false;
    for (int j = 0; j < n_mats; ++j) {
        float* m = &sc.mats[8 * size_t(j)];
        if (int(m[0]) == MAT_LIGHT) continue;  // emitters keep their Ke
        for (int k = 0; k < 4; ++k) {
            changed |= std::memcmp(&m[1 + k], &theta[4 * j + k], sizeof(float)) != 0;
            m[1 + k] = theta[4 * j + k];
        }
    }
    // setting the parameters a render already used keeps its stored samples for the adjoint
    if (!changed) return;
    ++sc.generation;
    if (!sc.built) return;
    const std::vector<MatRec> mr = sc.mat_records();
    sc.d_mats.upload(mr.data(), mr.size() * sizeof(MatRec), sc.stream);
    check_hip(hipStreamSynchronize(sc.stream), "params upload");
}

double loss_l2(Scene& sc, const float* target, float* dldi_host) {
    if (!sc.target || sc.target->spp_total == 0) throw std::runtime_error("render before computing the loss");
    RenderTarget& t = *sc.target;
    const size_t n3 = size_t(t.W) * t.H * 3;
    t.adj_target.upload(target, n3 * sizeof(float), sc.stream);
    t.adj_dldi.ensure(n3 * sizeof(float));
    t.adj_loss.ensure(sizeof(double));
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, sc.stream, t.adj_loss.as<double>(), 1);
    hipLaunchKernelGGL(k_loss_l2, dim3(256), dim3(256), 0, sc.stream, t.sum.as<double>(), 1.0 / double(t.spp_total),
                       t.adj_target.as<float>(), n3, t.adj_dldi.as<float>(), t.adj_loss.as<double>());
    double loss = 0;
    check_hip(hipMemcpyAsync(&loss, t.adj_loss.ptr, sizeof(double), hipMemcpyDeviceToHost, sc.stream), "D2H");
    if (dldi_host)
        check_hip(hipMemcpyAsync(dldi_host, t.adj_dldi.ptr, n3 * sizeof(float), hipMemcpyDeviceToHost, sc.stream), "D2H");
    check_hip(hipStreamSynchronize(sc.stream), "loss sync");
    return loss;
}

namespace {

// Launch shape of k_render_paths for the scene's material features (as the shade kernels).
struct PathsShape {
    unsigned feat;
    int grid;
    size_t shmem;
};
PathsShape paths_shape(const Scene& sc) {
    PathsShape L;
    L.shmem = size_t(sc.view.lane_slots) * ABLK * sizeof(int);
    if (L.shmem > device_max_lds()) throw std::runtime_error("BVH too deep for the LDS traversal stack");
    L.feat = sc.features & (F_GGX | F_EXT) ? F_ALL : (sc.features & F_TEX ? F_TEX : 0u);
    int per_cu = 0;
    check_hip(L.feat == F_ALL   ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_render_paths<F_ALL>, ABLK, L.shmem)
              : L.feat == F_TEX ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_render_paths<F_TEX>, ABLK, L.shmem)
                                : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_render_paths<0u>, ABLK, L.shmem),
              "occupancy");
    L.grid = std::max(1, per_cu) * device_cu_count();
    return L;
}
// Samples [a.slot0, a.slot1) into out[slot - a.rad_off].
void launch_paths(const Scene& sc, const PathsShape& L, const AdjArgs& a, float4* out) {
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, sc.stream, a.claim);
    if (L.feat == F_ALL) hipLaunchKernelGGL(k_render_paths<F_ALL>, dim3(L.grid), dim3(ABLK), L.shmem, sc.stream, a, out);
    else if (L.feat == F_TEX) hipLaunchKernelGGL(k_render_paths<F_TEX>, dim3(L.grid), dim3(ABLK), L.shmem, sc.stream, a, out);
    else hipLaunchKernelGGL(k_render_paths<0u>, dim3(L.grid), dim3(ABLK), L.shmem, sc.stream, a, out);
    check_hip(hipGetLastError(), "k_render_paths launch");
}

constexpr long long PATH_CHUNK = 16ll << 20;  // launches of at most 16 M samples

}  // namespace

void render_paths(Scene& sc, const CamView& cam, const RenderParams& p, uint32_t sample0, int P, float4* radiance,
                  int* err, HitRec* vrec, int vrec_cap) {
    if (p.env_is) throw std::runtime_error("the path kernel does not sample the sky (env_is)");
    if (!sc.target) sc.target = new RenderTarget();
    sc.target->path_claim.ensure(sizeof(int));
    AdjArgs a{sc.view, cam, p.W, p.H, p.W * p.H, P, p.seed, sample0, p.prr, p.mis, p.maxdepth, p.rrdepth, nullptr,
              0, P, nullptr, 0.0f, 0, nullptr, err, -1, nullptr, p.integrator == 1 ? 0 : 1, 0,
              sc.target->path_claim.as<unsigned>(), vrec, vrec ? vrec_cap : 0};
    const PathsShape L = paths_shape(sc);
    for (long long s0 = 0; s0 < P; s0 += PATH_CHUNK) {
        a.slot0 = int(s0);
        a.slot1 = int(std::min<long long>(P, s0 + PATH_CHUNK));
        launch_paths(sc, L, a, radiance);
    }
}

void adjoint_ptnee(Scene& sc, const CamView& cam, const RenderParams& p, const float* dldi_host, double* grad_host,
                   int tex_id, double* gtex_host) {
    if (!sc.built) throw std::runtime_error("scene not built");
    if (p.W <= 0 || p.H <= 0 || p.spp <= 0) throw std::runtime_error("bad image size or spp");
    if (p.env_is) throw std::runtime_error("the adjoint does not replay environment sampling");
    if (!sc.target) sc.target = new RenderTarget();
    RenderTarget& t = *sc.target;
    // the stored per-sample radiance serves when it belongs to exactly these samples
    const bool stored = t.adj_matches(cam, p, sc.generation);
    const size_t n3 = size_t(p.W) * p.H * 3;
    if (dldi_host) {
        t.adj_dldi.ensure(n3 * sizeof(float));
        t.adj_dldi.upload(dldi_host, n3 * sizeof(float), sc.stream);
    } else if (!t.adj_dldi.ptr || t.W != p.W || t.H != p.H) {
        throw std::runtime_error("no dL/dimage: pass one or call nr_loss_grad_l2");
    }
    const int n_params = int(sc.mats.size() / 8) * 4;
    t.adj_grad.ensure(sizeof(double) * n_params);
    hipLaunchKernelGGL(k_zero, dim3(64), dim3(256), 0, sc.stream, t.adj_grad.as<double>(), n_params);
    double* gtex = nullptr;
    size_t ntex = 0;
    if (tex_id >= 0 && gtex_host) {
        if (tex_id >= int(sc.textures.size()) || sc.textures[tex_id].w == 0)
            throw std::runtime_error("texel gradient of a texture that does not exist");
        ntex = size_t(sc.textures[tex_id].w) * sc.textures[tex_id].h * 3;
        t.adj_gtex.ensure(sizeof(double) * ntex);
        hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, sc.stream, t.adj_gtex.as<double>(), int(ntex));
        gtex = t.adj_gtex.as<double>();
    }
    t.adj_err.ensure(sizeof(int));
    int* const d_err = t.adj_err.as<int>();
    t.path_claim.ensure(sizeof(int));
    const int zero = 0;
    check_hip(hipMemcpyAsync(d_err, &zero, sizeof(int), hipMemcpyHostToDevice, sc.stream), "H2D");
    const long long P = (long long)p.W * p.H * p.spp;
    if (P > 0x7FFFFFFFll) throw std::runtime_error("adjoint: more than 2^31 samples");
    AdjArgs a{sc.view, cam, p.W, p.H, p.W * p.H, int(P), p.seed, uint32_t(p.spp_offset), p.prr, p.mis, p.maxdepth,
              p.rrdepth, stored ? t.radiance.as<float4>() : nullptr, 0, int(P), t.adj_dldi.as<float>(),
              1.0f / float(p.spp), n_params, t.adj_grad.as<double>(), d_err, tex_id, gtex, p.integrator == 1 ? 0 : 1,
              0, t.path_claim.as<unsigned>(), nullptr, 0};
    // the replay reads the vertex records of a one-lane-per-path render of these very samples
    if (stored && t.adj_vrec_valid) {
        a.vrec = t.adj_vrec.as<HitRec>();
        a.vrec_cap = VREC_CAP;
    }
    // Chunks of a whole number of samples per pixel, about 16 M samples (at least one per pixel).
    // Without the render's own per-sample radiance each chunk is first rendered forward, by the
    // form the render modes would pick: one lane per path into a chunk buffer when the scene has a
    // flat leaf list or p.path_kernel asks for it (1, or 2 below AUTO_PATHS_WORK per chunk), else
    // the wavefront into the render target's radiance buffer (the accumulated image is left alone).
    // Nothing per path outlives its chunk.
    const int npix = p.W * p.H;
    const long long chunk_spp = std::max<long long>(1, PATH_CHUNK / npix);
    const size_t chunk_max = size_t(std::min<long long>(P, chunk_spp * npix));
    const bool fwd_paths = sc.view.n_flat > 0 || p.path_kernel == 1 ||
                           (p.path_kernel == 2 && double(chunk_max) * sc.n_tris() <= AUTO_PATHS_WORK);
    PathsShape fwd{};
    // the forward pass of a chunk also writes vertex records when they fit, and the replay reads them
    const bool fwd_rec = chunk_max * VREC_CAP * sizeof(HitRec) <= VREC_BUDGET;
    if (!stored && fwd_paths) {
        fwd = paths_shape(sc);
        t.adj_rad.ensure(sizeof(float4) * chunk_max);
        if (fwd_rec) t.adj_vrec.ensure(chunk_max * VREC_CAP * sizeof(HitRec));
        t.adj_vrec_valid = false;  // the chunk's records replace the last render's
    }
    const size_t shmem = size_t(sc.view.lane_slots) * ABLK * sizeof(int) +
                         (n_params <= LDS_PARAMS ? size_t(n_params) * sizeof(float) : 0);
    if (shmem > device_max_lds()) throw std::runtime_error("adjoint LDS budget exceeded (BVH too deep)");
    // kernel instance for the material features present (as the shade kernels)
    const unsigned feat = sc.features & (F_GGX | F_EXT) ? F_ALL : (sc.features & F_TEX ? F_TEX : 0u);
    int per_cu = 0;
    check_hip(feat == F_ALL   ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_adjoint<F_ALL>, ABLK, shmem)
              : feat == F_TEX ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_adjoint<F_TEX>, ABLK, shmem)
                              : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_adjoint<0u>, ABLK, shmem),
              "occupancy");
    const int grid = std::max(1, per_cu) * device_cu_count();
    for (long long b = 0; b < p.spp; b += chunk_spp) {
        a.slot0 = int(b * npix);
        a.slot1 = int(std::min(P, (b + chunk_spp) * npix));
        if (!stored && fwd_paths) {
            a.radiance = t.adj_rad.as<float4>();
            a.rad_off = a.slot0;
            a.vrec = fwd_rec ? t.adj_vrec.as<HitRec>() : nullptr;
            a.vrec_cap = VREC_CAP;
            launch_paths(sc, fwd, a, t.adj_rad.as<float4>());
        } else if (!stored) {
            const bool rec = wavefront_radiance(sc, cam, p, uint32_t(p.spp_offset + b), a.slot1 - a.slot0, fwd_rec);
            a.radiance = t.radiance.as<float4>();
            a.rad_off = a.slot0;
            a.vrec = rec ? t.adj_vrec.as<HitRec>() : nullptr;
            a.vrec_cap = VREC_CAP;
        }
        hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, sc.stream, a.claim);
        if (feat == F_ALL) hipLaunchKernelGGL(k_adjoint<F_ALL>, dim3(grid), dim3(ABLK), shmem, sc.stream, a);
        else if (feat == F_TEX) hipLaunchKernelGGL(k_adjoint<F_TEX>, dim3(grid), dim3(ABLK), shmem, sc.stream, a);
        else hipLaunchKernelGGL(k_adjoint<0u>, dim3(grid), dim3(ABLK), shmem, sc.stream, a);
        check_hip(hipGetLastError(), "k_adjoint launch");
    }
    int err = 0;
    check_hip(hipMemcpyAsync(&err, d_err, sizeof(int), hipMemcpyDeviceToHost, sc.stream), "D2H");
    check_hip(hipMemcpyAsync(grad_host, t.adj_grad.ptr, sizeof(double) * n_params, hipMemcpyDeviceToHost, sc.stream),
              "D2H");
    if (gtex)
        check_hip(hipMemcpyAsync(gtex_host, gtex, sizeof(double) * ntex, hipMemcpyDeviceToHost, sc.stream), "D2H");
    check_hip(hipStreamSynchronize(sc.stream), "adjoint sync");
    if (err) throw std::runtime_error("traversal stack overflow in the adjoint");
}

}  // namespace nr
