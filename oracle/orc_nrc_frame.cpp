// ORACLE -- test infrastructure only (see orc_common.h).
//
// One radiance-cache frame, restated path by path in FP32 (forward form): the data path of the
// device's cache renderer.  It follows the reference's radiance-cache renderer
// (renderernrc.cpp:13-281: camera paths that stop at a cache query, training paths whose vertices
// become training records with target (L_path - L_before) / beta, renderernrc.cpp:128-133,
// 175-182) with the path-spread termination rule of the radiance-cache method, on top of the ptnee
// vertex of rendererptls.cpp:13-93 (NEE with MIS, Russian roulette, BSDF continuation).
//
//   * slots [0, P): sample s / npix of pixel s % npix; slots [P, P + T): training path i = s - P from
//     pixel philox((i, sample0, 0x6E7263, 0), (seed, 0x74726E67)).x % npix, sample index
//     0x80000000 | (sample0 T + i); every draw comes from the Philox stream of (seed, pixel, sample).
//   * spread at a surface vertex (cos = |n . d|, t the hit distance): at the first vertex
//     a = sqrt(c t^2 / (4 pi max(cos, 1e-6))) and s = 0; later s += sqrt(t^2 / max(p cos, 1e-20)), p
//     the pdf of the direction that led there.  A camera path queries the cache (and stops: no NEE,
//     no continuation) at depth >= min_depth when s > a; a training path queries at depth + 1 >=
//     maxlen and records every earlier surface vertex with the radiance collected so far.
//   * query / record features: hit point, roughness (0 for Lambert), spherical angles of wo and n
//     ((acos y / pi, atan2(z, x) / 2 pi + 1/2)), albedo, slot, throughput beta on arrival.
//   * the cache's answer beta * max(0, y) is added to the path's radiance; y is the FP64 network
//     evaluated at the device's BF16 rounding points (orc_nrc.cpp).
#include <omp.h>
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include "orc_handle.h"

extern "C" void orc_nrc_forward(const float* w, const float* q, int n, int bf16, double* y);

namespace orc {
namespace {

struct PathOut {
    V3f L;                     // radiance collected before any cache answer
    bool queried = false;
    int qdepth = -1;           // vertex index of the query
    float q[16];               // the query that ended the path
    std::vector<float> recs;   // training records: 16 query floats + 3 radiance-before each
};

void spherical(float x, float y, float z, float& th, float& ph) {
    const float yc = std::fmin(std::fmax(y, -1.0f), 1.0f);
    th = std::acos(yc) * 0.31830988618f;
    ph = std::atan2(z, x) * 0.15915494309f + 0.5f;
}

struct FrameOpts {
    int W, H, P, T, maxlen, use_cache, min_depth;
    uint32_t seed, sample0;
    float c_spread;
    RenderOpts ro;
};

PathOut nrc_path(const Tracer<float>& tr, const Camera& cam, const FrameOpts& fo, int slot) {
    const Scene& sc = tr.sc;
    const int npix = fo.W * fo.H;
    const bool train = slot >= fo.P;
    int pixel;
    uint32_t sample;
    if (!train) {
        pixel = slot % npix;
        sample = fo.sample0 + uint32_t(slot / npix);
    } else {
        const uint32_t i = uint32_t(slot - fo.P);
        sample = 0x80000000u | (fo.sample0 * uint32_t(fo.T) + i);
        const uint32_t ctr[4] = {i, fo.sample0, 0x6E7263u, 0u}, key[2] = {fo.seed, 0x74726E67u};
        uint32_t out[4];
        philox4x32_10(ctr, key, out);
        pixel = int(out[0] % uint32_t(npix));
    }
    const Stream rng(fo.seed, uint32_t(pixel), sample);
    float u0[4];
    rng.draw(0, 0, u0);
    V3f o = cam.pos;
    V3f d = cam.ray(float(pixel % fo.W) + u0[0], float(pixel / fo.W) + u0[1], float(fo.W), float(fo.H));
    V3f beta(1.0f);
    float fac = 1.0f, s_sum = 0.0f, s_lim = 0.0f, p_prev = 1.0f;
    PathOut r;
    r.L = V3f(0.0f);
    for (int depth = 0;; ++depth) {
        float t, b1, b2;
        int ti;
        if (!tr.hit(o, d, t, b1, b2, ti)) {
            r.L += beta * tr.env(d);
            return r;
        }
        const Tri& tri = sc.tris[ti];
        const Material& M = sc.mats[tri.mat];
        MatEval<float> mat{&sc, &M};
        if (M.type == MAT_LIGHT) {
            r.L += beta * (V3f(M.kd[0], M.kd[1], M.kd[2]) * fac);
            return r;
        }
        const V3f n = tri.normal<float>(b1, b2);
        if (n.dot(d) > 0 && !mat.transmissive()) return r;
        const V3f uv = tri.texc<float>(b1, b2);
        const V3f hp = o + d * t, wo = -d;
        // spread and the cache's decision
        const float cs = std::fabs(n.dot(d)), t2 = t * t;
        if (depth == 0) {
            s_sum = 0.0f;
            s_lim = std::sqrt(fo.c_spread * t2 / (4.0f * 3.14159265f * std::fmax(cs, 1e-6f)));
        } else {
            s_sum += std::sqrt(t2 / std::fmax(p_prev * cs, 1e-20f));
        }
        const bool query = train ? depth + 1 >= fo.maxlen : (fo.use_cache && depth >= fo.min_depth && s_sum > s_lim);
        if (query || train) {
            float th, ph, nth, nph;
            spherical(wo.x, wo.y, wo.z, th, ph);
            spherical(n.x, n.y, n.z, nth, nph);
            const V3f alb = mat.albedo(uv);
            int slot_bits = slot;
            float slot_f;
            std::memcpy(&slot_f, &slot_bits, 4);
            const float vals[16] = {hp.x, hp.y, hp.z, M.type == MAT_LAMBERT ? 0.0f : M.alpha, th, ph, nth, nph,
                                    alb.x, alb.y, alb.z, slot_f, beta.x, beta.y, beta.z, 0.0f};
            if (query) {
                std::memcpy(r.q, vals, sizeof(vals));
                r.queried = true;
                r.qdepth = depth;
                return r;
            }
            r.recs.insert(r.recs.end(), vals, vals + 16);
            r.recs.push_back(r.L.x);
            r.recs.push_back(r.L.y);
            r.recs.push_back(r.L.z);
        }
        float ua[4];
        rng.draw(uint32_t(1 + depth), 0, ua);
        float fac_next = 1.0f;
        if (fo.ro.nee && mat.needs_nee()) {
            const int li = sc.pick_light(ua[0]);
            if (li >= 0) {
                const Tri& L = sc.tris[li];
                float r1 = ua[1], r2 = ua[2];
                if (r1 + r2 > 1.0f) { r1 = 1.0f - r1; r2 = 1.0f - r2; }
                const V3f lp = (1.0f - r1 - r2) * L.p[0] + r1 * L.p[1] + r2 * L.p[2];
                const V3f ln = L.normal<float>(r1, r2);
                const V3f lv = lp - hp, wl = lv.normalized();
                const Material& LM = sc.mats[L.mat];
                const V3f ke(LM.kd[0], LM.kd[1], LM.kd[2]);
                const float lpdf = sc.light_p();
                const float lcos = ln.dot(-wl), ncos = n.dot((lp - hp).normalized());
                const float lop = lpdf * lv.dot(lv) / (1e-6f < lcos ? lcos : 1e-6f);
                float mis = lop / (lop + mat.pdf(wo, n, wl));
                if (fo.ro.mis == 0) mis = 0.5f;
                fac_next = 1.0f - mis;
                const V3f Ll = ke / lv.dot(lv) * ((0.0f < lcos ? lcos : 0.0f) / lpdf);
                const V3f c = beta * Ll * mat.f(wo, n, wl, uv) * ((0.0f < ncos ? ncos : 0.0f) * mis);
                if (c.x != 0.0f || c.y != 0.0f || c.z != 0.0f) {
                    float st, sb1, sb2;
                    int sti;
                    tr.hit(hp + wl * 1e-3f, wl, st, sb1, sb2, sti);
                    if (sti < 0 || double(st) + 2e-3 > double(lv.norm())) r.L += c;
                }
            }
        }
        if (fo.ro.maxdepth > 0 && depth + 1 >= fo.ro.maxdepth) return r;
        float prr = 1.0f;
        if (depth >= fo.ro.rrdepth) {
            if (ua[3] > fo.ro.prr) return r;
            prr = fo.ro.prr;
        }
        float v[4];
        rng.draw(uint32_t(1 + depth), 1, v);
        const V3f wi = mat.sample(v, wo, n);
        const float pdf = mat.pdf(wo, n, wi);
        const V3f f = mat.f(wo, n, wi, uv);
        const float c = std::fabs(wi.dot(n));
        const V3f nb(beta.x * c * f.x / pdf / prr, beta.y * c * f.y / pdf / prr, beta.z * c * f.z / pdf / prr);
        if (nb.x == 0.0f && nb.y == 0.0f && nb.z == 0.0f) return r;
        o = hp + wi * 1e-3f;
        d = wi;
        beta = nb;
        fac = fac_next;
        p_prev = pdf;
    }
}

}  // namespace
}  // namespace orc

extern "C" {

// One cache frame of W x H x spp camera paths and T training paths (the device's render_nrc with
// the network w answering every query).  sum / sum2 (3 W H doubles, added to): per-pixel sums of
// max(0, L) and its square.  recs (rec_cap x 16 floats), lb and tgt (rec_cap x 3 floats each): the
// training records in slot and depth order, the radiance their path had collected before them,
// and their targets; *n_recs (all of them, also past rec_cap),
// *n_queries: the queries made (camera paths and training-path tails); qdepth (optional, W H spp
// ints): the vertex index at which each camera path queried the cache, -1 if it did not.
int orc_nrc_frame(void* h, const float* cam, int W, int H, int spp, uint32_t sample0, uint32_t seed, float prr,
                  int mis, int maxdepth, int rrdepth, int T, int maxlen, float c_spread, int use_cache, int min_depth,
                  const float* w, int nthreads, double* sum, double* sum2, float* recs, float* lb, float* tgt,
                  int rec_cap, int* n_recs, int* n_queries, int* qdepth) {
    using namespace orc;
    OScene* o = static_cast<OScene*>(h);
    if (!o->built) return -1;
    Camera c;
    c.pos = V3f(cam[0], cam[1], cam[2]);
    c.from_euler(cam[3], cam[4], cam[5]);
    c.fov = cam[6];
    c.aspect = cam[7];
    FrameOpts fo{W, H, spp * W * H, T, maxlen, use_cache, std::max(1, min_depth), seed, sample0, c_spread, {}};
    fo.ro.prr = prr; fo.ro.mis = mis; fo.ro.maxdepth = maxdepth; fo.ro.rrdepth = rrdepth; fo.ro.seed = seed;
    const int S = fo.P + T;
    std::vector<PathOut> paths(S);
    if (nthreads > 0) omp_set_num_threads(nthreads);
    const Tracer<float> tr{o->sc, o->bvh, fo.ro};
#pragma omp parallel for schedule(dynamic, 64)
    for (int s = 0; s < S; ++s) paths[s] = nrc_path(tr, c, fo, s);
    // the cache answers every query at once
    std::vector<float> q;
    std::vector<int> qslot;
    for (int s = 0; s < S; ++s)
        if (paths[s].queried) {
            q.insert(q.end(), paths[s].q, paths[s].q + 16);
            qslot.push_back(s);
        }
    std::vector<double> y(3 * qslot.size());
    if (!qslot.empty()) orc_nrc_forward(w, q.data(), int(qslot.size()), 1, y.data());
    for (size_t k = 0; k < qslot.size(); ++k) {
        PathOut& p = paths[qslot[k]];
        const float* qd = p.q + 12;
        p.L.x += qd[0] * std::fmax(float(y[3 * k]), 0.0f);
        p.L.y += qd[1] * std::fmax(float(y[3 * k + 1]), 0.0f);
        p.L.z += qd[2] * std::fmax(float(y[3 * k + 2]), 0.0f);
    }
    *n_queries = int(qslot.size());
    if (qdepth)
        for (int s = 0; s < fo.P; ++s) qdepth[s] = paths[s].qdepth;
    const int npix = W * H;
    for (int s = 0; s < fo.P; ++s) {  // per pixel in sample order, as the device's finalize
        const V3f& L = paths[s].L;
        const double v[3] = {0.0f < L.x ? L.x : 0.0f, 0.0f < L.y ? L.y : 0.0f, 0.0f < L.z ? L.z : 0.0f};
        for (int ch = 0; ch < 3; ++ch) {
            sum[3 * size_t(s % npix) + ch] += v[ch];
            sum2[3 * size_t(s % npix) + ch] += v[ch] * v[ch];
        }
    }
    int nr = 0;
    for (int s = fo.P; s < S; ++s) {
        const PathOut& p = paths[s];
        for (size_t k = 0; k + 19 <= p.recs.size(); k += 19, ++nr) {
            if (nr >= rec_cap) continue;
            const float* r = &p.recs[k];
            std::memcpy(recs + size_t(nr) * 16, r, 16 * sizeof(float));
            const float Lp[3] = {p.L.x, p.L.y, p.L.z};
            for (int ch = 0; ch < 3; ++ch) {
                const float b = r[12 + ch];
                lb[size_t(nr) * 3 + ch] = r[16 + ch];
                tgt[size_t(nr) * 3 + ch] = b > 1e-8f ? (Lp[ch] - r[16 + ch]) / b : 0.0f;
            }
        }
    }
    *n_recs = nr;
    return 0;
}

}  // extern "C"
