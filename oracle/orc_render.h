// ORACLE -- test infrastructure only (see orc_common.h).
//
// Camera and the ptnee estimator restated from src/scene/camera.cpp and
// src/renderer/rendererptls.cpp, templated on the arithmetic type: T = float is the
// reference-semantics FP32 mode, T = double is the FP64 oracle of the same estimator.
#pragma once
#include "orc_bvh.h"

namespace orc {

// src/scene/camera.cpp:8-32 and :88-105.
struct Camera {
    V3f pos, gaze, up;
    float fov = 90.0f, aspect = 1.0f;
    void from_euler(float yaw, float pitch, float roll) {
        float e[3] = {yaw, pitch, roll};
        float c[3], s[3];
        for (int i = 0; i < 3; ++i) {
            c[i] = float(std::cos(double(e[i] / 180) * 3.14159));
            s[i] = float(std::sin(double(e[i] / 180) * 3.14159));
        }
        float c1 = c[0], s1 = s[0], c2 = c[1], s2 = s[1], c3 = c[2], s3 = s[2];
        float m12 = c3 * s1 * s2 - c1 * s3;
        float m13 = c2 * s1;
        float m22 = c2 * c3;
        float m23 = -s2;
        float m32 = s1 * s3 + c1 * c3 * s2;
        float m33 = c1 * c2;
        up = {m12, m22, m32};
        gaze = {-m13, -m23, -m33};
    }
    V3f ray(float ix, float iy, float W, float H) const {
        float cx = ix / W * 2 - 1;
        float cy = iy / H * 2 - 1;
        V3f hand = gaze.cross(up).normalized();
        float fh = float(2 * std::tan(double(fov) * 3.14159 / 180 / 2));
        float fw = fh * aspect;
        return (gaze + hand * fw * cx * 0.5f - up * fh * cy * 0.5f).normalized();
    }
};

struct RenderOpts {
    float prr = 0.8f;   // rendererptls.cpp:81
    int mis = 1;        // rendererptls.cpp:65
    int maxdepth = 0;   // 0 = unbounded (reference); else vertices with index >= maxdepth are not created
    int rrdepth = 0;    // Russian roulette applies at vertices with index >= rrdepth (reference: 0)
    uint32_t seed = 0;
    int nee = 1;        // 0: RendererPT (rendererpt.cpp:13-61): no light sampling, emitters at full weight
    int env_is = 0;     // 1: sky sampling at Lambert / GGX-reflect vertices (same expected image)
};

// Counters of the instrumented traversal: node-pair fetches and triangle tests.
struct TravStats {
    uint64_t rays = 0, pair_fetch = 0, tri_tests = 0, shade = 0;
};

template <class T>
struct Tracer {
    const Scene& sc;
    const BVH& bvh;
    RenderOpts opt;

    bool hit(const V3<T>& o, const V3<T>& d, T& t, T& b1, T& b2, int& tri) const;

    V3<T> env(const V3<T>& d) const {  // rendererptls.cpp:18-28
        if (!sc.has_env) return V3<T>(0);
        T v = T(1) - T(std::acos(d.y)) / T(3.14159);
        T u = T(std::atan2(d.z, d.x)) / T(2) / T(3.14159) + T(0.5);
        return V3<T>(sc.env.uv(float(u), float(v))) * T(sc.env_gain);
    }

    // RendererPTLS::trace, rendererptls.cpp:13-93 (recursive, one vertex per call).  skyw is the
    // balance-heuristic weight of a BSDF-sampled escape when the previous vertex also sampled the
    // sky (1 otherwise).
    V3<T> trace(const V3<T>& o, const V3<T>& d, T fac, int depth, const Stream& rng, T skyw = T(1)) const {
        T t, b1, b2;
        int ti;
        if (!hit(o, d, t, b1, b2, ti)) return env(d) * skyw;
        const Tri& tri = sc.tris[ti];
        const Material& M = sc.mats[tri.mat];
        MatEval<T> mat{&sc, &M};
        V3<T> wo = -d;
        V3<T> n = tri.normal<T>(b1, b2);
        if (M.type == MAT_LIGHT) return V3<T>(T(M.kd[0]), T(M.kd[1]), T(M.kd[2])) * fac;
        V3<T> uv = tri.texc<T>(b1, b2);
        if (n.dot(d) > 0 && !mat.transmissive()) return V3<T>(0);
        V3<T> hp = o + d * t;
        V3<T> result(0);
        T fac_next = T(1);
        float u[4];
        rng.draw(uint32_t(1 + depth), 0, u);
        if (opt.nee && mat.needs_nee()) {
            int li = sc.pick_light(u[0]);
            if (li >= 0) {
                const Tri& L = sc.tris[li];
                T r1 = T(u[1]), r2 = T(u[2]);
                if (r1 + r2 > 1) { r1 = T(1) - r1; r2 = T(1) - r2; }
                V3<T> lp = (T(1) - r1 - r2) * V3<T>(L.p[0]) + r1 * V3<T>(L.p[1]) + r2 * V3<T>(L.p[2]);
                V3<T> ln = L.normal<T>(r1, r2);
                V3<T> lv = lp - hp;
                V3<T> wl = lv.normalized();
                const Material& LM = sc.mats[L.mat];
                V3<T> le(T(LM.kd[0]), T(LM.kd[1]), T(LM.kd[2]));
                T lpdf = T(sc.light_p());
                T lop = lpdf * lv.dot(lv) / std::max(T(1e-6f), ln.dot(-wl));
                T bop = mat.pdf(wo, n, wl);
                T mis = lop / (lop + bop);
                if (opt.mis == 0) mis = T(0.5f);
                fac_next = T(1) - mis;
                T st, sb1, sb2;
                int sti;
                hit(hp + wl * T(1e-3), wl, st, sb1, sb2, sti);
                if (double(st) + 2e-3 > double(lv.norm())) {
                    V3<T> f = mat.f(wo, n, wl, uv);
                    V3<T> Ll = le / lv.norm2() * std::max(T(0), ln.dot(-wl)) / lpdf;
                    result += Ll * f * std::max(T(0), n.dot((lp - hp).normalized())) * mis;
                }
            }
        }
        // Sky sampling: the reference's sky term at this vertex is the integral of g L V with
        // g = q / p_rep * f |cos| (q the sampler's true density, p_rep the reported pdf); the sky
        // sample adds g L V / (p_env + q), the BSDF escape is weighted by q / (q + p_env).
        const bool can_next = !(opt.maxdepth > 0 && depth + 1 >= opt.maxdepth);
        const bool env_vertex = opt.env_is && opt.nee && sc.env_sampler.ready() && can_next &&
                                (M.type == MAT_LAMBERT || M.type == MAT_GGXREFL);
        if (env_vertex) {
            float ue[4];
            rng.draw(uint32_t(1 + depth), 2, ue);
            V3<T> we;
            T pe = 0;
            if (sc.env_sampler.sample(ue, we, pe)) {
                const T qe = mat.sample_density(wo, n, we);
                if (qe > 0) {
                    T st, sb1, sb2;
                    int sti;
                    if (!hit(hp + we * T(1e-3), we, st, sb1, sb2, sti)) {
                        const V3<T> f = mat.f(wo, n, we, uv);
                        const T pr = mat.pdf(wo, n, we);
                        result += f * env(we) * (qe / pr * std::fabs(we.dot(n)) / (pe + qe));
                    }
                }
            }
        }
        T prr = T(1);
        if (depth >= opt.rrdepth) {
            if (u[3] > opt.prr) return result;
            prr = T(opt.prr);
        }
        if (!can_next) return result;
        float v[4];
        rng.draw(uint32_t(1 + depth), 1, v);
        V3<T> wi = mat.sample(v, wo, n);
        T pdf = mat.pdf(wo, n, wi);
        V3<T> f = mat.f(wo, n, wi, uv);
        T sw = T(1);
        if (env_vertex) {
            const T qi = mat.sample_density(wo, n, wi), pi_env = sc.env_sampler.density(wi);
            sw = qi + pi_env > 0 ? qi / (qi + pi_env) : T(1);
        }
        V3<T> Li = trace(hp + wi * T(1e-3), wi, fac_next, depth + 1, rng, sw);
        result += Li * std::fabs(wi.dot(n)) * f / pdf / prr;
        return result;
    }

    // One camera sample: rendererptls.cpp:170-174, max(0, .) per component (NaN -> 0).
    V3<T> sample(const Camera& cam, int x, int y, int W, int H, uint32_t s) const {
        Stream rng(opt.seed, uint32_t(y) * uint32_t(W) + uint32_t(x), s);
        float u[4];
        rng.draw(0, 0, u);
        V3f dir = cam.ray(float(x) + u[0], float(y) + u[1], float(W), float(H));
        V3<T> L = trace(V3<T>(cam.pos), V3<T>(dir), T(1), 0, rng);
        return {T(0) < L.x ? L.x : T(0), T(0) < L.y ? L.y : T(0), T(0) < L.z ? L.z : T(0)};
    }
};

template <>
inline bool Tracer<float>::hit(const V3f& o, const V3f& d, float& t, float& b1, float& b2, int& tri) const {
    Hit h = bvh.intersect(o, d);
    t = h.t; b1 = h.b1; b2 = h.b2; tri = h.tri;
    return h.tri >= 0;
}
template <>
inline bool Tracer<double>::hit(const V3d& o, const V3d& d, double& t, double& b1, double& b2, int& tri) const {
    return bvh.intersect_d(o, d, t, b1, b2, tri);
}

}  // namespace orc
