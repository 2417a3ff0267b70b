// ORACLE -- test infrastructure only (see orc_common.h).
//
// Scene records, textures, materials and the area-light sampler of the reference, restated.
#pragma once
#include "orc_common.h"

namespace orc {

enum MatType { MAT_LAMBERT = 0, MAT_LIGHT = 1, MAT_GGXREFL = 2, MAT_GGX = 3, MAT_BLINNPHONG = 4, MAT_GLASS = 5 };

// One material row as passed through the C API: type, rgb (Kd / F0 / Ke), alpha, ior, texture.
struct Material {
    int type = MAT_LAMBERT;
    float kd[3] = {0, 0, 0};   // Kd / Ke / F0
    float alpha = 0.5f;        // GGX roughness; Blinn-Phong exponent Ns
    float ior = 1.0f;
    int tex = -1;              // Kd / F0 texture
    float ks[3] = {0, 0, 0};   // Blinn-Phong Ks
    int tex_ks = -1;           // Blinn-Phong Ks texture
};

// src/scene/triangle.h:8-36, triangle.cpp:5-25.  A flat normal normalize((p1-p0)x(p2-p0))
// replaces all three vertex normals when |n0|^2 < 1e-4 (:14-17).
struct Tri {
    V3f p[3], t[3], n[3];
    int mat = 0;
    void fix_normal() {
        if (double(n[0].norm2()) < 1e-4) {
            n[0] = (p[1] - p[0]).cross(p[2] - p[0]).normalized();
            n[1] = n[0];
            n[2] = n[0];
        }
    }
    V3f pmin() const { return vmin(p[0], vmin(p[1], p[2])); }  // triangle.cpp:68-71
    V3f pmax() const { return vmax(p[0], vmax(p[1], p[2])); }  // triangle.cpp:73-76
    // triangle.cpp:57-60: |e1 x e2|, twice the geometric area.
    float area() const { return (p[1] - p[0]).cross(p[2] - p[0]).norm(); }
    template <class T>
    V3<T> normal(T b1, T b2) const {  // triangle.cpp:35-38
        return ((T(1) - b1 - b2) * V3<T>(n[0]) + b1 * V3<T>(n[1]) + b2 * V3<T>(n[2])).normalized();
    }
    template <class T>
    V3<T> texc(T b1, T b2) const {  // triangle.cpp:40-43
        return (T(1) - b1 - b2) * V3<T>(t[0]) + b1 * V3<T>(t[1]) + b2 * V3<T>(t[2]);
    }
};

// src/scene/texture.cpp:54-72: texels outside [0,W)x[0,H) read as black; bilinear weights between
// the floor texel and floor+1 (no wrap of the +1 neighbour); pixelUV wraps u and 1-v into [0,1).
struct Texture {
    int w = 0, h = 0;
    std::vector<float> rgb;  // linear RGB, row-major, y down
    V3f px(int x, int y) const {
        if (x < 0 || y < 0 || x >= w || y >= h) return V3f(0.0f);
        const float* q = &rgb[3 * (size_t(y) * w + x)];
        return {q[0], q[1], q[2]};
    }
    V3f bilinear(float x, float y) const {
        int x0 = int(std::floor(x)), y0 = int(std::floor(y));
        int x1 = x0 + 1, y1 = y0 + 1;
        return px(x1, y1) * (x - x0) * (y - y0) + px(x1, y0) * (x - x0) * (y1 - y) +
               px(x0, y1) * (x1 - x) * (y - y0) + px(x0, y0) * (x1 - x) * (y1 - y);
    }
    V3f uv(float u, float v) const {
        float fu = std::fmod(std::fmod(u, 1.0f) + 1.0f, 1.0f);
        float fv = std::fmod(std::fmod(1.0f - v, 1.0f) + 1.0f, 1.0f);
        return bilinear(fu * w, fv * h);
    }
};
// texture.cpp:18-20: QColor::redF() is c/255 in double, raised to the float literal 2.2f.
inline float srgb8_decode(int c) { return float(std::pow(double(c) / 255.0, double(2.2f))); }

struct Light {  // one emissive triangle (lightsampler.cpp:13-22)
    int tri;
};

// Sky sampler of the optional environment sampling (not in the reference: a variance reduction of
// its estimator that keeps its expected image).  The lookup reads the continuous texel coordinate
// x = u W, y = (1 - v) H, u = atan2(z, x) / (2 pi') + 1/2, 1 - v = acos(y) / pi' (pi' = 3.14159),
// blending texels floor(x) and floor(x) + 1 (black past the last one).  Cell (i, j) =
// [j, j+1) x [i, i+1) is bounded by its four corner texels; its weight is the largest corner
// channel times |gain| times sin(theta) at the cell centre.  Cells are drawn with Vose's alias
// method built in double (small and large lists as stacks in index order), then a point uniform in
// the cell; the solid-angle density is pdf_cell W H / (2 pi'^2 sin theta).
struct EnvSampler {
    int w = 0, h = 0;
    std::vector<float> prob, pdf;
    std::vector<int> alias;
    float norm = 0.0f;

    void build(const Texture& t, float gain) {
        w = h = 0;
        prob.clear(); pdf.clear(); alias.clear();
        if (t.w <= 0 || t.h <= 0) return;
        const size_t N = size_t(t.w) * t.h;
        auto corner = [&](int x, int y) -> float {
            if (x >= t.w || y >= t.h) return 0.0f;
            const float* q = &t.rgb[3 * (size_t(y) * t.w + x)];
            return std::max(std::fabs(q[0]), std::max(std::fabs(q[1]), std::fabs(q[2])));
        };
        std::vector<double> wt(N);
        double sum = 0;
        for (int i = 0; i < t.h; ++i) {
            const double st = std::sin((i + 0.5) / t.h * double(3.14159f));
            for (int j = 0; j < t.w; ++j) {
                const float c = std::max(std::max(corner(j, i), corner(j + 1, i)),
                                         std::max(corner(j, i + 1), corner(j + 1, i + 1)));
                wt[size_t(i) * t.w + j] = double(c) * std::fabs(double(gain)) * st;
                sum += wt[size_t(i) * t.w + j];
            }
        }
        if (!(sum > 0) || !std::isfinite(sum)) return;
        std::vector<double> scaled(N);
        prob.assign(N, 1.0f);
        alias.resize(N);
        pdf.resize(N);
        std::vector<int> lo, hi;
        for (size_t k = 0; k < N; ++k) {
            scaled[k] = wt[k] / sum * double(N);
            alias[k] = int(k);
            pdf[k] = float(wt[k] / sum);
            if (scaled[k] < 1.0) lo.push_back(int(k));
            else hi.push_back(int(k));
        }
        while (!lo.empty() && !hi.empty()) {
            const int a = lo.back(), b = hi.back();
            lo.pop_back();
            hi.pop_back();
            prob[a] = float(scaled[a]);
            alias[a] = b;
            scaled[b] = (scaled[b] + scaled[a]) - 1.0;
            if (scaled[b] < 1.0) lo.push_back(b);
            else hi.push_back(b);
        }
        w = t.w;
        h = t.h;
        norm = float(double(w) * h / (2.0 * double(3.14159f) * double(3.14159f)));
    }
    bool ready() const { return w > 0; }

    // Direction and its solid-angle density from four uniforms (cell, alias choice, point in cell).
    template <class T>
    bool sample(const float u[4], V3<T>& d, T& density) const {
        const int N = w * h;
        int k = std::min(int(u[0] * float(N)), N - 1);
        if (!(u[1] < prob[k])) k = alias[k];
        const int i = k / w, j = k - i * w;
        const T th = (T(i) + T(u[3])) / T(h) * T(3.14159f);
        const T ph = ((T(j) + T(u[2])) / T(w) - T(0.5f)) * (T(2.0f) * T(3.14159f));
        const T st = std::sin(th);
        d = V3<T>(st * std::cos(ph), std::cos(th), st * std::sin(ph));
        if (!(st > T(1e-7f))) return false;
        density = T(pdf[k]) * T(norm) / st;
        return true;
    }
    // Density of `sample` at a unit direction (sin theta from x and z: precise near the poles).
    template <class T>
    T density(const V3<T>& d) const {
        const T st = std::sqrt(d.x * d.x + d.z * d.z);
        if (!(st > T(1e-7f))) return T(0);
        const T th = std::atan2(st, d.y), ph = std::atan2(d.z, d.x);
        const int j = std::min(std::max(int(std::floor((ph / (T(2.0f) * T(3.14159f)) + T(0.5f)) * T(w))), 0), w - 1);
        const int i = std::min(std::max(int(std::floor(th / T(3.14159f) * T(h))), 0), h - 1);
        return T(pdf[size_t(i) * w + j]) * T(norm) / st;
    }
};

struct Scene {
    std::vector<Tri> tris;
    std::vector<Material> mats;
    std::vector<Texture> textures;
    Texture env;
    float env_gain = 1.0f;
    bool has_env = false;
    EnvSampler env_sampler;   // built with the map (orc_scene_set_envmap)
    // light sampler state (lightsampler.cpp:4-39)
    std::vector<int> light_tris;
    std::vector<float> light_cdf;
    float light_sum = 0.0f;

    bool is_emitter(int tri) const { return mats[tris[tri].mat].type == MAT_LIGHT; }

    // LightSampler::setPrimitives, lightsampler.cpp:4-39.
    void build_light_sampler() {
        light_tris.clear();
        light_cdf.clear();
        std::vector<float> pdf;
        float sum = 0;
        for (size_t i = 0; i < tris.size(); ++i) {
            if (is_emitter(int(i))) {
                float w = tris[i].area();
                pdf.push_back(w);
                sum += w;
                light_tris.push_back(int(i));
            }
        }
        sum += 1e-8f;
        for (auto& x : pdf) x /= sum;
        light_sum = sum;
        float acc = 1e-6f;
        for (float x : pdf) {
            acc += x;
            light_cdf.push_back(acc);
        }
    }
    // LightSampler::sampleLight, lightsampler.cpp:49-58: lower_bound over the CDF.
    int pick_light(float u) const {
        if (light_tris.empty()) return -1;
        size_t id = std::lower_bound(light_cdf.begin(), light_cdf.end(), u) - light_cdf.begin();
        if (id >= light_tris.size()) id = light_tris.size() - 1;
        return light_tris[id];
    }
    // LightSampler::p, lightsampler.cpp:100-103 (double division, stored as float).
    float light_p() const { return float(1.0 / (double(light_sum) + 1e-8)); }
};

// ---------------------------------------------------------------- materials
// Shading frame shared by MatLambert (matlambert.cpp:11-13) and MatGGXRefl (matggxrefl.cpp:20-22).
// The reference's `abs` is read as the floating-point absolute value: a g++ build without the
// Qt headers in the material translation units would bind int abs, but the reference's published
// renders are only consistent with the floating-point one.
template <class T>
inline void frame(const V3<T>& n, V3<T>& ax1, V3<T>& ax2) {
    V3<T> ax0 = double(std::fabs(n.x)) < 0.8 ? V3<T>(1, 0, 0) : V3<T>(0, 1, 0);
    ax1 = ax0.cross(n).normalized();
    ax2 = n.cross(ax1).normalized();
}

template <class T>
struct MatEval {
    const Scene* sc;
    const Material* m;

    V3<T> albedo(const V3<T>& uv) const {
        if (m->tex >= 0) return V3<T>(sc->textures[m->tex].uv(float(uv.x), float(uv.y)));
        return V3<T>(T(m->kd[0]), T(m->kd[1]), T(m->kd[2]));
    }
    // requireLightSampling / isTransmission (matblinnphong.cpp:88-91, matglass.cpp:55-69)
    bool needs_nee() const { return m->type == MAT_LAMBERT || m->type == MAT_GGXREFL || m->type == MAT_BLINNPHONG; }
    bool transmissive() const { return m->type == MAT_LIGHT || m->type == MAT_GGX || m->type == MAT_GLASS; }
    V3<T> specular(const V3<T>& uv) const {
        if (m->tex_ks >= 0) return V3<T>(sc->textures[m->tex_ks].uv(float(uv.x), float(uv.y)));
        return V3<T>(T(m->ks[0]), T(m->ks[1]), T(m->ks[2]));
    }

    // True solid-angle density of sample() at wi, for the models that take part in the optional
    // sky sampling (0 otherwise).  Lambert (matlambert.cpp:5-18): r^2 = 0.99f U, so cos^2 is
    // uniform on [1 - 0.99f, 1]: q = cos / (0.99f pi) for cos^2 > 1 - 0.99f.  GGX-reflect
    // (matggxrefl.cpp:5-34): theta_h = s atan(alpha sqrt(U/(1-U))), s = 0.9999, phi uniform on
    // [0, 2 * 3.14159f): theta0 = theta_h / s has density
    // p0 = 2 a^2 tan(theta0) (1 + tan^2 theta0) / (a^2 + tan^2 theta0)^2, the half vector
    // q_h = p0 / (s 2 (3.14159f) sin theta_h), the reflection q = q_h / (4 wo.h).
    T sample_density(const V3<T>& wo, const V3<T>& n, const V3<T>& wi) const {
        const T pi = T(3.14159265358979);
        if (m->type == MAT_LAMBERT) {
            const T c = n.dot(wi);
            return c > T(0.1f) ? c / (T(0.99f) * pi) : T(0);
        }
        if (m->type != MAT_GGXREFL) return T(0);
        const V3<T> hv = wo + wi;
        const T hl = std::sqrt(hv.dot(hv));
        if (!(hl > T(1e-7f))) return T(0);
        const V3<T> h = hv * (T(1) / hl);
        const T ch = n.dot(h), woh = wo.dot(h);
        if (!(ch > 0) || !(woh > 0)) return T(0);
        const V3<T> cr = n.cross(h);
        const T sh = std::sqrt(cr.dot(cr));
        const T s = T(0.9999);
        const T th0 = std::atan2(sh, ch) / s;
        if (!(th0 < T(1.5707963f))) return T(0);
        const T t = std::tan(th0);
        const T a2 = T(m->alpha) * T(m->alpha);
        const T k = a2 + t * t;
        const T tan_over_sin = sh > T(1e-6f) ? t / sh : T(1) / s;
        const T qh = T(2) * a2 * tan_over_sin * (T(1) + t * t) / (k * k) / (s * T(2.0f) * T(3.14159f));
        return qh / (T(4) * woh);
    }

    // sampleBxdf: matlambert.cpp:5-18, matggxrefl.cpp:5-34, matggx.cpp:5-47.
    V3<T> sample(const float u[3], const V3<T>& wo, const V3<T>& n) const {
        if (m->type == MAT_LAMBERT) {
            T r2 = T(u[0]) * T(0.99f);
            T phi = T(double(u[1]) * 3.14159 * 2);
            T r = std::sqrt(r2), h = std::sqrt(T(1) - r2);
            V3<T> ax1, ax2;
            frame(n, ax1, ax2);
            return h * n + (r * T(std::cos(phi))) * ax1 + (r * T(std::sin(phi))) * ax2;
        }
        if (m->type == MAT_GGXREFL) {
            V3<T> ax1, ax2;
            frame(n, ax1, ax2);
            T a = T(m->alpha);
            T r1 = T(u[0]), r2 = T(u[1]);
            T theta = T(double(std::atan(a * std::sqrt(r1 / (T(1) - r1)))) * 0.9999);
            T phi = T(2.0f * 3.14159f) * r2;
            V3<T> hv = T(std::cos(theta)) * n + (T(std::sin(theta)) * T(std::cos(phi))) * ax1 +
                       (T(std::sin(theta)) * T(std::sin(phi))) * ax2;
            return (T(2) * hv.dot(wo)) * hv - wo;
        }
        if (m->type == MAT_GGX) {
            V3<T> ax1, ax2;
            frame(n, ax1, ax2);
            T a = T(m->alpha);
            T r1 = T(u[0]), r2 = T(u[1]);
            T ct = std::sqrt((T(1) - r1) / (T(1) + (a * a - T(1)) * r1));
            T theta = T(double(std::acos(ct)) * 0.9999);
            T phi = T(2.0f * 3.14159f) * r2;
            V3<T> hv = T(std::cos(theta)) * n + (T(std::sin(theta)) * T(std::cos(phi))) * ax1 +
                       (T(std::sin(theta)) * T(std::sin(phi))) * ax2;
            V3<T> wi;
            if (u[2] < 0.5f) {
                wi = (T(2) * hv.dot(wo)) * hv - wo;
            } else {
                // matggx.cpp:43 reads the still-zero `wi`, so eta is always 1/ior.
                T eta = T(1) / T(m->ior);
                T nh = std::fabs((eta + T(1)) * wo.dot(hv));
                wi = -eta * wo - hv * nh;
            }
            return wi.normalized();
        }
        if (m->type == MAT_BLINNPHONG) {  // matblinnphong.cpp:5-34: lobe 1/2 each, wi not normalised
            V3<T> ax1, ax2;
            frame(n, ax1, ax2);
            if (u[0] < 0.5f) {
                T r2 = T(u[1]) * T(0.99f);
                T phi = T(double(u[2]) * 3.14159 * 2);
                T r = std::sqrt(r2), h = std::sqrt(T(1) - r2);
                return h * n + (r * T(std::cos(phi))) * ax1 + (r * T(std::sin(phi))) * ax2;
            }
            const double r1 = u[1], r2 = u[2];
            const double ct = std::pow(r1, double(1.0f / (m->alpha + 2.0f)));
            const double st = std::sqrt(1 - ct);  // the reference's sqrt(1 - cos), not sqrt(1 - cos^2)
            const double phi = double(2.0f * 3.14159f) * r2;
            V3<T> h = T(float(ct)) * n + T(float(st * std::cos(phi))) * ax1 + T(float(st * std::sin(phi))) * ax2;
            return (T(2) * h.dot(wo)) * h - wo;
        }
        if (m->type == MAT_GLASS) {  // matglass.cpp:5-24
            const bool out = wo.dot(n) > 0;
            const T I = out ? T(1) / T(m->ior) : T(m->ior);
            const V3<T> nn = out ? n : -n;
            if (u[0] < 0.5f) {
                const V3<T> t = nn.cross(wo).cross(nn).normalized();
                const T so = wo.dot(t);
                const T si = std::min(so * I, T(1));
                const T ci = std::sqrt(T(1) - si * si);
                return -(ci * nn + si * t);
            }
            return (T(2) * wo.dot(nn)) * nn - wo;
        }
        return V3<T>(0);
    }

    static T ggx_d(T a, T c) {  // D of matggxrefl.cpp:46 / matggx.cpp:77-82 (for n.m >= 0)
        T a2 = a * a;
        T k = c * c * (a2 - T(1)) + T(1);
        return a2 / pi_f() / (k * k);
    }
    static T pi_f() { return T(3.14159f); }

    // bxdf: matlambert.cpp:20-32, matggxrefl.cpp:36-52, matggx.cpp:54-121.
    V3<T> f(const V3<T>& wo, const V3<T>& n, const V3<T>& wi, const V3<T>& uv) const {
        if (m->type == MAT_LAMBERT) {
            if (wi.dot(n) * wo.dot(n) < 0) return V3<T>(0);
            return albedo(uv) / T(3.14159);
        }
        if (m->type == MAT_GGXREFL) {
            if (wi.dot(n) * wo.dot(n) < 0) return V3<T>(0);
            V3<T> kd = albedo(uv);
            T a = T(m->alpha);
            V3<T> wh = (wo + wi).normalized();
            T D = ggx_d(a, n.dot(wh));
            T k = a * T(0.5) + T(0.5);
            T ci = n.dot(wi), co = n.dot(wo);
            T Gl = T(2) * ci / (ci + std::sqrt(k * k + (T(1) - k * k * ci * ci)));
            T Gv = T(2) * co / (co + std::sqrt(k * k + (T(1) - k * k * co * co)));
            T fr = std::pow(T(1) - wh.dot(wi), T(5));
            V3<T> F = kd + (V3<T>(1) - kd) * fr;
            T s = D * Gl * Gv / T(4) / (co + T(1e-6f)) / (ci + T(1e-6f));
            return F * s;
        }
        if (m->type == MAT_GGX) {
            V3<T> kd = albedo(uv);
            T bi = wi.dot(n), bo = wo.dot(n);
            if (bi * bo >= 0) return ggx_fr(wi, wo, n, kd);
            return ggx_ft(wi, wo, n, kd);
        }
        if (m->type == MAT_BLINNPHONG) {  // matblinnphong.cpp:36-49
            if (wi.dot(n) * wo.dot(n) < 0) return V3<T>(0);
            if (wi.dot(n) <= T(1e-4) || wo.dot(n) <= T(1e-4f)) return V3<T>(0);
            const V3<T> h = (wo + wi).normalized();
            const T ns = T(m->alpha);
            return albedo(uv) / T(3.14159) +
                   specular(uv) * ((ns + T(2)) / T(3.14159) / T(8) * T(std::pow(n.dot(h), ns)));
        }
        if (m->type == MAT_GLASS) {  // matglass.cpp:26-44
            if (std::fabs(wi.dot(n)) < T(1e-3f)) return V3<T>(0);
            const V3<T> f0 = albedo(uv);
            const V3<T> F = f0 + (V3<T>(1) - f0) * T(std::pow(T(1) - std::fabs(n.dot(wo)), T(5)));
            if (wi.dot(n) * wo.dot(n) > 0) return F;
            const T I = wo.dot(n) > 0 ? T(1) / T(m->ior) : T(m->ior);
            return (V3<T>(1) - F) * T(std::pow(I, T(2)));
        }
        return V3<T>(0);
    }

    // pdf: matlambert.cpp:34-37, matggxrefl.cpp:54-64, matggx.cpp:123-148, matlight.cpp:15-18.
    T pdf(const V3<T>& wo, const V3<T>& n, const V3<T>& wi) const {
        if (m->type == MAT_LAMBERT) return T(1.0 / 3.14159) * (std::fabs(wi.dot(n)) + T(1e-4f));
        if (m->type == MAT_GGXREFL) {
            if (double(wi.dot(n)) <= 1e-4 || wo.dot(n) <= T(1e-4f)) return T(1e18f);
            V3<T> wh = (wo + wi).normalized();
            T D = ggx_d(T(m->alpha), n.dot(wh));
            return D * n.dot(wh) / T(4) + T(1e-8f);
        }
        if (m->type == MAT_GGX) {
            if (wi.dot(n) * wo.dot(n) > 0) {
                V3<T> wh = (wo + wi).normalized();
                T d = ggx_D_signed(wh, n);
                if (d < T(1e-9f)) return T(1e18f);
                return d * n.dot(wh) / T(4);
            }
            T eta = wi.dot(n) > 0 ? T(m->ior) : T(1) / T(m->ior);
            V3<T> wh = -(wi + eta * wo).normalized();
            if (wh.dot(n) < 0) wh = -wh;
            T d = ggx_D_signed(wh, n);
            if (d < T(1e-6f)) return T(1e18f);
            return std::max(d * n.dot(wh) / T(4), T(1));
        }
        if (m->type == MAT_BLINNPHONG) {  // matblinnphong.cpp:51-59
            if (wi.dot(n) <= T(1e-4) || wo.dot(n) <= T(1e-4f)) return T(1e-18f);
            const T pd = T(1.0 / 3.14159) * (std::fabs(wi.dot(n)) + T(1e-6f));
            const V3<T> h = (wo + wi).normalized();
            const T ns = T(m->alpha);
            const T ps = (ns + T(2)) / T(3.14159) / T(8) * T(std::pow(std::fabs(n.dot(h)), ns + T(1)));
            return T(0.5f) * pd + (T(1) - T(0.5f)) * ps;
        }
        if (m->type == MAT_GLASS) return T(0.5);  // matglass.cpp:46-49
        return T(1.0 / 3.14159);
    }

    // MatGGX helpers (matggx.cpp:49-121).
    T ggx_D_signed(const V3<T>& mm, const V3<T>& n) const {
        if (n.dot(mm) < 0) return 0;
        return ggx_d(T(m->alpha), n.dot(mm));
    }
    V3<T> ggx_F(const V3<T>& v, const V3<T>& mm, const V3<T>& kd) const {
        if (v.dot(mm) < 0) return V3<T>(0);
        return kd + (V3<T>(1) - kd) * std::pow(T(1) - v.dot(mm), T(5));
    }
    T ggx_G1(const V3<T>& v, const V3<T>& mm, const V3<T>& n) const {
        T vm = v.dot(mm), vn = v.dot(n);
        if (vm * vn < 0) return 0;
        T tn = T(m->alpha) * std::tan(std::acos(vn));
        T r2 = T(1) + tn * tn;
        return T(2) / (T(1) + std::sqrt(r2));
    }
    V3<T> ggx_fr(const V3<T>& i, const V3<T>& o, const V3<T>& n, const V3<T>& kd) const {
        if (n.dot(i) * n.dot(o) < 0) return V3<T>(0);
        V3<T> h = (i + o).normalized();
        V3<T> nn = i.dot(n) > 0 ? n : -n;
        V3<T> F = ggx_F(i, h, kd);
        T g = ggx_G1(i, h, nn) * ggx_G1(o, h, nn);
        T d = ggx_D_signed(h, nn);
        T denom = T(4) * i.dot(nn) * o.dot(nn);
        return F * (g * d) / (denom + T(1e-6f));
    }
    V3<T> ggx_ft(const V3<T>& i, const V3<T>& o, const V3<T>& n, const V3<T>& kd) const {
        if (n.dot(i) * n.dot(o) > 0) return V3<T>(0);
        T eta = i.dot(n) > 0 ? T(m->ior) : T(1) / T(m->ior);
        V3<T> h = -(i + eta * o).normalized();
        T A = std::fabs(i.dot(h) * o.dot(h) / i.dot(n) / o.dot(n));
        V3<T> F = ggx_F(i, h, kd);
        T g = ggx_G1(i, h, n) * ggx_G1(o, h, n);
        T d = ggx_D_signed(h, n);
        V3<T> numer = (eta * eta) * ((V3<T>(1) - F) * (g * d));
        T c = i.dot(h) + eta * o.dot(h);
        return numer * (A / (c * c + T(1e-8f)));
    }
};

}  // namespace orc
