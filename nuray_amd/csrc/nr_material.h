// Device materials, textures and environment lookup (FP32), with the reference's semantics.
//
//   MatLambert   src/material/matlambert.cpp:5-79
//   MatGGXRefl   src/material/matggxrefl.cpp:5-88   (Kd is F0; no diffuse lobe)
//   MatGGX       src/material/matggx.cpp:5-173      (translucent; eta always 1/ior, matggx.cpp:43)
//   MatLight     src/material/matlight.cpp:1-42
//   MatBlinnPhong src/material/matblinnphong.cpp:5-85 (sampled wi not normalised, sin = sqrt(1 - cos))
//   MatGlass     src/material/matglass.cpp:5-48
//   Texture      src/scene/texture.cpp:54-72        (black outside [0,W)x[0,H), no wrap of x0+1)
// `abs` is the floating-point absolute value throughout.
#pragma once
#include "nr_device.h"

namespace nr {

// Optional material features a kernel instance is compiled with (the host launches the instance
// matching the scene): F_GGX = translucent GGX dielectric present, F_TEX = albedo textures present,
// F_EXT = Blinn-Phong or glass present.
constexpr unsigned F_GGX = 1u, F_TEX = 2u, F_EXT = 4u, F_ALL = 7u;

struct Mat {
    int type;
    F3 kd;
    float alpha, ior;
    int tex;
    F3 ks;       // Blinn-Phong
    int tex_ks;
};
__device__ __forceinline__ Mat load_mat(const SceneView& sv, int id) {
    const MatRec m = sv.mats[id];
    return Mat{int(m.a.x), f3(m.a.y, m.a.z, m.a.w), m.b.x, m.b.y, int(m.b.z), f3(m.c.x, m.c.y, m.c.z), int(m.c.w)};
}
// requireLightSampling / isTransmission (matlambert.cpp:54-57, matggxrefl.cpp, matblinnphong.cpp:88-91,
// matggx.cpp:155-169, matglass.cpp:55-69, matlight.cpp:25-28)
__device__ __forceinline__ bool mat_needs_nee(const Mat& m) {
    return m.type == MAT_LAMBERT || m.type == MAT_GGXREFL || m.type == MAT_BLINNPHONG;
}
__device__ __forceinline__ bool mat_transmissive(const Mat& m) {
    return m.type == MAT_LIGHT || m.type == MAT_GGX || m.type == MAT_GLASS;
}

// ---------------------------------------------------------------- textures
__device__ __forceinline__ F3 texel(const SceneView& sv, const TexDesc& t, int x, int y) {
    if (x < 0 || y < 0 || x >= t.w || y >= t.h) return f3(0.0f, 0.0f, 0.0f);
    const size_t k = size_t(y) * t.w + x;
    if (t.fmt == 0) {
        const uchar4 c = reinterpret_cast<const uchar4*>(sv.texpool + t.offset)[k];
        return f3(sv.lut[c.x], sv.lut[c.y], sv.lut[c.z]);
    }
    const float* q = reinterpret_cast<const float*>(sv.texpool + t.offset) + 3 * k;
    return f3(q[0], q[1], q[2]);
}
// Texture::pixelUV -> pixelBi (texture.cpp:61-72).
__device__ __forceinline__ F3 tex_uv(const SceneView& sv, int id, float u, float v) {
    const TexDesc t = sv.tex[id];
    const float fu = fmodf(fmodf(u, 1.0f) + 1.0f, 1.0f);
    const float fv = fmodf(fmodf(1.0f - v, 1.0f) + 1.0f, 1.0f);
    const float x = fu * float(t.w), y = fv * float(t.h);
    const int x0 = int(floorf(x)), y0 = int(floorf(y));
    const int x1 = x0 + 1, y1 = y0 + 1;
    F3 r = mul(mul(texel(sv, t, x1, y1), x - float(x0)), y - float(y0));
    r = add(r, mul(mul(texel(sv, t, x1, y0), x - float(x0)), float(y1) - y));
    r = add(r, mul(mul(texel(sv, t, x0, y1), float(x1) - x), y - float(y0)));
    r = add(r, mul(mul(texel(sv, t, x0, y0), float(x1) - x), float(y1) - y));
    return r;
}
// Miss shading, rendererptls.cpp:18-28.  The reference's double promotion of acos / atan2 is a
// C++ literal artefact; FP32 here moves u, v by at most a couple of ulps.
__device__ __forceinline__ F3 env_radiance(const SceneView& sv, F3 d) {
    if (sv.env_tex < 0) return f3(0.0f, 0.0f, 0.0f);
    const float v = 1.0f - acosf(d.y) / 3.14159f;
    const float u = atan2f(d.z, d.x) / 2.0f / 3.14159f + 0.5f;
    return mul(tex_uv(sv, sv.env_tex, u, v), sv.env_gain);
}
template <unsigned F = F_ALL>
__device__ __forceinline__ F3 mat_albedo(const SceneView& sv, const Mat& m, float u, float v) {
    if (F & F_TEX) return m.tex >= 0 ? tex_uv(sv, m.tex, u, v) : m.kd;
    return m.kd;
}
template <unsigned F = F_ALL>
__device__ __forceinline__ F3 mat_specular(const SceneView& sv, const Mat& m, float u, float v) {
    if (F & F_TEX) return m.tex_ks >= 0 ? tex_uv(sv, m.tex_ks, u, v) : m.ks;
    return m.ks;
}

// ---------------------------------------------------------------- BSDFs
__device__ __forceinline__ void frame(F3 n, F3& ax1, F3& ax2) {
    const F3 ax0 = fabsf(n.x) < 0.8f ? f3(1.0f, 0.0f, 0.0f) : f3(0.0f, 1.0f, 0.0f);
    ax1 = normalized(cross(ax0, n));
    ax2 = normalized(cross(n, ax1));
}
__device__ __forceinline__ float ggx_d(float a, float c) {
    const float a2 = a * a;
    const float k = c * c * (a2 - 1.0f) + 1.0f;
    return a2 / 3.14159f / (k * k);
}

// sampleBxdf; u = (r1, r2, lobe)
template <unsigned F = F_ALL>
__device__ __forceinline__ F3 bsdf_sample(const Mat& m, float u0, float u1, float u2, F3 wo, F3 n) {
    F3 ax1, ax2;
    frame(n, ax1, ax2);
    if (m.type == MAT_LAMBERT) {
        const float r2 = u0 * 0.99f;
        const float phi = float(double(u1) * 3.14159 * 2);
        const float r = sqrtf(r2), h = sqrtf(1.0f - r2);
        float sp, cp;
        sincosf(phi, &sp, &cp);
        return add(add(mul(n, h), mul(ax1, r * cp)), mul(ax2, r * sp));
    }
    if (m.type == MAT_GGXREFL) {
        const float theta = float(double(atanf(m.alpha * sqrtf(u0 / (1.0f - u0)))) * 0.9999);
        const float phi = (2.0f * 3.14159f) * u1;
        float st, ct, sp, cp;
        sincosf(theta, &st, &ct);
        sincosf(phi, &sp, &cp);
        const F3 h = add(add(mul(n, ct), mul(ax1, st * cp)), mul(ax2, st * sp));
        return sub(mul(h, 2.0f * dot(h, wo)), wo);
    }
    if ((F & F_GGX) && m.type == MAT_GGX) {
        const float a = m.alpha;
        const float c = sqrtf((1.0f - u0) / (1.0f + (a * a - 1.0f) * u0));
        const float theta = float(double(acosf(c)) * 0.9999);
        const float phi = (2.0f * 3.14159f) * u1;
        float st, ct, sp, cp;
        sincosf(theta, &st, &ct);
        sincosf(phi, &sp, &cp);
        const F3 h = add(add(mul(n, ct), mul(ax1, st * cp)), mul(ax2, st * sp));
        F3 wi;
        if (u2 < 0.5f) {
            wi = sub(mul(h, 2.0f * dot(h, wo)), wo);
        } else {
            const float eta = 1.0f / m.ior;
            const float nh = fabsf((eta + 1.0f) * dot(wo, h));
            wi = sub(mul(wo, -eta), mul(h, nh));
        }
        return normalized(wi);
    }
    if ((F & F_EXT) && m.type == MAT_BLINNPHONG) {  // matblinnphong.cpp:5-34, u0 picks the lobe
        if (u0 < 0.5f) {
            const float r2 = u1 * 0.99f;
            const float phi = float(double(u2) * 3.14159 * 2);
            const float r = sqrtf(r2), h = sqrtf(1.0f - r2);
            float sp, cp;
            sincosf(phi, &sp, &cp);
            return add(add(mul(n, h), mul(ax1, r * cp)), mul(ax2, r * sp));
        }
        const float ct = powf(u1, 1.0f / (m.alpha + 2.0f));
        const float st = sqrtf(1.0f - ct);  // the reference's sqrt(1 - cos)
        const float phi = (2.0f * 3.14159f) * u2;
        float sp, cp;
        sincosf(phi, &sp, &cp);
        const F3 h = add(add(mul(n, ct), mul(ax1, st * cp)), mul(ax2, st * sp));
        return sub(mul(h, 2.0f * dot(h, wo)), wo);  // not normalised, as the reference
    }
    if ((F & F_EXT) && m.type == MAT_GLASS) {  // matglass.cpp:5-24, u0 picks refraction / reflection
        const bool out = dot(wo, n) > 0.0f;
        const float I = out ? 1.0f / m.ior : m.ior;
        const F3 nn = out ? n : neg(n);
        if (u0 < 0.5f) {
            const F3 t = normalized(cross(cross(nn, wo), nn));
            const float so = dot(wo, t);
            const float si = fminf(so * I, 1.0f);
            const float ci = sqrtf(1.0f - si * si);
            return neg(add(mul(nn, ci), mul(t, si)));
        }
        return sub(mul(nn, 2.0f * dot(wo, nn)), wo);
    }
    return f3(0.0f, 0.0f, 0.0f);
}

__device__ __forceinline__ float ggx_D_signed(float a, F3 mm, F3 n) {
    const float c = dot(n, mm);
    return c < 0.0f ? 0.0f : ggx_d(a, c);
}
__device__ __forceinline__ F3 ggx_F(F3 v, F3 mm, F3 kd) {
    const float c = dot(v, mm);
    if (c < 0.0f) return f3(0.0f, 0.0f, 0.0f);
    const float f = powf(1.0f - c, 5.0f);
    return f3(kd.x + (1.0f - kd.x) * f, kd.y + (1.0f - kd.y) * f, kd.z + (1.0f - kd.z) * f);
}
__device__ __forceinline__ float ggx_G1(float a, F3 v, F3 mm, F3 n) {
    const float vm = dot(v, mm), vn = dot(v, n);
    if (vm * vn < 0.0f) return 0.0f;
    const float tn = a * tanf(acosf(vn));
    return 2.0f / (1.0f + sqrtf(1.0f + tn * tn));
}

template <unsigned F = F_ALL>
__device__ __forceinline__ F3 bsdf_eval(const SceneView& sv, const Mat& m, F3 wo, F3 n, F3 wi, float u, float v) {
    if (m.type == MAT_LAMBERT) {
        if (dot(wi, n) * dot(wo, n) < 0.0f) return f3(0.0f, 0.0f, 0.0f);
        return divs(mat_albedo<F>(sv, m, u, v), 3.14159f);
    }
    if (m.type == MAT_GGXREFL) {
        if (dot(wi, n) * dot(wo, n) < 0.0f) return f3(0.0f, 0.0f, 0.0f);
        const F3 kd = mat_albedo<F>(sv, m, u, v);
        const F3 wh = normalized(add(wo, wi));
        const float D = ggx_d(m.alpha, dot(n, wh));
        const float k = m.alpha * 0.5f + 0.5f;
        const float ci = dot(n, wi), co = dot(n, wo);
        const float Gl = 2.0f * ci / (ci + sqrtf(k * k + (1.0f - k * k * ci * ci)));
        const float Gv = 2.0f * co / (co + sqrtf(k * k + (1.0f - k * k * co * co)));
        const float fr = powf(1.0f - dot(wh, wi), 5.0f);
        const float s = D * Gl * Gv / 4.0f / (co + 1e-6f) / (ci + 1e-6f);
        return f3((kd.x + (1.0f - kd.x) * fr) * s, (kd.y + (1.0f - kd.y) * fr) * s, (kd.z + (1.0f - kd.z) * fr) * s);
    }
    if ((F & F_GGX) && m.type == MAT_GGX) {
        const F3 kd = mat_albedo<F>(sv, m, u, v);
        const float bi = dot(wi, n), bo = dot(wo, n);
        if (bi * bo >= 0.0f) {  // fr
            if (dot(n, wi) * dot(n, wo) < 0.0f) return f3(0.0f, 0.0f, 0.0f);
            const F3 h = normalized(add(wi, wo));
            const F3 nn = dot(wi, n) > 0.0f ? n : neg(n);
            const F3 Fr = ggx_F(wi, h, kd);
            const float g = ggx_G1(m.alpha, wi, h, nn) * ggx_G1(m.alpha, wo, h, nn);
            const float dd = ggx_D_signed(m.alpha, h, nn);
            const float denom = 4.0f * dot(wi, nn) * dot(wo, nn);
            return mul(Fr, g * dd / (denom + 1e-6f));
        }
        // ft
        const float eta = dot(wi, n) > 0.0f ? m.ior : 1.0f / m.ior;
        const F3 h = neg(normalized(add(wi, mul(wo, eta))));
        const float A = fabsf(dot(wi, h) * dot(wo, h) / dot(wi, n) / dot(wo, n));
        const F3 Fr = ggx_F(wi, h, kd);
        const float g = ggx_G1(m.alpha, wi, h, n) * ggx_G1(m.alpha, wo, h, n);
        const float dd = ggx_D_signed(m.alpha, h, n);
        const float c = dot(wi, h) + eta * dot(wo, h);
        const float s = A * (eta * eta) * g * dd / (c * c + 1e-8f);
        return f3((1.0f - Fr.x) * s, (1.0f - Fr.y) * s, (1.0f - Fr.z) * s);
    }
    if ((F & F_EXT) && m.type == MAT_BLINNPHONG) {  // matblinnphong.cpp:36-49
        if (dot(wi, n) * dot(wo, n) < 0.0f) return f3(0.0f, 0.0f, 0.0f);
        if (double(dot(wi, n)) <= 1e-4 || dot(wo, n) <= 1e-4f) return f3(0.0f, 0.0f, 0.0f);
        const F3 h = normalized(add(wo, wi));
        const float sp = (m.alpha + 2.0f) / 3.14159f / 8.0f * powf(dot(n, h), m.alpha);
        return add(divs(mat_albedo<F>(sv, m, u, v), 3.14159f), mul(mat_specular<F>(sv, m, u, v), sp));
    }
    if ((F & F_EXT) && m.type == MAT_GLASS) {  // matglass.cpp:26-44
        if (fabsf(dot(wi, n)) < 1e-3f) return f3(0.0f, 0.0f, 0.0f);
        const F3 f0 = mat_albedo<F>(sv, m, u, v);
        const float fr = powf(1.0f - fabsf(dot(n, wo)), 5.0f);
        const F3 Fr = f3(f0.x + (1.0f - f0.x) * fr, f0.y + (1.0f - f0.y) * fr, f0.z + (1.0f - f0.z) * fr);
        if (dot(wi, n) * dot(wo, n) > 0.0f) return Fr;
        const float I = dot(wo, n) > 0.0f ? 1.0f / m.ior : m.ior;
        return mul(f3(1.0f - Fr.x, 1.0f - Fr.y, 1.0f - Fr.z), I * I);
    }
    return f3(0.0f, 0.0f, 0.0f);
}

template <unsigned F = F_ALL>
__device__ __forceinline__ float bsdf_pdf(const Mat& m, F3 wo, F3 n, F3 wi) {
    if (m.type == MAT_LAMBERT) return float(1.0 / 3.14159) * (fabsf(dot(wi, n)) + 1e-4f);
    if (m.type == MAT_GGXREFL) {
        if (double(dot(wi, n)) <= 1e-4 || dot(wo, n) <= 1e-4f) return 1e18f;
        const F3 wh = normalized(add(wo, wi));
        const float c = dot(n, wh);
        return ggx_d(m.alpha, c) * c / 4.0f + 1e-8f;
    }
    if ((F & F_GGX) && m.type == MAT_GGX) {
        if (dot(wi, n) * dot(wo, n) > 0.0f) {
            const F3 wh = normalized(add(wo, wi));
            const float d = ggx_D_signed(m.alpha, wh, n);
            if (d < 1e-9f) return 1e18f;
            return d * dot(n, wh) / 4.0f;
        }
        const float eta = dot(wi, n) > 0.0f ? m.ior : 1.0f / m.ior;
        F3 wh = neg(normalized(add(wi, mul(wo, eta))));
        if (dot(wh, n) < 0.0f) wh = neg(wh);
        const float d = ggx_D_signed(m.alpha, wh, n);
        if (d < 1e-6f) return 1e18f;
        return fmaxf(d * dot(n, wh) / 4.0f, 1.0f);
    }
    if ((F & F_EXT) && m.type == MAT_BLINNPHONG) {  // matblinnphong.cpp:51-59
        if (double(dot(wi, n)) <= 1e-4 || dot(wo, n) <= 1e-4f) return 1e-18f;
        const float pd = float(1.0 / 3.14159 * double(fabsf(dot(wi, n)) + 1e-6f));
        const F3 h = normalized(add(wo, wi));
        const float ps = (m.alpha + 2.0f) / 3.14159f / 8.0f * powf(fabsf(dot(n, h)), m.alpha + 1.0f);
        return 0.5f * pd + (1.0f - 0.5f) * ps;
    }
    if ((F & F_EXT) && m.type == MAT_GLASS) return 0.5f;  // matglass.cpp:46-49
    return float(1.0 / 3.14159);
}

// ---------------------------------------------------------------- light sampler
// LightSampler::sampleLight: lower_bound over the CDF (lightsampler.cpp:49-58).
__device__ __forceinline__ int pick_light(const SceneView& sv, float u) {
    int lo = 0, hi = sv.n_lights;  // first index with cdf >= u
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (sv.light_cdf[mid] < u) lo = mid + 1;
        else hi = mid;
    }
    return lo < sv.n_lights ? lo : sv.n_lights - 1;
}

}  // namespace nr
