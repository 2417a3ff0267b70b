// Environment importance sampling for ptnee: the sky is sampled like a light at Lambert and
// GGX-reflect vertices and combined with BSDF sampling by the balance heuristic.
//
// The reference has no such strategy (its sky is reached by BSDF sampling only,
// rendererptls.cpp:16-28, 84-88), so the weights are built to keep the reference's expected
// image.  At a vertex the reference's sky term is the integral of
//     g(w) L(w) V(w),   g = q(w) / p_rep(w) * f(w) |cos|,
// where q is the true density of the material's sampler and p_rep the pdf the material reports
// (they differ: the Lambert sampler stops at cos = 0.1, the GGX-reflect pdf has no 1 / (4 wo.h)).
// With the sky sampler's density p_env the two strategies contribute
//     sky sample w_e:    g(w_e) L V / (p_env(w_e) + q(w_e))
//     BSDF escape w_i:   (f |cos| / p_rep / prr) L * q(w_i) / (q(w_i) + p_env(w_i))
// whose expectations add up to the integral above: a variance reduction of the reference's own
// estimator, not a different image.  (The samplers' azimuth ranges use 2 * 3.14159, leaving a
// sliver of about 1e-6 of the circle that the weights do not single out.)
#pragma once
#include "nr_material.h"

namespace nr {

// True solid-angle density of the reference's BSDF sampler at wi for the two models that use the
// sky sampler (0 for any other model).
//  * Lambert (matlambert.cpp:5-18): r^2 = 0.99 U, so cos^2 is uniform on [0.01, 1] and
//    q = cos / (0.99 pi) for cos > 0.1.
//  * GGX-reflect (matggxrefl.cpp:5-34): theta_h = s * atan(alpha sqrt(U / (1 - U))) with
//    s = 0.9999, phi uniform on [0, 2 * 3.14159).  theta0 = theta_h / s has the GGX density
//    p0 = 2 alpha^2 tan(theta0) (1 + tan^2 theta0) / (alpha^2 + tan^2 theta0)^2, so the half vector
//    has q_h = p0(theta_h / s) / (s * 2 * 3.14159 * sin theta_h) and the reflected direction
//    q = q_h / (4 wo.h).
__device__ __forceinline__ float bsdf_sample_density(const Mat& m, F3 wo, F3 n, F3 wi) {
    if (m.type == MAT_LAMBERT) {
        const float c = dot(n, wi);
        return c > 0.1f ? c * (1.0f / (0.99f * 3.14159265f)) : 0.0f;
    }
    if (m.type != MAT_GGXREFL) return 0.0f;
    const F3 hv = add(wo, wi);
    const float hl = norm(hv);
    if (!(hl > 1e-7f)) return 0.0f;
    const F3 h = mul(hv, 1.0f / hl);
    const float ch = dot(n, h), woh = dot(wo, h);
    if (!(ch > 0.0f) || !(woh > 0.0f)) return 0.0f;
    const float sh = norm(cross(n, h));
    const float th0 = atan2f(sh, ch) / 0.9999f;
    if (!(th0 < 1.5707963f)) return 0.0f;
    const float t = tanf(th0);
    const float a2 = m.alpha * m.alpha;
    const float k = a2 + t * t;
    const float tan_over_sin = sh > 1e-6f ? t / sh : 1.0f / 0.9999f;  // tan(theta/s)/sin(theta) -> 1/s
    const float qh = 2.0f * a2 * tan_over_sin * (1.0f + t * t) / (k * k) / (0.9999f * 2.0f * 3.14159f);
    return qh / (4.0f * woh);
}

__device__ __forceinline__ bool env_sampling_vertex(const Mat& m) {
    return m.type == MAT_LAMBERT || m.type == MAT_GGXREFL;
}

// Draws a sky direction from four uniforms: cell by the alias table (u.x cell, u.y alias choice),
// then a point uniform in the cell (u.z, u.w).  Returns false for a direction on a pole.
__device__ __forceinline__ bool env_sample(const SceneView& sv, float4 u, F3& d, float& pdf) {
    const int W = sv.env_w, N = sv.env_w * sv.env_h;
    int k = min(int(u.x * float(N)), N - 1);
    const float4 e = sv.env_tab[k];
    float pc = e.z;
    if (!(u.y < e.x)) {
        k = __float_as_int(e.y);
        pc = e.w;
    }
    const int i = k / W, j = k - i * W;
    const float th = (float(i) + u.w) / float(sv.env_h) * 3.14159f;
    const float ph = ((float(j) + u.z) / float(W) - 0.5f) * (2.0f * 3.14159f);
    const float st = sinf(th);
    d = f3(st * cosf(ph), cosf(th), st * sinf(ph));
    if (!(st > 1e-7f)) return false;
    pdf = pc * sv.env_norm / st;
    return true;
}

// Solid-angle density of env_sample at direction d (unit length).  sin(theta) from the x, z
// components keeps its precision near the poles.
__device__ __forceinline__ float env_pdf(const SceneView& sv, F3 d) {
    const float st = sqrtf(d.x * d.x + d.z * d.z);
    if (!(st > 1e-7f)) return 0.0f;
    const float th = atan2f(st, d.y);
    const float ph = atan2f(d.z, d.x);
    const int W = sv.env_w, H = sv.env_h;
    const int j = min(max(int(floorf((ph / (2.0f * 3.14159f) + 0.5f) * float(W))), 0), W - 1);
    const int i = min(max(int(floorf(th / 3.14159f * float(H))), 0), H - 1);
    return sv.env_tab[i * W + j].z * sv.env_norm / st;
}

}  // namespace nr
