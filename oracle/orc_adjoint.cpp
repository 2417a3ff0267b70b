// ORACLE -- test infrastructure only (see orc_common.h).
//
// FP64 gradient of an image loss with respect to per-material {Kd.r, Kd.g, Kd.b, alpha} for the
// ptnee estimator of rendererptls.cpp:13-93, computed by a forward (tangent) recursion: every call
// of the recursive estimator returns its radiance and the sparse derivative of that radiance.
// Same derivative definition as the GPU adjoint (sampling detached, plus the log-derivative of the
// GGX sampler's density for alpha), computed in a different way (no replay, no radiance
// subtraction), so agreement checks the GPU's path replay.
#include <omp.h>
#include <map>
#include "orc_handle.h"

namespace orc {
namespace {

struct SG {  // sparse gradient: parameter -> dL (RGB)
    std::vector<std::pair<int, V3d>> e;
    void add(int k, const V3d& v) {
        if (v.x == 0 && v.y == 0 && v.z == 0) return;
        for (auto& p : e)
            if (p.first == k) { p.second += v; return; }
        e.push_back({k, v});
    }
    void add_scaled(const SG& o, const V3d& s) {
        for (const auto& p : o.e) add(p.first, p.second * s);
    }
};

constexpr double PI_R = 3.14159f;  // the reference's float literal 3.14159f in D (matggxrefl.cpp:46)

struct MatD {  // f, pdf and their parameter derivatives at one direction
    V3d f, df_dkd, df_da;
    double pdf = 0, dpdf_da = 0;
};

// kd: the material's albedo at the hit (Kd, or its texture's bilinear value).
MatD mat_deriv(const Material& m, const V3d& kd, const V3d& wo, const V3d& n, const V3d& wi) {
    MatD r;
    if (m.type == MAT_LAMBERT) {
        const bool same = !(wi.dot(n) * wo.dot(n) < 0);
        const double inv = same ? 1.0 / 3.14159 : 0.0;
        r.f = V3d(kd.x * inv, kd.y * inv, kd.z * inv);
        r.df_dkd = V3d(inv);
        r.pdf = 1.0 / 3.14159 * (std::fabs(wi.dot(n)) + 1e-4);
        return r;
    }
    // MatGGXRefl
    const double a = m.alpha;
    auto D = [&](double c, double& dD) {
        const double k = c * c * (a * a - 1) + 1;
        dD = 2 * a / (PI_R * k * k * k) * (k - 2 * a * a * c * c);
        return a * a / PI_R / (k * k);
    };
    auto G = [&](double c, double& dG) {
        const double kap = a * 0.5 + 0.5;
        const double S = std::sqrt(kap * kap + (1 - kap * kap * c * c));
        dG = -2 * c / ((c + S) * (c + S)) * (kap * (1 - c * c) / (2 * S));
        return 2 * c / (c + S);
    };
    const double ci = n.dot(wi), co = n.dot(wo);
    if (ci <= 1e-4 || co <= 1e-4) {
        r.pdf = 1e18;
    } else {
        const V3d wh = (wo + wi).normalized();
        double dD;
        const double d0 = D(n.dot(wh), dD);
        r.pdf = d0 * n.dot(wh) / 4 + 1e-8;
        r.dpdf_da = dD * n.dot(wh) / 4;
    }
    if (wi.dot(n) * wo.dot(n) < 0) return r;
    const V3d wh = (wo + wi).normalized();
    double dD, dGl, dGv;
    const double Dv = D(n.dot(wh), dD), Gl = G(ci, dGl), Gv = G(co, dGv);
    const double fr = std::pow(1 - wh.dot(wi), 5.0);
    const double den = 4 * (co + 1e-6) * (ci + 1e-6);
    const double s = Dv * Gl * Gv / den;
    const double ds = (dD * Gl * Gv + Dv * dGl * Gv + Dv * Gl * dGv) / den;
    const V3d F(kd.x + (1 - kd.x) * fr, kd.y + (1 - kd.y) * fr, kd.z + (1 - kd.z) * fr);
    r.f = F * s;
    r.df_dkd = V3d(s * (1 - fr));
    r.df_da = F * ds;
    return r;
}

// log-derivative in alpha of the reference GGX sampler's density (matggxrefl.cpp:24-29)
double ggx_score(double a, const V3d& wo, const V3d& n, const V3d& wi) {
    const V3d h = (wo + wi).normalized();
    const double ch = std::min(1.0, std::max(-1.0, n.dot(h)));
    const double th0 = std::acos(ch) / 0.9999;
    if (!(th0 < 1.5707963)) return 0.0;
    const double t = std::tan(th0);
    return 2 / a - 4 * a / (a * a + t * t);
}

// Bilinear corners of Texture::pixelUV / pixelBi (texture.cpp:54-72) at (u, v): texel indices
// (-1 outside the image, which reads black) and weights, in the order of Texture::bilinear.
void texel_corners(const Texture& t, float u, float v, int idx[4], double w[4]) {
    const float fu = std::fmod(std::fmod(u, 1.0f) + 1.0f, 1.0f);
    const float fv = std::fmod(std::fmod(1.0f - v, 1.0f) + 1.0f, 1.0f);
    const float x = fu * t.w, y = fv * t.h;
    const int x0 = int(std::floor(x)), y0 = int(std::floor(y)), x1 = x0 + 1, y1 = y0 + 1;
    const int xs[4] = {x1, x1, x0, x0}, ys[4] = {y1, y0, y1, y0};
    const float ws[4] = {(x - x0) * (y - y0), (x - x0) * (y1 - y), (x1 - x) * (y - y0), (x1 - x) * (y1 - y)};
    for (int k = 0; k < 4; ++k) {
        const bool in = xs[k] >= 0 && ys[k] >= 0 && xs[k] < t.w && ys[k] < t.h;
        idx[k] = in ? ys[k] * t.w + xs[k] : -1;
        w[k] = ws[k];
    }
}

struct GradTracer {
    const Scene& sc;
    const BVH& bvh;
    RenderOpts opt;
    Tracer<double> tr;
    int P;        // material parameters; texel keys start here
    int tex_id;   // texture whose texels are differentiated (-1: none)

    // f and pdf, with derivatives for the differentiated models (Lambert, GGX-reflect); the
    // other models are replayed without derivatives
    static bool differentiated(const Material& M) { return M.type == MAT_LAMBERT || M.type == MAT_GGXREFL; }
    MatD parts(const Material& M, const MatEval<double>& me, const V3d& kd, const V3d& wo, const V3d& n,
               const V3d& wi, const V3d& tc) const {
        if (differentiated(M)) return mat_deriv(M, kd, wo, n, wi);
        MatD md;
        md.f = me.f(wo, n, wi, tc);
        md.pdf = me.pdf(wo, n, wi);
        return md;
    }

    // albedo derivative v (RGB) of material mi at texture coordinates (u, v)
    void add_albedo(SG& dL, const Material& M, int mi, float u, float vv, const V3d& v) const {
        if (!differentiated(M)) return;
        if (M.tex < 0) {
            dL.add(4 * mi + 0, V3d(v.x, 0, 0));
            dL.add(4 * mi + 1, V3d(0, v.y, 0));
            dL.add(4 * mi + 2, V3d(0, 0, v.z));
        } else if (M.tex == tex_id) {
            int idx[4];
            double w[4];
            texel_corners(sc.textures[M.tex], u, vv, idx, w);
            for (int k = 0; k < 4; ++k) {
                if (idx[k] < 0 || w[k] == 0) continue;
                dL.add(P + 3 * idx[k] + 0, V3d(v.x * w[k], 0, 0));
                dL.add(P + 3 * idx[k] + 1, V3d(0, v.y * w[k], 0));
                dL.add(P + 3 * idx[k] + 2, V3d(0, 0, v.z * w[k]));
            }
        }
    }

    V3d trace(const V3d& o, const V3d& d, double fac, int prev_mat, double dfac, int depth, const Stream& rng,
              SG& dL) const {
        double t, b1, b2;
        int ti;
        if (!tr.hit(o, d, t, b1, b2, ti)) return tr.env(d);
        const Tri& tri = sc.tris[ti];
        const Material& M = sc.mats[tri.mat];
        const V3d wo = -d;
        const V3d n = tri.normal<double>(b1, b2);
        if (M.type == MAT_LIGHT) {
            const V3d le(M.kd[0], M.kd[1], M.kd[2]);
            if (prev_mat >= 0) dL.add(4 * prev_mat + 3, le * dfac);
            return le * fac;
        }
        MatEval<double> me{&sc, &M};
        if (n.dot(d) > 0 && !me.transmissive()) return V3d(0);   // back faces of opaque models
        const V3d hp = o + d * t;
        const int mi = tri.mat;
        const V3d tc = tri.texc<double>(b1, b2);
        const float tu = float(tc.x), tv = float(tc.y);
        const V3d kd = me.albedo(tc);
        V3d result(0);
        double fac_next = 1, dfac_next = 0;
        float u[4];
        rng.draw(uint32_t(1 + depth), 0, u);
        if (opt.nee && me.needs_nee() && !sc.light_tris.empty()) {
            const Tri& L = sc.tris[sc.pick_light(u[0])];
            double r1 = u[1], r2 = u[2];
            if (r1 + r2 > 1) { r1 = 1 - r1; r2 = 1 - r2; }
            const V3d lp = (1 - r1 - r2) * V3d(L.p[0]) + r1 * V3d(L.p[1]) + r2 * V3d(L.p[2]);
            const V3d ln = L.normal<double>(r1, r2);
            const V3d lv = lp - hp;
            const V3d wl = lv.normalized();
            const Material& LM = sc.mats[L.mat];
            const V3d le(LM.kd[0], LM.kd[1], LM.kd[2]);
            const double lpdf = sc.light_p();
            const double lop = lpdf * lv.dot(lv) / std::max(1e-6, ln.dot(-wl));
            const MatD md = parts(M, me, kd, wo, n, wl, tc);
            double mis = lop / (lop + md.pdf);
            double dmis = -lop / ((lop + md.pdf) * (lop + md.pdf)) * md.dpdf_da;
            if (opt.mis == 0) mis = 0.5, dmis = 0;
            fac_next = 1 - mis;
            dfac_next = -dmis;
            double st, sb1, sb2;
            int sti;
            tr.hit(hp + wl * 1e-3, wl, st, sb1, sb2, sti);
            if (st + 2e-3 > lv.norm()) {
                const V3d Ll = le / lv.norm2() * std::max(0.0, ln.dot(-wl)) / lpdf;
                const double cn = std::max(0.0, n.dot((lp - hp).normalized()));
                result += Ll * md.f * cn * mis;
                const V3d k = Ll * cn;
                add_albedo(dL, M, mi, tu, tv, V3d(k.x * md.df_dkd.x * mis, k.y * md.df_dkd.y * mis, k.z * md.df_dkd.z * mis));
                if (M.type == MAT_GGXREFL) dL.add(4 * mi + 3, k * (md.df_da * mis + md.f * dmis));
            }
        }
        double prr = 1;
        if (depth >= opt.rrdepth) {
            if (u[3] > opt.prr) return result;
            prr = opt.prr;
        }
        if (opt.maxdepth > 0 && depth + 1 >= opt.maxdepth) return result;
        float v[4];
        rng.draw(uint32_t(1 + depth), 1, v);
        const V3d wi = me.sample(v, wo, n);
        const MatD md = parts(M, me, kd, wo, n, wi, tc);
        SG dLc;
        const V3d Lc = trace(hp + wi * 1e-3, wi, fac_next, mi, dfac_next, depth + 1, rng, dLc);
        const double cw = std::fabs(wi.dot(n));
        const V3d T = md.f * (cw / md.pdf / prr);
        result += Lc * T;
        dL.add_scaled(dLc, T);
        if (differentiated(M)) {
            const V3d dT = md.df_dkd * (cw / md.pdf / prr);
            add_albedo(dL, M, mi, tu, tv, V3d(dT.x * Lc.x, dT.y * Lc.y, dT.z * Lc.z));
        }
        if (M.type == MAT_GGXREFL) {
            const V3d dTa = (md.df_da * md.pdf - md.f * md.dpdf_da) * (cw / (md.pdf * md.pdf) / prr);
            const double sc_ = ggx_score(M.alpha, wo, n, wi);
            dL.add(4 * mi + 3, dTa * Lc + T * Lc * sc_);
        }
        return result;
    }
};

}  // namespace
}  // namespace orc

using namespace orc;

extern "C" {

// theta: n_mats rows {r, g, b, alpha}; emitters keep their Ke.
void orc_set_params(void* h, const float* theta, int n_mats) {
    Scene& sc = static_cast<OScene*>(h)->sc;
    for (int j = 0; j < n_mats && j < int(sc.mats.size()); ++j) {
        Material& m = sc.mats[j];
        if (m.type == MAT_LIGHT) continue;
        m.kd[0] = theta[4 * j]; m.kd[1] = theta[4 * j + 1]; m.kd[2] = theta[4 * j + 2];
        m.alpha = theta[4 * j + 3];
    }
}

// grad (n_mats x 4) of sum_p,c dldi[p,c] * mean_s max(0, L_s[p,c]), FP64; with tex_id >= 0 also
// grad_tex (w x h x 3) with respect to that texture's linear texel values.
int orc_adjoint(void* h, const float* cam, int W, int H, int spp, int spp0, uint32_t seed, float prr, int mis,
                int maxdepth, int rrdepth, const float* dldi, double* grad, int n_mats, int nthreads, int tex_id,
                double* grad_tex, int integrator) {
    OScene* o = static_cast<OScene*>(h);
    if (!o->built) return -1;
    Camera c;
    c.pos = V3f(cam[0], cam[1], cam[2]);
    c.from_euler(cam[3], cam[4], cam[5]);
    c.fov = cam[6];
    c.aspect = cam[7];
    RenderOpts opt;
    opt.prr = prr; opt.mis = mis; opt.maxdepth = maxdepth; opt.rrdepth = rrdepth; opt.seed = seed;
    opt.nee = integrator == 1 ? 0 : 1;
    const int P = 4 * n_mats;
    const int NT = (tex_id >= 0 && grad_tex) ? 3 * o->sc.textures[tex_id].w * o->sc.textures[tex_id].h : 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
    std::vector<double> acc(P + NT, 0.0);
#pragma omp parallel
    {
        std::vector<double> loc(P + NT, 0.0);
        GradTracer gt{o->sc, o->bvh, opt, Tracer<double>{o->sc, o->bvh, opt}, P, NT ? tex_id : -1};
#pragma omp for schedule(dynamic, 16)
        for (int px = 0; px < W * H; ++px) {
            const int x = px % W, y = px / W;
            for (int s = 0; s < spp; ++s) {
                Stream rng(seed, uint32_t(px), uint32_t(spp0 + s));
                float u[4];
                rng.draw(0, 0, u);
                V3f dir = c.ray(float(x) + u[0], float(y) + u[1], float(W), float(H));
                SG dL;
                V3d L = gt.trace(V3d(c.pos), V3d(dir), 1.0, -1, 0.0, 0, rng, dL);
                const double g[3] = {L.x > 0 ? dldi[3 * px] / spp : 0.0, L.y > 0 ? dldi[3 * px + 1] / spp : 0.0,
                                     L.z > 0 ? dldi[3 * px + 2] / spp : 0.0};
                for (const auto& e : dL.e)
                    if (e.first < P + NT) loc[e.first] += g[0] * e.second.x + g[1] * e.second.y + g[2] * e.second.z;
            }
        }
#pragma omp critical
        for (int k = 0; k < P + NT; ++k) acc[k] += loc[k];
    }
    for (int k = 0; k < P; ++k) grad[k] = acc[k];
    for (int k = 0; k < NT; ++k) grad_tex[k] = acc[P + k];
    return 0;
}

}  // extern "C"
