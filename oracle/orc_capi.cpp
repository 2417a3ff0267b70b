// ORACLE -- test infrastructure only (see orc_common.h).
//
// C entry points of liboracle.so, loaded by tests/ and bench.py through ctypes.
#include "orc_handle.h"
#include <omp.h>

using namespace orc;

namespace {
OScene* S(void* h) { return static_cast<OScene*>(h); }

void count_rec(const BVH& b, int i, const V3f& o, const V3f& d, TravStats& st) {
    const BVH::Node& nd = b.nodes[i];
    float mn[3] = {nd.mn.x, nd.mn.y, nd.mn.z}, mx[3] = {nd.mx.x, nd.mx.y, nd.mx.z};
    if (!slab_hit<float>(mn, mx, o, d)) return;
    if (nd.l < 0) { st.tri_tests += nd.prims.size(); return; }
    st.pair_fetch += 1;
    count_rec(b, nd.l, o, d, st);
    count_rec(b, nd.r, o, d, st);
}
}  // namespace

extern "C" {

// Triangles in loader order: p (N x 9), optional n (N x 9) and uv (N x 6), material id per
// triangle; materials as M rows of 8 floats {type, r, g, b, alpha, ior, texture, unused}.
void* orc_scene_create(const float* p, const float* n, const float* uv, const int32_t* mat, int N,
                       const float* mats, int M) {
    OScene* o = new OScene();
    o->sc.tris.resize(N);
    for (int i = 0; i < N; ++i) {
        Tri& t = o->sc.tris[i];
        for (int k = 0; k < 3; ++k) {
            t.p[k] = V3f(p[9 * i + 3 * k], p[9 * i + 3 * k + 1], p[9 * i + 3 * k + 2]);
            if (n) t.n[k] = V3f(n[9 * i + 3 * k], n[9 * i + 3 * k + 1], n[9 * i + 3 * k + 2]);
            if (uv) t.t[k] = V3f(uv[6 * i + 2 * k], uv[6 * i + 2 * k + 1], 0.0f);
        }
        t.mat = mat[i];
        t.fix_normal();
    }
    o->sc.mats.resize(M);
    for (int j = 0; j < M; ++j) {
        Material& m = o->sc.mats[j];
        m.type = int(mats[8 * j]);
        m.kd[0] = mats[8 * j + 1]; m.kd[1] = mats[8 * j + 2]; m.kd[2] = mats[8 * j + 3];
        m.alpha = mats[8 * j + 4];
        m.ior = mats[8 * j + 5];
        m.tex = int(mats[8 * j + 6]);
    }
    o->sc.build_light_sampler();
    return o;
}

void orc_scene_destroy(void* h) { delete S(h); }

// Blinn-Phong's second colour: ext = {Ks r, g, b, Ks texture or -1}.
void orc_scene_set_material_ext(void* h, int id, const float* ext) {
    Material& m = S(h)->sc.mats[id];
    m.ks[0] = ext[0]; m.ks[1] = ext[1]; m.ks[2] = ext[2];
    m.tex_ks = int(ext[3]);
}

void orc_scene_set_envmap(void* h, int w, int hh, const float* rgb, float gain) {
    Texture& e = S(h)->sc.env;
    e.w = w; e.h = hh;
    e.rgb.assign(rgb, rgb + size_t(w) * hh * 3);
    S(h)->sc.env_gain = gain;
    S(h)->sc.has_env = true;
    S(h)->sc.env_sampler.build(e, gain);
}

// RGBA8 texture, decoded like Texture::load (texture.cpp:5-31).
void orc_scene_set_texture_rgba8(void* h, int id, int w, int hh, const uint8_t* rgba) {
    auto& v = S(h)->sc.textures;
    if (int(v.size()) <= id) v.resize(id + 1);
    Texture& t = v[id];
    t.w = w; t.h = hh;
    t.rgb.resize(size_t(w) * hh * 3);
    for (size_t i = 0; i < size_t(w) * hh; ++i)
        for (int c = 0; c < 3; ++c) t.rgb[3 * i + c] = srgb8_decode(rgba[4 * i + c]);
}

// Linear RGB float texture (3 floats per texel, first row at v = 1).
void orc_scene_set_texture_rgb32f(void* h, int id, int w, int hh, const float* rgb) {
    auto& v = S(h)->sc.textures;
    if (int(v.size()) <= id) v.resize(id + 1);
    Texture& t = v[id];
    t.w = w; t.h = hh;
    t.rgb.assign(rgb, rgb + size_t(w) * hh * 3);
}

int orc_bvh_build(void* h) {
    S(h)->bvh.build(S(h)->sc);
    if (S(h)->bvh.degenerate) return -1;
    S(h)->built = true;
    return int(S(h)->bvh.nodes.size());
}

int orc_bvh_depth(void* h) { return S(h)->bvh.depth(); }

// Canonical preorder export: node i has children left[i], right[i] (preorder ids, -1 for a leaf),
// box[6*i..] = mn, mx; leaves list their primitives at prim_order[leaf_first[i] .. +leaf_count[i]],
// concatenated in DFS (left-first) order.
static int export_rec(const BVH& b, int node, int& next, int& pcount, int32_t* left, int32_t* right,
                      float* box, int32_t* leaf_first, int32_t* leaf_count, int32_t* prim_order) {
    const BVH::Node& nd = b.nodes[node];
    int id = next++;
    box[6 * id + 0] = nd.mn.x; box[6 * id + 1] = nd.mn.y; box[6 * id + 2] = nd.mn.z;
    box[6 * id + 3] = nd.mx.x; box[6 * id + 4] = nd.mx.y; box[6 * id + 5] = nd.mx.z;
    if (nd.l < 0) {
        left[id] = right[id] = -1;
        leaf_first[id] = pcount;
        leaf_count[id] = int(nd.prims.size());
        for (int p : nd.prims) prim_order[pcount++] = p;
        return id;
    }
    leaf_first[id] = -1;
    leaf_count[id] = 0;
    left[id] = export_rec(b, nd.l, next, pcount, left, right, box, leaf_first, leaf_count, prim_order);
    right[id] = export_rec(b, nd.r, next, pcount, left, right, box, leaf_first, leaf_count, prim_order);
    return id;
}

int orc_bvh_export(void* h, int32_t* left, int32_t* right, float* box, int32_t* leaf_first,
                   int32_t* leaf_count, int32_t* prim_order) {
    const BVH& b = S(h)->bvh;
    if (b.root < 0) return 0;
    int next = 0, pcount = 0;
    export_rec(b, b.root, next, pcount, left, right, box, leaf_first, leaf_count, prim_order);
    return next;
}

void orc_intersect(void* h, int n, const float* o, const float* d, int32_t* tri, float* t, float* b1,
                   float* b2, int brute) {
    const BVH& b = S(h)->bvh;
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
        V3f O(o[3 * i], o[3 * i + 1], o[3 * i + 2]), D(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        Hit r = brute ? b.brute(O, D) : b.intersect(O, D);
        tri[i] = r.tri; t[i] = r.t; b1[i] = r.b1; b2[i] = r.b2;
    }
}

// Node-pair fetches and triangle tests of the reference traversal over a ray batch.
void orc_trav_stats(void* h, int n, const float* o, const float* d, uint64_t* out) {
    const BVH& b = S(h)->bvh;
    uint64_t pf = 0, tt = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : pf, tt)
    for (int i = 0; i < n; ++i) {
        TravStats st;
        V3f O(o[3 * i], o[3 * i + 1], o[3 * i + 2]), D(d[3 * i], d[3 * i + 1], d[3 * i + 2]);
        if (b.root >= 0) count_rec(b, b.root, O, D, st);
        pf += st.pair_fetch; tt += st.tri_tests;
    }
    out[0] = pf; out[1] = tt;
}

// cam = {pos.xyz, yaw, pitch, roll, fov, aspect}.  Renders pixels [x0,x1)x[y0,y1), samples
// [spp0, spp0+spp), accumulating per-pixel sums (W*H*3 doubles, full-image layout).
// integrator: 0 = RendererPTLS (ptnee), 1 = RendererPT (pt, rendererpt.cpp:13-61).
int orc_render(void* h, const float* cam, int W, int H, int x0, int y0, int x1, int y1, int spp0,
               int spp, uint32_t seed, float prr, int mis, int maxdepth, int rrdepth, int fp64,
               int nthreads, double* sum, double* sum2, int integrator, int env_is) {
    OScene* o = S(h);
    if (!o->built) return -1;
    Camera c;
    c.pos = V3f(cam[0], cam[1], cam[2]);
    c.from_euler(cam[3], cam[4], cam[5]);
    c.fov = cam[6];
    c.aspect = cam[7];
    RenderOpts opt;
    opt.prr = prr; opt.mis = mis; opt.maxdepth = maxdepth; opt.rrdepth = rrdepth; opt.seed = seed;
    opt.nee = integrator == 1 ? 0 : 1;
    opt.env_is = env_is;
    const int B = 8;
    int tx = (x1 - x0 + B - 1) / B, ty = (y1 - y0 + B - 1) / B;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
    for (int tile = 0; tile < tx * ty; ++tile) {
        int bx = x0 + (tile % tx) * B, by = y0 + (tile / tx) * B;
        for (int y = by; y < std::min(by + B, y1); ++y)
            for (int x = bx; x < std::min(bx + B, x1); ++x) {
                double s[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
                for (int k = 0; k < spp; ++k) {
                    double L[3];
                    if (fp64) {
                        Tracer<double> tr{o->sc, o->bvh, opt};
                        V3d v = tr.sample(c, x, y, W, H, uint32_t(spp0 + k));
                        L[0] = v.x; L[1] = v.y; L[2] = v.z;
                    } else {
                        Tracer<float> tr{o->sc, o->bvh, opt};
                        V3f v = tr.sample(c, x, y, W, H, uint32_t(spp0 + k));
                        L[0] = v.x; L[1] = v.y; L[2] = v.z;
                    }
                    for (int ch = 0; ch < 3; ++ch) { s[ch] += L[ch]; s2[ch] += L[ch] * L[ch]; }
                }
                size_t px = size_t(y) * W + x;
                for (int ch = 0; ch < 3; ++ch) { sum[3 * px + ch] += s[ch]; sum2[3 * px + ch] += s2[ch]; }
            }
    }
    return 0;
}

// ---- small probes used by the oracle's own pinning tests
int orc_slab(const float* mn, const float* mx, const float* o, const float* d) {
    return slab_hit<float>(mn, mx, V3f(o[0], o[1], o[2]), V3f(d[0], d[1], d[2])) ? 1 : 0;
}
void orc_tri_test(const float* p, const float* o, const float* d, float* out3) {
    V3f p0(p[0], p[1], p[2]), p1(p[3], p[4], p[5]), p2(p[6], p[7], p[8]);
    tri_test_f(p0, p1 - p0, p2 - p0, V3f(o[0], o[1], o[2]), V3f(d[0], d[1], d[2]), out3[0], out3[1], out3[2]);
}
void orc_philox(const uint32_t* ctr, const uint32_t* key, uint32_t* out) { philox4x32_10(ctr, key, out); }
float orc_srgb8_decode(int c) { return srgb8_decode(c); }

void orc_camera_rays(const float* cam, int n, const float* px, const float* py, float W, float H, float* dirs) {
    Camera c;
    c.pos = V3f(cam[0], cam[1], cam[2]);
    c.from_euler(cam[3], cam[4], cam[5]);
    c.fov = cam[6];
    c.aspect = cam[7];
    for (int i = 0; i < n; ++i) {
        V3f d = c.ray(px[i], py[i], W, H);
        dirs[3 * i] = d.x; dirs[3 * i + 1] = d.y; dirs[3 * i + 2] = d.z;
    }
}

void orc_texture_lookup(void* h, int id, int n, const float* u, const float* v, float* out) {
    const Texture& t = id < 0 ? S(h)->sc.env : S(h)->sc.textures[id];
    for (int i = 0; i < n; ++i) {
        V3f c = t.uv(u[i], v[i]);
        out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
    }
}

// Material probes: for material `m`, evaluate f and pdf (mode 0) or sample wi then pdf/f (mode 1)
// for n directions (wo, n, wi / u3) at FP32 (fp64=0) or FP64.
void orc_material_eval(void* h, int m, int n, const float* wo, const float* nn, const float* wi,
                       const float* u3, int mode, float* wi_out, float* f_out, float* pdf_out) {
    OScene* o = S(h);
    MatEval<float> me{&o->sc, &o->sc.mats[m]};
    for (int i = 0; i < n; ++i) {
        V3f WO(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), N(nn[3 * i], nn[3 * i + 1], nn[3 * i + 2]);
        V3f WI;
        if (mode == 1) {
            float u[3] = {u3[3 * i], u3[3 * i + 1], u3[3 * i + 2]};
            WI = me.sample(u, WO, N);
        } else {
            WI = V3f(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]);
        }
        V3f f = me.f(WO, N, WI, V3f(0.0f));
        wi_out[3 * i] = WI.x; wi_out[3 * i + 1] = WI.y; wi_out[3 * i + 2] = WI.z;
        f_out[3 * i] = f.x; f_out[3 * i + 1] = f.y; f_out[3 * i + 2] = f.z;
        pdf_out[i] = me.pdf(WO, N, WI);
    }
}

// Sky sampler of the optional environment sampling: n directions from n x 4 uniforms (FP32 mode),
// ok[i] = 0 for a rejected (polar) sample; and the density at n given directions.
void orc_env_sample(void* h, int n, const float* u4, float* d_out, float* pdf_out, int* ok) {
    const EnvSampler& es = S(h)->sc.env_sampler;
    for (int i = 0; i < n; ++i) {
        V3f d(0.0f);
        float p = 0.0f;
        ok[i] = es.ready() && es.sample(u4 + 4 * i, d, p) ? 1 : 0;
        d_out[3 * i] = d.x; d_out[3 * i + 1] = d.y; d_out[3 * i + 2] = d.z;
        pdf_out[i] = p;
    }
}
void orc_env_density(void* h, int n, const double* d, double* out) {
    const EnvSampler& es = S(h)->sc.env_sampler;
    for (int i = 0; i < n; ++i) out[i] = es.ready() ? es.density(V3d(d[3 * i], d[3 * i + 1], d[3 * i + 2])) : 0.0;
}
// True density of material m's sampler at wi (FP64).
void orc_material_sample_density(void* h, int m, int n, const double* wo, const double* nn, const double* wi,
                                 double* out) {
    OScene* o = S(h);
    MatEval<double> me{&o->sc, &o->sc.mats[m]};
    for (int i = 0; i < n; ++i)
        out[i] = me.sample_density(V3d(wo[3 * i], wo[3 * i + 1], wo[3 * i + 2]), V3d(nn[3 * i], nn[3 * i + 1], nn[3 * i + 2]),
                                   V3d(wi[3 * i], wi[3 * i + 1], wi[3 * i + 2]));
}

// Radiance the estimator returns for a ray that leaves the scene along d (Tracer::env,
// rendererptls.cpp:18-28): n unit directions in, n x 3 RGB out (FP32 mode).
void orc_env_radiance(void* h, int n, const float* d, float* out) {
    Tracer<float> tr{S(h)->sc, S(h)->bvh, RenderOpts{}};
    for (int i = 0; i < n; ++i) {
        V3f c = tr.env(V3f(d[3 * i], d[3 * i + 1], d[3 * i + 2]));
        out[3 * i] = c.x; out[3 * i + 1] = c.y; out[3 * i + 2] = c.z;
    }
}

int orc_light_pick(void* h, float u) { return S(h)->sc.pick_light(u); }
float orc_light_p(void* h) { return S(h)->sc.light_p(); }

}  // extern "C"
