// Host scene assembly and upload.
#include "scene.h"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include "render.h"

namespace nr {

void check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

DevBuf::~DevBuf() {
    if (ptr) (void)hipFree(ptr);
}
void DevBuf::ensure(size_t n) {
    if (n <= bytes && ptr) return;
    if (ptr) check_hip(hipFree(ptr), "hipFree");
    ptr = nullptr;
    bytes = 0;
    check_hip(hipMalloc(&ptr, n ? n : 64), "hipMalloc");
    bytes = n;
}
void DevBuf::upload(const void* src, size_t n, hipStream_t s) {
    ensure(n);
    if (n) check_hip(hipMemcpyAsync(ptr, src, n, hipMemcpyHostToDevice, s), "hipMemcpyAsync H2D");
}

namespace {
struct V {
    float x, y, z;
};
inline V vsub(V a, V b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V vcross(V a, V b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline float vnorm2(V a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline V vnormalized(V a) {  // vec3::normalized, vec3.cpp:70-73
    float d = std::sqrt(vnorm2(a)) + 1e-8f;
    return {a.x / d, a.y / d, a.z / d};
}
inline V at(const float* q, int k) { return {q[3 * k], q[3 * k + 1], q[3 * k + 2]}; }
}  // namespace

Scene::Scene(const float* P, const float* N, const float* UV, const int32_t* M, uint32_t ntri, const float* MATS,
             uint32_t nmat)
    : p(P, P + 9 * size_t(ntri)), n(9 * size_t(ntri), 0.0f), uv(6 * size_t(ntri), 0.0f), mat(M, M + ntri),
      mats(MATS, MATS + 8 * size_t(nmat)) {
    mats_ext.assign(4 * size_t(nmat), 0.0f);
    for (uint32_t j = 0; j < nmat; ++j) mats_ext[4 * size_t(j) + 3] = -1.0f;
    if (N) std::memcpy(n.data(), N, sizeof(float) * 9 * size_t(ntri));
    if (UV) std::memcpy(uv.data(), UV, sizeof(float) * 6 * size_t(ntri));
    for (uint32_t i = 0; i < ntri; ++i) {
        if (M[i] < 0 || uint32_t(M[i]) >= nmat) throw std::runtime_error("material id out of range");
        // Triangle::Triangle: flat normal when |n0|^2 < 1e-4 (triangle.cpp:14-25)
        float* nn = &n[9 * size_t(i)];
        if (double(vnorm2(at(nn, 0))) < 1e-4) {
            const float* q = &p[9 * size_t(i)];
            V f = vnormalized(vcross(vsub(at(q, 1), at(q, 0)), vsub(at(q, 2), at(q, 0))));
            for (int k = 0; k < 3; ++k) { nn[3 * k] = f.x; nn[3 * k + 1] = f.y; nn[3 * k + 2] = f.z; }
        }
    }
    check_hip(hipGetDevice(&device), "hipGetDevice");
    check_hip(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
    check_hip(hipStreamCreateWithFlags(&stream2, hipStreamNonBlocking), "hipStreamCreate");
    check_hip(hipEventCreateWithFlags(&shadow_ready, hipEventDisableTiming), "hipEventCreate");
    check_hip(hipEventCreateWithFlags(&shadow_done, hipEventDisableTiming), "hipEventCreate");
}

Scene::~Scene() {
    delete target;
    if (shadow_ready) (void)hipEventDestroy(shadow_ready);
    if (shadow_done) (void)hipEventDestroy(shadow_done);
    if (stream2) (void)hipStreamDestroy(stream2);
    if (stream) (void)hipStreamDestroy(stream);
}

std::vector<MatRec> Scene::mat_records() const {
    std::vector<MatRec> mr(mats.size() / 8);
    for (size_t j = 0; j < mr.size(); ++j) {
        const float* m = &mats[8 * j];
        const float* e = &mats_ext[4 * j];
        mr[j].a = make_float4(m[0], m[1], m[2], m[3]);
        mr[j].b = make_float4(m[4], m[5], m[6], 0);
        mr[j].c = make_float4(e[0], e[1], e[2], e[3]);
    }
    return mr;
}

void Scene::set_material_ext(int id, const float* ext) {
    if (id < 0 || size_t(id) >= mats.size() / 8) throw std::runtime_error("material id out of range");
    for (int k = 0; k < 4; ++k) mats_ext[4 * size_t(id) + k] = ext[k];
    ++generation;
    if (built) upload_materials();
}

int Scene::add_texture(int id, int w, int h, int fmt, const void* data) {
    if (id < 0 || w <= 0 || h <= 0) throw std::runtime_error("bad texture");
    ++generation;
    if (int(textures.size()) <= id) textures.resize(id + 1);
    Texture& t = textures[id];
    // same size and format on a built scene: rewrite the texels in place, no rebuild
    const bool in_place = built && t.w == w && t.h == h && t.fmt == fmt && size_t(id) < tex_offsets.size();
    t.w = w; t.h = h; t.fmt = fmt;
    size_t bytes = size_t(w) * h * (fmt == 0 ? 4 : 12);
    t.bytes.assign(static_cast<const uint8_t*>(data), static_cast<const uint8_t*>(data) + bytes);
    if (in_place) {
        check_hip(hipMemcpyAsync(static_cast<uint8_t*>(d_texpool.ptr) + tex_offsets[id], t.bytes.data(), bytes,
                                 hipMemcpyHostToDevice, stream), "texture update");
        check_hip(hipStreamSynchronize(stream), "texture update sync");
        if (id == env_tex) build_env_sampler();
        return id;
    }
    if (built) upload_textures();  // a new or resized texture: new pool, the BVH stays
    return id;
}

// Every texture a material or the environment names must exist (kernels index without checks);
// the material feature bits select the shade kernels.
void Scene::check_materials() {
    const int N = n_tris();
    for (size_t j = 0; j < mats.size() / 8; ++j) {
        const int tex = int(mats[8 * j + 6]);
        if (tex >= 0 && (tex >= int(textures.size()) || textures[tex].w <= 0))
            throw std::runtime_error("material " + std::to_string(j) + " names missing texture " + std::to_string(tex));
        const int type = int(mats[8 * j]);
        if (type < MAT_LAMBERT || type > MAT_GLASS) throw std::runtime_error("unknown material type");
        const int tks = int(mats_ext[4 * j + 3]);
        if (tks >= 0 && (tks >= int(textures.size()) || textures[tks].w <= 0))
            throw std::runtime_error("material " + std::to_string(j) + " names missing Ks texture " + std::to_string(tks));
    }
    features = 0;
    for (int i = 0; i < N; ++i) {
        const float* m = &mats[8 * size_t(mat[i])];
        if (int(m[0]) == MAT_GGX) features |= 1u;
        if (int(m[6]) >= 0 || int(mats_ext[4 * size_t(mat[i]) + 3]) >= 0) features |= 2u;
        if (int(m[0]) == MAT_BLINNPHONG || int(m[0]) == MAT_GLASS) features |= 4u;
    }
    if (env_tex >= 0 && (env_tex >= int(textures.size()) || textures[env_tex].w <= 0))
        throw std::runtime_error("environment map missing");
}

void Scene::upload_materials(bool check) {
    if (check) check_materials();
    const std::vector<MatRec> mr = mat_records();
    d_mats.upload(mr.data(), mr.size() * sizeof(MatRec), stream);
    check_hip(hipStreamSynchronize(stream), "material upload");
    view.mats = d_mats.as<MatRec>();
}

// Texture pool (every texture 256-byte aligned), descriptors, the RGBA8 decode table and the sky
// sampler; points the view at them.
void Scene::upload_textures(bool check) {
    if (check) check_materials();
    std::vector<TexDesc> td(textures.size());
    size_t pool = 0;
    for (size_t j = 0; j < textures.size(); ++j) {
        td[j] = TexDesc{textures[j].w, textures[j].h, textures[j].fmt, 0, pool};
        pool += (textures[j].bytes.size() + 255) & ~size_t(255);
    }
    tex_offsets.resize(textures.size());
    for (size_t j = 0; j < textures.size(); ++j) tex_offsets[j] = td[j].offset;
    std::vector<uint8_t> poolh(pool);
    for (size_t j = 0; j < textures.size(); ++j)
        std::memcpy(poolh.data() + td[j].offset, textures[j].bytes.data(), textures[j].bytes.size());
    // texture.cpp:18-20 decode: pow(c / 255, 2.2f) in double, stored as float
    float lut[256];
    for (int c = 0; c < 256; ++c) lut[c] = float(std::pow(double(c) / 255.0, double(2.2f)));
    d_tex.upload(td.data(), td.size() * sizeof(TexDesc), stream);
    d_texpool.upload(poolh.data(), poolh.size(), stream);
    d_lut.upload(lut, sizeof(lut), stream);
    check_hip(hipStreamSynchronize(stream), "texture upload");
    view.tex = d_tex.as<TexDesc>();
    view.texpool = d_texpool.as<uint8_t>();
    view.lut = d_lut.as<float>();
    view.env_tex = env_tex;
    view.env_gain = env_gain;
    build_env_sampler();
}

void Scene::build(int nthreads) {
    ++generation;
    const int N = n_tris();
    check_materials();
    bvh = build_sah(p.data(), N, nthreads);

    // LightSampler::setPrimitives (lightsampler.cpp:4-39): emissive triangles in loader order,
    // weight = area() = |e1 x e2| (twice the area), CDF from 1e-6.
    lights.clear();
    light_cdf.clear();
    std::vector<float> w;
    float sum = 0.0f;
    for (int i = 0; i < N; ++i) {
        const float* m = &mats[8 * size_t(mat[i])];
        if (int(m[0]) != MAT_LIGHT) continue;
        const float* q = &p[9 * size_t(i)];
        const float* nn = &n[9 * size_t(i)];
        float a = std::sqrt(vnorm2(vcross(vsub(at(q, 1), at(q, 0)), vsub(at(q, 2), at(q, 0)))));
        w.push_back(a);
        sum += a;
        LightRec L;
        L.p0 = make_float4(q[0], q[1], q[2], 0); L.p1 = make_float4(q[3], q[4], q[5], 0);
        L.p2 = make_float4(q[6], q[7], q[8], 0);
        L.n0 = make_float4(nn[0], nn[1], nn[2], 0); L.n1 = make_float4(nn[3], nn[4], nn[5], 0);
        L.n2 = make_float4(nn[6], nn[7], nn[8], 0);
        L.ke = make_float4(m[1], m[2], m[3], 0);
        lights.push_back(L);
    }
    sum += 1e-8f;
    light_sum = sum;
    float acc = 1e-6f;
    for (float x : w) {
        acc += x / sum;
        light_cdf.push_back(acc);
    }

    // Geometry and shading records in DFS leaf order.
    std::vector<TriGeom> g(N);
    std::vector<TriShade> s(N);
    for (int k = 0; k < N; ++k) {
        int i = bvh.order[k];
        const float* q = &p[9 * size_t(i)];
        float e1[3] = {q[3] - q[0], q[4] - q[1], q[5] - q[2]};
        float e2[3] = {q[6] - q[0], q[7] - q[1], q[8] - q[2]};
        int32_t id = i;
        float idf;
        std::memcpy(&idf, &id, 4);
        g[k].g0 = make_float4(q[0], q[1], q[2], e1[0]);
        g[k].g1 = make_float4(e1[1], e1[2], e2[0], e2[1]);
        g[k].g2 = make_float4(e2[2], idf, 0, 0);
        const float* nn = &n[9 * size_t(i)];
        const float* t = &uv[6 * size_t(i)];
        int32_t m = mat[i];
        float mf;
        std::memcpy(&mf, &m, 4);
        s[k].s0 = make_float4(nn[0], nn[1], nn[2], nn[3]);
        s[k].s1 = make_float4(nn[4], nn[5], nn[6], nn[7]);
        s[k].s2 = make_float4(nn[8], t[0], t[1], t[2]);
        s[k].s3 = make_float4(t[3], t[4], t[5], mf);
    }
    d_wide.upload(bvh.wide.data(), bvh.wide.size() * sizeof(Node4), stream);
    d_flat.upload(bvh.flat.data(), bvh.flat.size() * sizeof(float4), stream);
    d_tris.upload(g.data(), g.size() * sizeof(TriGeom), stream);
    d_shade.upload(s.data(), s.size() * sizeof(TriShade), stream);
    d_lights.upload(lights.data(), lights.size() * sizeof(LightRec), stream);
    d_cdf.upload(light_cdf.data(), light_cdf.size() * sizeof(float), stream);
    check_hip(hipStreamSynchronize(stream), "upload sync");

    view = SceneView{};
    view.wide = d_wide.as<Node4>();
    view.flat = d_flat.as<float4>();
    view.n_flat = int(bvh.flat.size() / 2);
    view.tris = d_tris.as<TriGeom>();
    view.shade = d_shade.as<TriShade>();
    view.lights = d_lights.as<LightRec>();
    view.light_cdf = d_cdf.as<float>();
    for (int j = 0; j < 3; ++j) { view.root_min[j] = bvh.root_min[j]; view.root_max[j] = bvh.root_max[j]; }
    view.n_lights = int(lights.size());
    view.light_p = float(1.0 / (double(light_sum) + 1e-8));  // LightSampler::p, lightsampler.cpp:100-103
    view.short_stack = short_stack;
    view.wide_root = N ? bvh.wide_root : make_leaf(0, 0);
    view.wide_stack_cap = 3 * bvh.wide_depth + 2;
    view.lane_slots = view.wide_stack_cap + 3;
    upload_materials(false);  // checked above
    upload_textures(false);
    built = true;
}

// Environment sampler.  The lookup (envmap.cpp:17-24, texture.cpp:54-72) maps a direction to the
// continuous texel coordinate x = u W, y = (1 - v) H with u = atan2(z, x) / (2 pi') + 1/2 and
// 1 - v = acos(y) / pi' (pi' = 3.14159), and blends texels floor(x), floor(x) + 1 (a texel past
// the last one reads black).  Cell (i, j) = [j, j+1) x [i, i+1) therefore sees at most the largest
// of its four corner texels; its weight is that bound (largest channel, times |gain|) times
// sin(theta) at the cell centre, so that the density in solid angle follows the radiance.  A cell
// is drawn by Vose's alias method (built in double, in the order below), then a point uniformly
// inside it; the direction's solid-angle density is pdf_cell * W H / (2 pi'^2 sin theta).
void Scene::build_env_sampler() {
    env_tab.clear();
    view.env_tab = nullptr;
    view.env_w = view.env_h = 0;
    view.env_norm = 0.0f;
    if (env_tex < 0 || env_tex >= int(textures.size())) return;
    const Texture& t = textures[env_tex];
    const int W = t.w, H = t.h;
    if (W <= 0 || H <= 0) return;
    float lut[256];
    for (int c = 0; c < 256; ++c) lut[c] = float(std::pow(double(c) / 255.0, double(2.2f)));
    // largest channel magnitude of every texel (0 outside the map)
    std::vector<float> mx(size_t(W) * H);
    for (size_t k = 0; k < mx.size(); ++k) {
        float m = 0.0f;
        for (int c = 0; c < 3; ++c) {
            float v;
            if (t.fmt == 0) v = lut[t.bytes[4 * k + c]];
            else std::memcpy(&v, &t.bytes[12 * k + 4 * c], 4);
            m = std::max(m, std::fabs(v));
        }
        mx[k] = m;
    }
    auto at = [&](int x, int y) { return (x < W && y < H) ? mx[size_t(y) * W + x] : 0.0f; };
    const size_t N = size_t(W) * H;
    std::vector<double> w(N);
    double sum = 0.0;
    const double gain = std::fabs(double(env_gain));
    for (int i = 0; i < H; ++i) {
        const double st = std::sin((i + 0.5) / H * double(3.14159f));
        for (int j = 0; j < W; ++j) {
            const float c = std::max(std::max(at(j, i), at(j + 1, i)), std::max(at(j, i + 1), at(j + 1, i + 1)));
            const double x = double(c) * gain * st;
            w[size_t(i) * W + j] = x;
            sum += x;
        }
    }
    if (!(sum > 0.0) || !std::isfinite(sum)) return;   // black map: the sky is never sampled
    std::vector<double> sc(N);
    std::vector<float> prob(N, 1.0f);
    std::vector<int32_t> alias(N);
    std::vector<int32_t> small, large;
    for (size_t k = 0; k < N; ++k) {
        sc[k] = w[k] / sum * double(N);
        alias[k] = int32_t(k);
        (sc[k] < 1.0 ? small : large).push_back(int32_t(k));
    }
    while (!small.empty() && !large.empty()) {
        const int32_t l = small.back();
        small.pop_back();
        const int32_t g = large.back();
        large.pop_back();
        prob[l] = float(sc[l]);
        alias[l] = g;
        sc[g] = (sc[g] + sc[l]) - 1.0;
        (sc[g] < 1.0 ? small : large).push_back(g);
    }
    env_tab.resize(N);
    for (size_t k = 0; k < N; ++k) {
        int32_t a = alias[k];
        float af;
        std::memcpy(&af, &a, 4);
        env_tab[k] = make_float4(prob[k], af, float(w[k] / sum), float(w[size_t(a)] / sum));
    }
    d_env.upload(env_tab.data(), env_tab.size() * sizeof(float4), stream);
    check_hip(hipStreamSynchronize(stream), "env sampler upload");
    view.env_tab = d_env.as<float4>();
    view.env_w = W;
    view.env_h = H;
    view.env_norm = float(double(W) * H / (2.0 * double(3.14159f) * double(3.14159f)));
}

}  // namespace nr
