// Drives the oracle's C entry points (oracle/orc_*.cpp, compiled into this program with
// -fsanitize=address,undefined) over a small scene that uses every material type, a texture and
// a sky: SAH build, traversal (tree and brute force), FP32 and FP64 renders with both integrators
// and sky sampling, the adjoint with texel gradients, the radiance-cache network's forward and
// gradient.  Prints OK when every call returned normally; the sanitizers report anything else.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

extern "C" {
void* orc_scene_create(const float* p, const float* n, const float* uv, const int32_t* mat, int N, const float* mats,
                       int M);
void orc_scene_destroy(void* h);
void orc_scene_set_material_ext(void* h, int id, const float* ext);
void orc_scene_set_envmap(void* h, int w, int hh, const float* rgb, float gain);
void orc_scene_set_texture_rgba8(void* h, int id, int w, int hh, const uint8_t* rgba);
int orc_bvh_build(void* h);
void orc_intersect(void* h, int n, const float* o, const float* d, int32_t* tri, float* t, float* b1, float* b2,
                   int brute);
int orc_render(void* h, const float* cam, int W, int H, int x0, int y0, int x1, int y1, int spp0, int spp,
               uint32_t seed, float prr, int mis, int maxdepth, int rrdepth, int fp64, int nthreads, double* sum,
               double* sum2, int integrator, int env_is);
void orc_set_params(void* h, const float* theta, int n_mats);
int orc_adjoint(void* h, const float* cam, int W, int H, int spp, int spp0, uint32_t seed, float prr, int mis,
                int maxdepth, int rrdepth, const float* dldi, double* grad, int n_mats, int nthreads, int tex_id,
                double* gtex, int integrator);
double orc_nrc_grad(const float* w, const float* q, const float* t, int n, int bf16, double* grad);
void orc_nrc_forward(const float* w, const float* q, int n, int bf16, double* y);
}

int main() {
    std::mt19937 g(7);
    std::uniform_real_distribution<float> u(0.0f, 1.0f);
    // a box of 12 quads around random triangles: floor, walls, a ceiling light, six material types
    std::vector<float> p, uv;
    std::vector<int32_t> mat;
    auto quad = [&](float x0, float y0, float z0, float ax, float ay, float az, float bx, float by, float bz, int m) {
        const float c[4][3] = {{x0, y0, z0}, {x0 + ax, y0 + ay, z0 + az}, {x0 + ax + bx, y0 + ay + by, z0 + az + bz},
                               {x0 + bx, y0 + by, z0 + bz}};
        const int tri[2][3] = {{0, 1, 2}, {2, 3, 0}};
        for (auto& t : tri) {
            for (int k : t) p.insert(p.end(), c[k], c[k] + 3);
            const float tc[4][2] = {{0, 0}, {1, 0}, {1, 1}, {0, 1}};
            for (int k : t) uv.insert(uv.end(), tc[k], tc[k] + 2);
            mat.push_back(m);
        }
    };
    quad(-10, 0, -10, 20, 0, 0, 0, 0, 20, 0);     // floor, Lambert with a texture
    quad(-10, 0, -10, 0, 20, 0, 20, 0, 0, 2);     // back wall, GGX-reflect
    quad(-10, 0, -10, 0, 0, 20, 0, 20, 0, 4);     // left wall, Blinn-Phong
    quad(-2, 19.9f, -2, 4, 0, 0, 0, 0, 4, 1);     // light
    for (int i = 0; i < 60; ++i) {                // small triangles of every type
        const float c[3] = {u(g) * 16 - 8, u(g) * 12 + 1, u(g) * 16 - 8};
        for (int k = 0; k < 3; ++k)
            for (int a = 0; a < 3; ++a) p.push_back(c[a] + u(g) * 2 - 1);
        for (int k = 0; k < 6; ++k) uv.push_back(u(g));
        const int types[5] = {0, 2, 3, 4, 5};
        mat.push_back(types[i % 5]);   // material rows are in type order
    }
    const float mats[6][8] = {{0, .6f, .5f, .4f, .5f, 1, 0, 0},  {1, 8, 7, 6, .5f, 1, -1, 0},
                              {2, .8f, .7f, .6f, .3f, 1, -1, 0}, {3, .9f, .9f, .9f, .2f, 1.5f, -1, 0},
                              {4, .3f, .4f, .5f, 40, 1, -1, 0},  {5, .04f, .04f, .04f, .5f, 1.5f, -1, 0}};
    const int N = int(mat.size());
    void* s = orc_scene_create(p.data(), nullptr, uv.data(), mat.data(), N, &mats[0][0], 6);
    const float ext[4] = {.4f, .4f, .4f, -1};
    orc_scene_set_material_ext(s, 4, ext);
    std::vector<uint8_t> tex(8 * 8 * 4);
    for (auto& b : tex) b = uint8_t(g() & 255);
    orc_scene_set_texture_rgba8(s, 0, 8, 8, tex.data());
    std::vector<float> sky(32 * 16 * 3);
    for (auto& v : sky) v = u(g) * 2;
    orc_scene_set_envmap(s, 32, 16, sky.data(), 1.2f);
    if (orc_bvh_build(s) <= 0) return 1;
    // traversal: random rays, the tree against brute force
    const int R = 4096;
    std::vector<float> o(3 * R), d(3 * R), t(R), b1(R), b2(R), t2(R), c1(R), c2(R);
    std::vector<int32_t> tri(R), tri2(R);
    for (int i = 0; i < R; ++i) {
        float v[3] = {u(g) - .5f, u(g) - .5f, u(g) - .5f};
        const float l = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        for (int a = 0; a < 3; ++a) { o[3 * i + a] = u(g) * 18 - 9 + (a == 1 ? 10 : 0); d[3 * i + a] = v[a] / l; }
    }
    orc_intersect(s, R, o.data(), d.data(), tri.data(), t.data(), b1.data(), b2.data(), 0);
    orc_intersect(s, R, o.data(), d.data(), tri2.data(), t2.data(), c1.data(), c2.data(), 1);
    int differ = 0;
    for (int i = 0; i < R; ++i) differ += t[i] != t2[i];
    // renders: FP32 / FP64, ptnee / pt, sky sampling
    const float cam[8] = {0, 10, 30, 0, -10, 0, 50, 1};
    const int W = 24, H = 24;
    std::vector<double> sum(W * H * 3), sum2(W * H * 3);
    for (int fp64 = 0; fp64 < 2; ++fp64)
        for (int integ = 0; integ < 2; ++integ)
            orc_render(s, cam, W, H, 0, 0, W, H, 0, 2, 3, 0.8f, 1, 0, 0, fp64, 4, sum.data(), sum2.data(), integ, 0);
    orc_render(s, cam, W, H, 0, 0, W, H, 0, 2, 3, 0.8f, 1, 8, 2, 0, 4, sum.data(), sum2.data(), 0, 1);
    // adjoint with texel gradients
    std::vector<float> theta(6 * 4);
    for (int j = 0; j < 6; ++j)
        for (int k = 0; k < 4; ++k) theta[4 * j + k] = mats[j][1 + k];
    orc_set_params(s, theta.data(), 6);
    std::vector<float> dldi(W * H * 3);
    for (auto& v : dldi) v = u(g) - .5f;
    std::vector<double> grad(6 * 4), gtex(8 * 8 * 3);
    orc_adjoint(s, cam, W, H, 2, 0, 5, 0.8f, 1, 0, 0, dldi.data(), grad.data(), 6, 4, 0, gtex.data(), 0);
    orc_scene_destroy(s);
    // radiance-cache network
    const int NP = 5 * 4096 + 16 * 64, Q = 70;
    std::vector<float> w(NP), q(16 * Q), tg(3 * Q);
    for (auto& v : w) v = (u(g) - .5f) * .2f;
    for (int i = 0; i < Q; ++i)
        for (int k = 0; k < 16; ++k) q[16 * i + k] = k < 3 ? u(g) * 100 : u(g);
    for (auto& v : tg) v = u(g);
    std::vector<double> y(3 * Q), gw(NP);
    orc_nrc_forward(w.data(), q.data(), Q, 1, y.data());
    const double loss = orc_nrc_grad(w.data(), q.data(), tg.data(), Q, 0, gw.data());
    std::printf("OK %d triangles, %d tree/brute differences, loss %.4f\n", N, differ, loss);
    return differ == 0 ? 0 : 1;
}
