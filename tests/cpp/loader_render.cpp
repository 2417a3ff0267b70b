// Loads a scene with the C++ facade's Loader (include/nuray.hpp), writes what the loader produced
// (triangles, material rows, textures, the optional sky) and renders it with the facade's
// Renderer on the GPU, appending the render's FP64 per-pixel sums, so that a test can render the
// very same triangles with the oracle.
//
//   loader_render <scene description> <out.bin> <config k=v ...> <camera px,py,pz,yaw,pitch,roll,fov,asp> [sky]
//
// out.bin (little endian): int32 nt, nm, ntex, W, H, has_sky; nt x {p[9], n[9], uv[6] float, int32 mat};
// nm x {row[8], ext[4] float}; ntex x {int32 w, h, w*h*4 RGBA8}; [int32 sw, sh, float gain, sw*sh*3 float];
// W*H*3 double sum, W*H*3 double sum2.
#include <cstdio>
#include <fstream>
#include <sstream>
#include "../../include/nuray.hpp"

template <class T>
static void put(std::ofstream& os, const T& v) {
    os.write(reinterpret_cast<const char*>(&v), sizeof(T));
}

int main(int argc, char** argv) {
    if (argc < 5) return 2;
    try {
        nuray::Loader l;
        l.fromSceneDescription(argv[1]);
        const nuray::Config cfg(argv[3]);
        float c[8];
        std::stringstream cs(argv[4]);
        for (int i = 0; i < 8; ++i) {
            std::string tok;
            std::getline(cs, tok, ',');
            c[i] = float(std::atof(tok.c_str()));
        }
        const nuray::Camera cam(c[6], c[7], nuray::vec3(c[0], c[1], c[2]), c[3], c[4], c[5]);
        const bool sky = argc > 5 && std::string(argv[5]) == "sky";
        nuray::Envmap env;
        if (sky) {  // a smooth sky with a bright patch, linear RGB, first row at the zenith
            env.w = 64;
            env.h = 32;
            env.gain = 1.5f;
            env.rgb.resize(size_t(env.w) * env.h * 3);
            for (int y = 0; y < env.h; ++y)
                for (int x = 0; x < env.w; ++x) {
                    const float v = 1.0f - float(y) / float(env.h - 1);
                    const bool sun = x >= 40 && x < 44 && y >= 6 && y < 9;
                    float* q = &env.rgb[(size_t(y) * env.w + x) * 3];
                    q[0] = 0.2f + 0.5f * v + (sun ? 20.0f : 0.0f);
                    q[1] = 0.3f + 0.5f * v + (sun ? 18.0f : 0.0f);
                    q[2] = 0.5f + 0.4f * v + (sun ? 14.0f : 0.0f);
                }
        }
        nuray::Renderer r(cfg.getValueStr("renderer", "").empty() ? "ptnee" : cfg.getValueStr("renderer", ""));
        if (sky) r.setEnvmap(env);
        r.prepare(l);
        nuray::Image img;
        r.render(cam, img, cfg);
        const int W = img.w, H = img.h;
        std::vector<double> sum(size_t(W) * H * 3), sum2(sum.size());
        nuray::check(nr_render_fetch(r.handle(), nullptr, sum.data(), sum2.data(), nullptr));

        const auto& tris = l.getTriangles();
        const auto& mats = l.materials();
        const auto& texs = l.textures();
        std::ofstream os(argv[2], std::ios::binary);
        put(os, int32_t(tris.size()));
        put(os, int32_t(mats.size()));
        put(os, int32_t(texs.size()));
        put(os, int32_t(W));
        put(os, int32_t(H));
        put(os, int32_t(sky));
        for (const auto& t : tris) {
            for (int k = 0; k < 3; ++k) os.write(reinterpret_cast<const char*>(t.p[k].v), 12);
            for (int k = 0; k < 3; ++k) os.write(reinterpret_cast<const char*>(t.n[k].v), 12);
            for (int k = 0; k < 3; ++k) os.write(reinterpret_cast<const char*>(t.t[k].v), 8);
            put(os, int32_t(t.mat));
        }
        for (const auto& m : mats) {
            const float row[12] = {float(m.type), m.kd[0], m.kd[1], m.kd[2], m.alpha, m.ior, float(m.tex), 0.0f,
                                   m.ks[0], m.ks[1], m.ks[2], float(m.tex_ks)};
            os.write(reinterpret_cast<const char*>(row), sizeof(row));
        }
        for (const auto& t : texs) {
            put(os, int32_t(t.w));
            put(os, int32_t(t.h));
            os.write(reinterpret_cast<const char*>(t.rgba.data()), std::streamsize(t.rgba.size()));
        }
        if (sky) {
            put(os, int32_t(env.w));
            put(os, int32_t(env.h));
            put(os, env.gain);
            os.write(reinterpret_cast<const char*>(env.rgb.data()), std::streamsize(env.rgb.size() * 4));
        }
        os.write(reinterpret_cast<const char*>(sum.data()), std::streamsize(sum.size() * 8));
        os.write(reinterpret_cast<const char*>(sum2.data()), std::streamsize(sum2.size() * 8));
        std::printf("OK %zu triangles %zu materials %zu textures %dx%d\n", tris.size(), mats.size(), texs.size(), W, H);
        return os ? 0 : 1;
    } catch (const std::exception& e) {
        std::printf("FAILED %s\n", e.what());
        return 1;
    }
}
