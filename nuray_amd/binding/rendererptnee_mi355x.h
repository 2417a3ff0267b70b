// RendererPTNEE_MI355X: the reference's renderer interface (src/renderer/renderer.h:15-35) served
// by libnuray.so on one MI355X.
//
// Reference-side code: a maintainer copies this header into the reference tree (it includes the
// reference's headers by their src/-relative paths, so INCLUDEPATH += src and += <this repo>/include),
// links -lnuray, and adds one factory line (cmd/cmd.cpp:69-95, gui/widget.cpp):
//     if (name == "ptnee_gpu") renderer = new RendererPTNEE_MI355X(0);
//     if (name == "pt_gpu")    renderer = new RendererPTNEE_MI355X(1);
//
// What it maps:
//   prepare(triangles)  triangles in loader order, vertex normals and uv, one material row per
//                       distinct Material (classified by dynamic_cast over MatLambert, MatLight,
//                       MatGGXRefl, MatGGX, MatBlinnPhong, MatGlass), every map_Kd_ / map_Ks_ /
//                       map_F0_ texture uploaded as linear float texels (Texture::load already
//                       applied the 2.2 decode) -> nr_scene_create + textures + nr_bvh_build_sah
//   render(...)         Config keys imgw, imgh, spp, prr, mis (+ seed, maxdepth, rrdepth), the
//                       camera as position + Euler angles (Camera::toEuler), the sky from the same
//                       Config keys the command line loads it from (envmap, envmap_gain; cmd.cpp:59-65),
//                       progress in percent as RendererPTLS reports it (rendererptls.cpp:121), con_flag
//                       watched by a thread that raises the render's cancel word, the result written
//                       into the QImage under the framebuffer mutex with the reference's RGB888 transform
//   setEnvmap(env)      the sky's texels are private to Envmap, so they are taken at render() from the
//                       envmap / envmap_gain keys (the Qt front-end sets them on its Config)
// Errors follow the reference's style: printed to std::cerr, the render returns.
#pragma once
#include <atomic>
#include <chrono>
#include <cstdint>
#include <functional>
#include <iostream>
#include <map>
#include <string>
#include <thread>
#include <vector>

#include <QImage>
#include <QMutex>

#include "material/matblinnphong.h"
#include "material/matggx.h"
#include "material/matggxrefl.h"
#include "material/matglass.h"
#include "material/matlambert.h"
#include "material/matlight.h"
#include "renderer/renderer.h"
#include "scene/envmap.h"
#include "scene/texture.h"
#include "utils/config.h"

#include "nuray.h"

class RendererPTNEE_MI355X : public Renderer {
public:
    explicit RendererPTNEE_MI355X(int integrator = 0) : integrator_(integrator) {}
    // (Renderer has no virtual destructor: the scene is freed when the object itself is deleted)
    ~RendererPTNEE_MI355X() {
        if (scene_) nr_scene_destroy(scene_);
    }

    // Renderer::prepare (renderer.h:21, renderer.cpp:30-39): the scene goes to the device once.
    void prepare(const std::vector<Triangle>& triangles) override {
        if (scene_) nr_scene_destroy(scene_);
        scene_ = nullptr;
        env_key_.clear();
        if (!ready_device()) return;
        std::vector<float> p, n, uv, rows;
        std::vector<int32_t> mat_id;
        std::map<const Material*, int> row_of;
        std::vector<const Material*> order;
        p.reserve(triangles.size() * 9);
        n.reserve(triangles.size() * 9);
        uv.reserve(triangles.size() * 6);
        for (const Triangle& t : triangles) {
            for (int k = 0; k < 3; ++k) {
                for (int c = 0; c < 3; ++c) p.push_back(t.p[k][c]);
                for (int c = 0; c < 3; ++c) n.push_back(t.n[k][c]);
                uv.push_back(t.t[k][0]);
                uv.push_back(t.t[k][1]);
            }
            auto it = row_of.find(t.mat);
            if (it == row_of.end()) {
                it = row_of.emplace(t.mat, int(order.size())).first;
                order.push_back(t.mat);
            }
            mat_id.push_back(it->second);
        }
        std::vector<TexUpload> tex;
        std::vector<std::pair<int, std::vector<float>>> ext;  // Blinn-Phong Ks rows
        for (size_t j = 0; j < order.size(); ++j) {
            float row[8] = {0, 0, 0, 0, 0.5f, 1.0f, -1.0f, 0};
            if (!material_row(order[j], int(j), row, tex, ext)) {
                std::cerr << "RendererPTNEE_MI355X: unsupported material type" << std::endl;
                return;
            }
            rows.insert(rows.end(), row, row + 8);
        }
        scene_ = nr_scene_create(p.data(), n.data(), uv.data(), mat_id.data(), uint32_t(triangles.size()),
                                 rows.data(), uint32_t(order.size()));
        if (!scene_) {
            std::cerr << "RendererPTNEE_MI355X: " << nr_last_error() << std::endl;
            return;
        }
        bool ok = true;
        for (const TexUpload& u : tex)
            ok = ok && nr_scene_set_texture_rgb32f(scene_, u.id, u.w, u.h, u.rgb.data()) == NR_OK;
        for (const auto& e : ext) ok = ok && nr_scene_set_material_ext(scene_, e.first, e.second.data()) == NR_OK;
        ok = ok && nr_bvh_build_sah(scene_, 0) == NR_OK;
        if (!ok) {
            std::cerr << "RendererPTNEE_MI355X: " << nr_last_error() << std::endl;
            nr_scene_destroy(scene_);
            scene_ = nullptr;
        }
    }

    // Renderer::setEnvmap (renderer.h:30).  Envmap keeps its texture and gain private; the sky is
    // read at render() from the Config keys the command line builds its Envmap from.
    void setEnvmap(const Envmap& env_map) override { Renderer::setEnvmap(env_map); }

    // Renderer::render (renderer.h:22-29), as RendererPTLS::render (rendererptls.cpp:96-236).
    void render(const Camera& camera, const std::vector<Triangle>& triangles, QImage& img, Config& config,
                std::function<void(bool)> display_cb, std::atomic<int>& con_flag,
                std::function<void(float)> progress_cb, QMutex& framebuffer_mutex,
                const Envmap* env_map = nullptr) override {
        (void)triangles;  // the scene of the last prepare()
        if (!scene_) {
            std::cerr << "RendererPTNEE_MI355X: render before a successful prepare" << std::endl;
            return;
        }
        nr_params p{};
        p.width = config.getValueInt("imgw", 1);
        p.height = config.getValueInt("imgh", 1);
        p.spp = config.getValueInt("spp", 1);
        p.prr = config.getValueFloat("prr", 0.8f);
        p.mis = config.getValueInt("mis", 1);
        p.seed = uint32_t(config.getValueInt("seed", 0));
        p.maxdepth = config.getValueInt("maxdepth", 0);
        p.rrdepth = config.getValueInt("rrdepth", 0);
        p.integrator = integrator_;
        p.path_kernel = 2;  // automatic: one lane per path for small scenes and renders
        if (env_map && !upload_sky(config)) return;
        const vec3 e = camera.toEuler();
        nr_camera c{{camera.pos[0], camera.pos[1], camera.pos[2]}, {e[0], e[1], e[2]}, camera.fov_h, camera.aspect};
        display_cb(false);
        std::vector<uint8_t> rgb(size_t(p.width) * size_t(p.height) * 3);
        volatile int32_t cancel = 0;
        std::atomic<bool> done{false};
        // con_flag = 0 from any thread cancels (rendererptls.cpp:150,215,226-230)
        std::thread watcher([&] {
            while (!done.load()) {
                if (con_flag.load() == 0) {
                    cancel = 1;
                    return;
                }
                std::this_thread::sleep_for(std::chrono::milliseconds(2));
            }
        });
        auto progress = [](float fraction, void* user) {
            (*static_cast<std::function<void(float)>*>(user))(fraction * 100.0f);
        };
        const int rc = nr_render_ptnee(scene_, &c, &p, nullptr, nullptr, nullptr, rgb.data(), progress, &progress_cb,
                                       &cancel);
        done.store(true);
        watcher.join();
        if (rc < 0) {
            std::cerr << "RendererPTNEE_MI355X: " << nr_last_error() << std::endl;
            return;
        }
        if (rc == NR_CANCELLED) {
            std::cout << "Cancelled" << std::endl;
            return;
        }
        framebuffer_mutex.lock();
        img = QImage(QSize(p.width, p.height), QImage::Format_RGB888);
        for (int y = 0; y < p.height; ++y)
            for (int x = 0; x < p.width; ++x) {
                const uint8_t* q = &rgb[(size_t(y) * p.width + x) * 3];
                img.setPixel(x, y, qRgb(q[0], q[1], q[2]));
            }
        framebuffer_mutex.unlock();
        con_flag = 0;
        display_cb(true);
    }

private:
    struct TexUpload {
        int id, w, h;
        std::vector<float> rgb;
    };

    nr_scene* scene_ = nullptr;
    int integrator_ = 0;
    std::string env_key_;  // path and gain of the uploaded sky

    static bool ready_device() {
        static int state = -1;
        if (state < 0) {
            state = nr_init(0) == NR_OK ? 1 : 0;
            if (!state) std::cerr << "RendererPTNEE_MI355X: " << nr_last_error() << std::endl;
        }
        return state == 1;
    }

    // Texture keeps width and height private and exposes width(); the height is the first row
    // whose (non-const) pixel(0, y) throws "out of bound" (texture.cpp:47-51).
    static int texture_height(const Texture& t) {
        Texture& m = const_cast<Texture&>(t);  // pixel() only reads
        auto valid = [&m](int y) {
            try {
                (void)m.pixel(0, y);
                return true;
            } catch (const char*) {
                return false;
            }
        };
        if (t.width() <= 0 || !valid(0)) return 0;
        int lo = 0, hi = 1;  // valid(lo), and hi is tested next
        while (valid(hi)) {
            lo = hi;
            hi *= 2;
        }
        while (hi - lo > 1) {
            const int mid = lo + (hi - lo) / 2;
            (valid(mid) ? lo : hi) = mid;
        }
        return lo + 1;
    }

    static TexUpload texels(const Texture& t, int id) {
        TexUpload u{id, t.width(), texture_height(t), {}};
        u.rgb.resize(size_t(u.w) * size_t(u.h) * 3);
        for (int y = 0; y < u.h; ++y)
            for (int x = 0; x < u.w; ++x) {
                const vec3 c = t.pixel(x, y);
                for (int k = 0; k < 3; ++k) u.rgb[(size_t(y) * u.w + x) * 3 + k] = c[k];
            }
        return u;
    }

    // One material row {type, r, g, b, alpha, ior, texture, 0} (include/nuray.h).
    static bool material_row(const Material* m, int row_id, float* row, std::vector<TexUpload>& tex,
                             std::vector<std::pair<int, std::vector<float>>>& ext) {
        auto set_rgb = [row](const vec3& c) {
            row[1] = c[0];
            row[2] = c[1];
            row[3] = c[2];
        };
        auto add_tex = [&tex](const Texture& t) {
            tex.push_back(texels(t, int(tex.size())));
            return float(tex.back().id);
        };
        if (const auto* l = dynamic_cast<const MatLight*>(m)) {
            row[0] = NR_MAT_LIGHT;
            set_rgb(l->Ke_);
        } else if (const auto* d = dynamic_cast<const MatLambert*>(m)) {
            row[0] = NR_MAT_LAMBERT;
            set_rgb(d->Kd_);
            if (d->usetex_Kd_) row[6] = add_tex(d->map_Kd_);
        } else if (const auto* g = dynamic_cast<const MatGGXRefl*>(m)) {
            row[0] = NR_MAT_GGXREFL;
            set_rgb(g->Kd_);
            row[4] = g->alpha_;
            if (g->usetex_Kd_) row[6] = add_tex(g->map_Kd_);
        } else if (const auto* t = dynamic_cast<const MatGGX*>(m)) {
            row[0] = NR_MAT_GGX;
            set_rgb(t->Kd_);
            row[4] = t->alpha_;
            row[5] = t->ior_;
            if (t->usetex_Kd_) row[6] = add_tex(t->map_Kd_);
        } else if (const auto* b = dynamic_cast<const MatBlinnPhong*>(m)) {
            row[0] = NR_MAT_BLINNPHONG;
            set_rgb(b->Kd_);
            row[4] = b->Ns_;
            if (b->usetex_Kd_) row[6] = add_tex(b->map_Kd_);
            const float ks_tex = b->usetex_Ks_ ? add_tex(b->map_Ks_) : -1.0f;
            ext.push_back({row_id, {b->Ks_[0], b->Ks_[1], b->Ks_[2], ks_tex}});
        } else if (const auto* s = dynamic_cast<const MatGlass*>(m)) {
            row[0] = NR_MAT_GLASS;
            set_rgb(s->F0_);
            row[5] = s->Ior_;
            if (s->usetex_F0_) row[6] = add_tex(s->map_F0_);
        } else {
            return false;
        }
        return true;
    }

    // The sky of cmd.cpp:59-65: Texture::load of the envmap key, times envmap_gain.
    bool upload_sky(Config& config) {
        const std::string path = config.getValueStr("envmap", "");
        const float gain = config.getValueFloat("envmap_gain", 1.0f);
        if (path.empty()) {
            std::cerr << "RendererPTNEE_MI355X: an environment map needs the envmap Config key" << std::endl;
            return false;
        }
        const std::string key = path + "|" + std::to_string(gain);
        if (key == env_key_) return true;
        Texture sky(path);
        const TexUpload u = texels(sky, 0);
        if (u.w <= 0 || u.h <= 0 || nr_scene_set_envmap(scene_, u.w, u.h, u.rgb.data(), gain) != NR_OK) {
            std::cerr << "RendererPTNEE_MI355X: sky upload failed: " << nr_last_error() << std::endl;
            return false;
        }
        env_key_ = key;
        return true;
    }
};
