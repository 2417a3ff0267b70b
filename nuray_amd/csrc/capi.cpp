// extern "C" entry points of libnuray.so (include/nuray.h).
#include "../../include/nuray.h"
#include <cstring>
#include <stdexcept>
#include <string>
#include "adjoint.h"
#include "noalloc.h"
#include "nrc.h"
#include "render.h"
#include "scene.h"

namespace nr {
double bench_copy(size_t bytes, int reps);
double bench_atomics(int n_cnt, int stride, int iters);
}

struct nr_scene {
    nr::Scene* sc;
    int profiling = 0;
    double stats[16] = {0};
};

namespace {
thread_local std::string g_err;

void check_hip_sync(hipStream_t s) { nr::check_hip(hipStreamSynchronize(s), "hipStreamSynchronize"); }

template <class F>
int guard(F&& f) {
    try {
        return f();
    } catch (const std::exception& e) {
        g_err = e.what();
    } catch (...) {
        g_err = "unknown error";
    }
    return NR_ERROR;
}
}  // namespace

extern "C" {

const char* nr_last_error(void) { return g_err.c_str(); }

int nr_init(int32_t device) {
    return guard([&] {
        nr::check_hip(hipSetDevice(device), "hipSetDevice");
        nr::check_hip(hipFree(nullptr), "context init");
        nr::noalloc_setup(device);
        return NR_OK;
    });
}

int nr_shutdown(void) {
    return guard([&] {
        nr::check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
        nr::noalloc_forget();
        return NR_OK;
    });
}

int nr_synchronize(void) {
    return guard([&] {
        nr::check_hip(hipDeviceSynchronize(), "hipDeviceSynchronize");
        return NR_OK;
    });
}

nr_scene* nr_scene_create(const float* p, const float* n, const float* uv, const int32_t* mat_id, uint32_t n_tris,
                          const float* mats, uint32_t n_mats) {
    nr_scene* out = nullptr;
    guard([&] {
        if (!p || !mat_id || !mats) throw std::runtime_error("null scene arrays");
        out = new nr_scene{new nr::Scene(p, n, uv, mat_id, n_tris, mats, n_mats)};
        return NR_OK;
    });
    return out;
}

int nr_scene_destroy(nr_scene* s) {
    return guard([&] {
        if (s) {
            delete s->sc;
            delete s;
        }
        return NR_OK;
    });
}

int nr_scene_set_texture_rgba8(nr_scene* s, int32_t id, int32_t w, int32_t h, const uint8_t* rgba) {
    return guard([&] {
        s->sc->add_texture(id, w, h, 0, rgba);
        return NR_OK;
    });
}

int nr_scene_set_material_ext(nr_scene* s, int32_t id, const float* ext) {
    return guard([&] {
        s->sc->set_material_ext(id, ext);
        return NR_OK;
    });
}

int nr_scene_set_texture_rgb32f(nr_scene* s, int32_t id, int32_t w, int32_t h, const float* rgb) {
    return guard([&] {
        if (s->sc->env_tex >= 0 && id == s->sc->env_tex) throw std::runtime_error("texture id is the environment map");
        s->sc->add_texture(id, w, h, 1, rgb);
        return NR_OK;
    });
}

int nr_scene_set_envmap(nr_scene* s, int32_t w, int32_t h, const float* rgb, float gain) {
    return guard([&] {
        // the environment map lives in the texture pool after the albedo textures; on a built
        // scene the texels (and the sky sampler) are replaced without rebuilding the BVH
        nr::Scene& sc = *s->sc;
        const int id = sc.env_tex >= 0 ? sc.env_tex : int(sc.textures.size());
        sc.env_tex = id;
        sc.env_gain = gain;
        sc.view.env_gain = gain;
        sc.add_texture(id, w, h, 1, rgb);
        return NR_OK;
    });
}

int nr_bvh_build_sah(nr_scene* s, int32_t nthreads) {
    return guard([&] {
        s->sc->build(nthreads);
        return NR_OK;
    });
}

int nr_bvh_info(nr_scene* s, int32_t* n_nodes, int32_t* depth, double* build_ms) {
    return guard([&] {
        if (!s->sc->built) throw std::runtime_error("BVH not built");
        if (n_nodes) *n_nodes = int32_t(s->sc->bvh.left.size());
        if (depth) *depth = s->sc->bvh.depth;
        if (build_ms) *build_ms = s->sc->bvh.build_ms;
        return NR_OK;
    });
}

int nr_bvh_export(nr_scene* s, int32_t* left, int32_t* right, float* box, int32_t* leaf_first, int32_t* leaf_count,
                  int32_t* prim_order) {
    return guard([&] {
        const nr::BvhBuild& b = s->sc->bvh;
        if (!s->sc->built) throw std::runtime_error("BVH not built");
        const size_t nn = b.left.size();
        std::memcpy(left, b.left.data(), nn * 4);
        std::memcpy(right, b.right.data(), nn * 4);
        std::memcpy(box, b.box.data(), nn * 24);
        std::memcpy(leaf_first, b.leaf_first.data(), nn * 4);
        std::memcpy(leaf_count, b.leaf_count.data(), nn * 4);
        std::memcpy(prim_order, b.order.data(), b.order.size() * 4);
        return NR_OK;
    });
}

int nr_bvh_build_host(const float* p, uint32_t n_tris, int32_t nthreads, int32_t* depth, int32_t* left,
                      int32_t* right, float* box, int32_t* leaf_first, int32_t* leaf_count, int32_t* prim_order) {
    int nn = NR_ERROR;
    int rc = guard([&] {
        nr::BvhBuild b = nr::build_sah(p, int(n_tris), nthreads);
        nn = int(b.left.size());
        if (depth) *depth = b.depth;
        std::memcpy(left, b.left.data(), nn * 4);
        std::memcpy(right, b.right.data(), nn * 4);
        std::memcpy(box, b.box.data(), size_t(nn) * 24);
        std::memcpy(leaf_first, b.leaf_first.data(), nn * 4);
        std::memcpy(leaf_count, b.leaf_count.data(), nn * 4);
        std::memcpy(prim_order, b.order.data(), b.order.size() * 4);
        return NR_OK;
    });
    return rc == NR_OK ? nn : rc;
}

int nr_intersect(nr_scene* s, const float* o, const float* d, uint32_t n, int32_t any_hit, const float* tmax,
                 int32_t* tri, float* t, float* b1, float* b2) {
    return guard([&] {
        nr::Scene& sc = *s->sc;
        if (!sc.built) throw std::runtime_error("BVH not built");
        if (any_hit && !tmax) throw std::runtime_error("any_hit needs tmax");
        if (any_hit)
            for (uint32_t i = 0; i < n; ++i)
                if (!(tmax[i] == tmax[i])) throw std::runtime_error("tmax must not be NaN");
        if (n == 0) return NR_OK;
        std::vector<nr::Ray> rays(n);
        for (uint32_t i = 0; i < n; ++i) {
            rays[i].o = make_float4(o[3 * i], o[3 * i + 1], o[3 * i + 2], any_hit ? tmax[i] : 0.0f);
            rays[i].d = make_float4(d[3 * i], d[3 * i + 1], d[3 * i + 2], 0.0f);
        }
        nr::DevBuf dr, dh, dc, dretry;
        dr.upload(rays.data(), n * sizeof(nr::Ray), sc.stream);
        dh.ensure(n * sizeof(nr::HitRec));
        dretry.ensure(n * sizeof(int));
        // c[0] ray count, c[1..3] the launch's work counters, c[4] overflow flag
        int hc[8] = {int(n), 0, 0, 0, 0, 0, 0, 0};
        dc.ensure(sizeof(hc));
        nr::check_hip(hipMemcpyAsync(dc.ptr, hc, sizeof(hc), hipMemcpyHostToDevice, sc.stream), "H2D");
        int* c = dc.as<int>();
        if (any_hit)
            nr::launch_shadow(sc.view, dr.as<nr::Ray>(), c, nullptr, dh.as<nr::HitRec>(), nullptr, nullptr, c + 1,
                              c + 4, nullptr, sc.stream, dretry.as<int>());
        else
            nr::launch_closest(sc.view, dr.as<nr::Ray>(), c, dh.as<nr::HitRec>(), c + 1, c + 4, nullptr, sc.stream,
                               dretry.as<int>());
        std::vector<nr::HitRec> h(n);
        nr::check_hip(hipMemcpyAsync(h.data(), dh.ptr, n * sizeof(nr::HitRec), hipMemcpyDeviceToHost, sc.stream), "D2H");
        nr::check_hip(hipMemcpyAsync(hc, dc.ptr, sizeof(hc), hipMemcpyDeviceToHost, sc.stream), "D2H");
        nr::check_hip(hipStreamSynchronize(sc.stream), "intersect sync");
        if (hc[4]) throw std::runtime_error("traversal stack overflow");
        for (uint32_t i = 0; i < n; ++i) {
            tri[i] = h[i].prim >= 0 ? sc.bvh.order[h[i].prim] : -1;
            t[i] = h[i].t;
            b1[i] = h[i].b1;
            b2[i] = h[i].b2;
        }
        return NR_OK;
    });
}

int nr_intersect_noalloc(nr_scene* s, const float* o_device, const float* d_device, uint32_t n, int32_t any_hit,
                         const float* tmax_device, int32_t* tri_device, float* t_device, float* b1_device,
                         float* b2_device) {
    static_assert(NR_INTERSECT_NOALLOC_MAX == nr::NOALLOC_MAX_RAYS, "public and internal ray limits");
    return guard([&] {
        nr::intersect_noalloc(*s->sc, o_device, d_device, n, any_hit, tmax_device, tri_device, t_device, b1_device,
                              b2_device);
        return NR_OK;
    });
}

static nr::RenderParams to_params(const nr_params* pp);

int nr_render_ptnee(nr_scene* s, const nr_camera* cam, const nr_params* pp, float* rgb_linear, double* sum,
                    double* sum2, uint8_t* rgb8, nr_progress_fn progress, void* user, volatile int32_t* cancel) {
    return guard([&] {
        nr::CamView cv = nr::make_camera(cam->pos, cam->euler, cam->fov, cam->aspect);
        const nr::RenderParams p = to_params(pp);
        nr::RenderStats rs;
        nr::TraceStats ts[2] = {};
        nr::KernelTimes kt;
        int st = nr::render_ptnee(*s->sc, cv, p, progress, user, (volatile int*)cancel, &rs, s->profiling >= 2, ts,
                                  s->profiling >= 1 ? &kt : nullptr);
        double* o = s->stats;
        o[0] = rs.ms; o[1] = double(rs.samples); o[2] = rs.iterations;
        o[3] = kt.closest_ms; o[4] = kt.closest_launches;
        o[5] = double(ts[0].rays); o[6] = double(ts[0].node_fetch); o[7] = double(ts[0].tri_tests);
        o[8] = kt.shadow_ms; o[9] = double(ts[1].rays); o[10] = double(ts[1].node_fetch); o[11] = double(ts[1].tri_tests);
        o[12] = kt.shade_ms; o[13] = kt.shade_launches; o[14] = kt.shadow_launches;
        o[15] = double(ts[0].retried + ts[1].retried);
        if (rgb_linear || sum || sum2 || rgb8) nr::render_fetch(*s->sc, rgb_linear, sum, sum2, rgb8);
        return st ? NR_CANCELLED : NR_OK;
    });
}

int nr_render_fetch(nr_scene* s, float* rgb_linear, double* sum, double* sum2, uint8_t* rgb8) {
    return guard([&] {
        nr::render_fetch(*s->sc, rgb_linear, sum, sum2, rgb8);
        return NR_OK;
    });
}

int nr_render_stats(nr_scene* s, double* out16) {
    return guard([&] {
        std::memcpy(out16, s->stats, sizeof(s->stats));
        return NR_OK;
    });
}

int nr_set_profiling(nr_scene* s, int32_t mode) {
    return guard([&] {
        s->profiling = mode;
        return NR_OK;
    });
}

int nr_scene_set_short_stack(nr_scene* s, int32_t entries) {
    return guard([&] {
        if (entries < 1) throw std::runtime_error("short stack needs at least one entry");
        s->sc->short_stack = entries;
        s->sc->view.short_stack = entries;
        return NR_OK;
    });
}

int nr_scene_set_trace_order(nr_scene* s, int32_t mode) {
    return guard([&] {
        if (mode < 0 || mode > 4) throw std::runtime_error("trace order must be 0 to 4");
        s->sc->trace_order = mode;
        return NR_OK;
    });
}

int nr_scene_set_params(nr_scene* s, const float* theta, uint32_t n_mats) {
    return guard([&] {
        nr::set_material_params(*s->sc, theta, int(n_mats));
        return NR_OK;
    });
}

int nr_loss_grad_l2(nr_scene* s, const float* target, double* loss, float* dL_dimg) {
    return guard([&] {
        *loss = nr::loss_l2(*s->sc, target, dL_dimg);
        return NR_OK;
    });
}

static nr::RenderParams to_params(const nr_params* pp) {
    nr::RenderParams p;
    p.W = pp->width; p.H = pp->height; p.spp = pp->spp; p.spp_offset = pp->spp_offset;
    p.seed = pp->seed; p.prr = pp->prr; p.mis = pp->mis; p.maxdepth = pp->maxdepth; p.rrdepth = pp->rrdepth;
    p.batch_spp = pp->batch_spp; p.accumulate = pp->accumulate;
    p.integrator = pp->integrator;
    if (p.integrator != 0 && p.integrator != 1) throw std::runtime_error("integrator must be 0 (ptnee) or 1 (pt)");
    p.env_is = pp->env_is;
    if (p.env_is != 0 && p.env_is != 1) throw std::runtime_error("env_is must be 0 or 1");
    p.path_kernel = pp->path_kernel;
    if (p.path_kernel < 0 || p.path_kernel > 2) throw std::runtime_error("path_kernel must be 0, 1 or 2");
    return p;
}

int nr_adjoint_ptnee(nr_scene* s, const nr_camera* cam, const nr_params* params, const float* dL_dimg, double* grad) {
    return guard([&] {
        nr::CamView cv = nr::make_camera(cam->pos, cam->euler, cam->fov, cam->aspect);
        nr::adjoint_ptnee(*s->sc, cv, to_params(params), dL_dimg, grad);
        return NR_OK;
    });
}

int nr_adjoint_ptnee_tex(nr_scene* s, const nr_camera* cam, const nr_params* params, const float* dL_dimg,
                         double* grad, int32_t tex_id, double* grad_texels) {
    return guard([&] {
        nr::CamView cv = nr::make_camera(cam->pos, cam->euler, cam->fov, cam->aspect);
        nr::adjoint_ptnee(*s->sc, cv, to_params(params), dL_dimg, grad, tex_id, grad_texels);
        return NR_OK;
    });
}

int nr_bench_copy(uint64_t bytes, int32_t reps, double* gbs) {
    return guard([&] {
        *gbs = nr::bench_copy(size_t(bytes), reps);
        return NR_OK;
    });
}

int nr_bench_atomics(int32_t counters, int32_t stride_ints, int32_t iters, double* mops) {
    return guard([&] {
        *mops = nr::bench_atomics(counters, stride_ints, iters);
        return NR_OK;
    });
}

// ---- neural radiance cache
struct nr_nrc {
    nr::Nrc* n;
};

nr_nrc* nr_nrc_create(uint32_t seed, float lr, int32_t max_batch) {
    nr_nrc* out = nullptr;
    guard([&] {
        nr::NrcConfig c;
        c.lr = lr;
        if (max_batch > 0) c.max_batch = max_batch;
        out = new nr_nrc{new nr::Nrc(seed, c)};
        return NR_OK;
    });
    return out;
}

int nr_nrc_destroy(nr_nrc* h) {
    return guard([&] {
        if (h) {
            delete h->n;
            delete h;
        }
        return NR_OK;
    });
}

int nr_nrc_set_weights(nr_nrc* h, const float* w, uint32_t n) {
    return guard([&] {
        if (n != uint32_t(NR_NRC_NPARAM)) throw std::runtime_error("nrc: weight count must be NR_NRC_NPARAM");
        h->n->set_weights(w);
        return NR_OK;
    });
}

int nr_nrc_get_weights(nr_nrc* h, float* w, uint32_t n) {
    return guard([&] {
        if (n != uint32_t(NR_NRC_NPARAM)) throw std::runtime_error("nrc: weight count must be NR_NRC_NPARAM");
        h->n->get_weights(w);
        return NR_OK;
    });
}

namespace {
// Host query rows (16 floats) -> device staging.
const nr::NrcQuery* stage_queries(nr::Nrc& n, const float* q, uint32_t count) {
    n.q_tmp.ensure(sizeof(nr::NrcQuery) * size_t(count));
    n.q_tmp.upload(q, sizeof(nr::NrcQuery) * size_t(count), n.stream);
    return n.q_tmp.as<nr::NrcQuery>();
}
const float4* stage_targets(nr::Nrc& n, const float* t, uint32_t count) {
    std::vector<float> t4(size_t(count) * 4, 0.0f);
    for (size_t i = 0; i < count; ++i)
        for (int c = 0; c < 3; ++c) t4[i * 4 + c] = t[i * 3 + c];
    n.t_tmp.ensure(sizeof(float4) * size_t(count));
    n.t_tmp.upload(t4.data(), sizeof(float4) * size_t(count), n.stream);
    check_hip_sync(n.stream);
    return n.t_tmp.as<float4>();
}
}  // namespace

int nr_nrc_infer(nr_nrc* h, const float* queries, uint32_t n, float* rgb) {
    return guard([&] {
        if (n == 0) return NR_OK;
        nr::Nrc& c = *h->n;
        const nr::NrcQuery* q = stage_queries(c, queries, n);
        c.o_tmp.ensure(sizeof(float4) * size_t(n));
        c.infer(q, int(n), c.o_tmp.as<float4>());
        std::vector<float> o4(size_t(n) * 4);
        nr::check_hip(hipMemcpyAsync(o4.data(), c.o_tmp.ptr, sizeof(float4) * size_t(n), hipMemcpyDeviceToHost, c.stream),
                      "nrc infer readback");
        check_hip_sync(c.stream);
        for (size_t i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) rgb[i * 3 + k] = o4[i * 4 + k];
        return NR_OK;
    });
}

int nr_nrc_train_step(nr_nrc* h, const float* queries, const float* targets, uint32_t n, double* loss) {
    return guard([&] {
        nr::Nrc& c = *h->n;
        const nr::NrcQuery* q = stage_queries(c, queries, n);
        const float4* t = stage_targets(c, targets, n);
        const double l = c.train(q, t, int(n), loss != nullptr);
        if (loss) *loss = l;
        return NR_OK;
    });
}

int nr_nrc_gradient(nr_nrc* h, const float* queries, const float* targets, uint32_t n, float* grad, double* loss) {
    return guard([&] {
        nr::Nrc& c = *h->n;
        const nr::NrcQuery* q = stage_queries(c, queries, n);
        const float4* t = stage_targets(c, targets, n);
        const double l = c.train(q, t, int(n), true, grad);
        if (loss) *loss = l;
        return NR_OK;
    });
}

int nr_nrc_bench(nr_nrc* c, uint32_t n, int32_t reps, double* infer_ms, double* train_ms) {
    return guard([&] {
        if (n == 0 || reps <= 0) throw std::runtime_error("nrc bench: n and reps must be positive");
        const double ms = c->n->bench(int(n), reps, train_ms);
        if (infer_ms) *infer_ms = ms;
        return NR_OK;
    });
}

int nr_nrc_frame_records(nr_scene* s, float* recs, float* targets, int32_t cap, int32_t* n) {
    return guard([&] {
        if (!s || !n || cap < 0 || (cap > 0 && (!recs || !targets))) throw std::runtime_error("bad argument");
        *n = nr::nrc_frame_records(*s->sc, recs, targets, cap);
        return NR_OK;
    });
}

int nr_render_nrc(nr_scene* s, nr_nrc* c, const nr_camera* cam, const nr_params* params, const nr_nrc_params* np,
                  double* stats) {
    return guard([&] {
        if (!s || !c || !cam || !params || !np) throw std::runtime_error("null argument");
        nr::CamView cv = nr::make_camera(cam->pos, cam->euler, cam->fov, cam->aspect);
        nr::NrcRenderParams rp;
        rp.train_paths = np->train_paths;
        rp.train_maxlen = np->train_maxlen;
        rp.c_spread = np->spread;
        rp.use_cache = np->use_cache;
        rp.train = np->train;
        rp.train_steps = np->train_steps;
        rp.min_depth = np->min_depth;
        nr::NrcFrameStats fs;
        nr::RenderStats st;
        nr::render_nrc(*s->sc, *c->n, cv, to_params(params), rp, &fs, &st);
        s->stats[0] = st.ms;
        s->stats[1] = double(st.samples);
        s->stats[2] = st.iterations;
        if (stats) {
            stats[0] = fs.train_loss;
            stats[1] = double(fs.queries);
            stats[2] = double(fs.records);
            stats[3] = double(fs.trained);
        }
        return NR_OK;
    });
}

}  // extern "C"
