// Wavefront ptnee renderer: buffers and host orchestration.
#pragma once
#include <cstring>
#include "scene.h"
#include "trace.h"

namespace nr {

// Camera::generateRay inputs (camera.cpp:24-32), precomputed once per render on the host.
struct CamView {
    float pos[3], gaze[3], up[3], hand[3];
    float fw, fh;  // film width / height
};
// Camera(fov, asp, pos, yaw, pitch, roll) with fromEuler (camera.cpp:88-105).
CamView make_camera(const float pos[3], const float euler[3], float fov, float aspect);

struct RenderParams {
    int W = 1, H = 1, spp = 1, spp_offset = 0;
    uint32_t seed = 0;
    float prr = 0.8f;
    int mis = 1;
    int maxdepth = 0;   // 0: unbounded (reference)
    int rrdepth = 0;    // Russian roulette from this vertex index (reference: 0)
    int batch_spp = 0;  // samples per pixel per wavefront batch (0: auto)
    int accumulate = 0; // keep previous sums instead of clearing them
    int integrator = 0; // 0: RendererPTLS (ptnee), 1: RendererPT (pt, no light sampling)
    int env_is = 0;     // 1: sample the sky at Lambert / GGX-reflect vertices (same expected image)
    int path_kernel = 0; // 1: one lane per path instead of the wavefront (same radiance, bit for bit);
                         // 2: one lane per path for small scenes (flat leaf list) or small renders
                         // (samples x triangles <= 2e12) without env_is
};

struct RenderStats {
    double ms = 0;
    long long samples = 0, extension_rays = 0, shadow_rays = 0;
    int iterations = 0;
};

// HIP-event timing of the traversal launches of one render (on the scene's stream).
struct KernelTimes {
    double closest_ms = 0, shadow_ms = 0, shade_ms = 0;
    int closest_launches = 0, shadow_launches = 0, shade_launches = 0;
};

// Vertex records kept by a one-lane-per-path render for the adjoint: the first VREC_CAP vertices of
// every sample (16 B each), when the whole render fits VREC_BUDGET bytes of them.
constexpr int VREC_CAP = 16;
constexpr size_t VREC_BUDGET = size_t(4) << 30;

// Automatic render mode (nr_params::path_kernel = 2): one lane per path while samples x triangles
// stay below this, where the wavefront's per-bounce launches cost more than the per-lane walks'
// divergence.  Measured (maxdepth 10): one lane per path still wins at 1.7e12 (114 k triangles x
// 14.7 M samples, 684 k x 2.1 M) and loses at 2.5e12 (684 k x 3.7 M).
constexpr double AUTO_PATHS_WORK = 2e12;

struct RenderTarget {
    int W = 0, H = 0;
    size_t cap = 0;
    DevBuf sum, sum2;                  // double, W*H*3
    DevBuf rays[2], beta[2];           // extension-ray queues and their (throughput, MIS factor)
    DevBuf hits, shadow, contrib;      // closest hits; shadow rays and their NEE contributions
    DevBuf radiance;                   // float4 per path slot (x2 with sky sampling: sky rays' region)
    DevBuf skyw;                       // sky sampling: BSDF-escape MIS weight per path slot
    DevBuf retry, retry_sh;            // rays the closest-hit / shadow launch hands to its full-stack retry launch
    DevBuf order;                      // traversal order of the ray queue, then of the shadow queue
    DevBuf counters;                   // the wavefront's counters, one 128-byte line each
    DevBuf stats;                      // TraceStats[2]
    int* pinned = nullptr;             // host-visible counter readback
    long long spp_total = 0;
    // adjoint: the per-sample radiance of the last render is replayable when it ran as one batch
    bool adj_valid = false;
    RenderParams adj_p;   // the render those samples came from
    uint64_t adj_gen = 0; // and the scene generation it rendered
    CamView adj_cam{};
    DevBuf adj_target, adj_dldi, adj_loss, adj_grad, adj_gtex;
    DevBuf adj_rad;                    // float4 per sample of one adjoint chunk (recomputed radiance)
    DevBuf adj_err;                    // int: the adjoint's traversal-overflow flag
    DevBuf path_claim;                 // unsigned: samples claimed by a per-lane path kernel launch
    // vertex records of the last render when it ran as one batch (see adjoint.hip)
    DevBuf adj_vrec;
    bool adj_vrec_valid = false;
    // radiance-cache rendering: spread state per queued ray, cache queries, training vertices
    DevBuf nstate[2], nrc_queries, nrc_recs, nrc_lb, nrc_tgt, nrc_batch, nrc_batch_tgt;
    int nrc_last_records = 0;  // records of the last cache frame that have targets (nrc_recs, nrc_tgt)
    ~RenderTarget();
    void ensure(int W, int H, size_t cap);  // image accumulators for W x H and queues for cap paths
    void ensure_queues(size_t cap);         // queues only (the accumulators keep their contents)
    // the stored per-sample radiance belongs to exactly the samples of (cam, p) on the scene as
    // it is now (no material, parameter, texture or sky change since that render)
    bool adj_matches(const CamView& cam, const RenderParams& p, uint64_t scene_generation) const {
        const RenderParams& q = adj_p;
        return adj_valid && adj_gen == scene_generation && W == p.W && H == p.H && q.W == p.W && q.H == p.H && q.spp == p.spp &&
               q.spp_offset == p.spp_offset && q.seed == p.seed && q.prr == p.prr && q.mis == p.mis &&
               q.maxdepth == p.maxdepth && q.rrdepth == p.rrdepth && q.integrator == p.integrator &&
               std::memcmp(&adj_cam, &cam, sizeof(CamView)) == 0;
    }
};

// RendererNRC::render parameters (renderernrc.cpp:185-281), with the spread rule of the
// radiance-cache method in place of the reference's fixed depth limit.
struct NrcRenderParams {
    int train_paths = 65536;  // training paths per frame (pixels drawn uniformly)
    int train_maxlen = 5;     // vertices of a training path; the last one is answered by the cache
    float c_spread = 0.01f;   // a render path stops at x_n when a(x1..xn) > c a0
    int use_cache = 1;        // 0: render paths are plain ptnee (training still runs)
    int train = 1;            // 0: no training step this frame
    int train_steps = 4;      // Adam steps per frame over disjoint parts of the recorded vertices
    int min_depth = 1;        // render paths query the cache only from this vertex index on
};
struct NrcFrameStats {
    double train_loss = 0;
    long long queries = 0, records = 0, trained = 0;
};

class Nrc;

typedef void (*ProgressFn)(float, void*);

// Renders p.spp samples per pixel into the scene's render target.  Returns 0, or 1 if cancelled.
int render_ptnee(Scene& sc, const CamView& cam, const RenderParams& p, ProgressFn cb, void* user,
                 volatile int* cancel, RenderStats* st, bool trace_stats = false, TraceStats* tstats = nullptr,
                 KernelTimes* kt = nullptr);
// One radiance-cache frame: p.spp samples per pixel whose paths stop at the spread rule and take
// the cache's answer there, plus the training paths, then one training step on their vertices.
// Accumulates into the same sums as render_ptnee (p.accumulate).
int render_nrc(Scene& sc, Nrc& nrc, const CamView& cam, const RenderParams& p, const NrcRenderParams& np,
               NrcFrameStats* fs, RenderStats* st);
// The last radiance-cache frame's training records (NrcQuery, 16 floats each) and their targets
// (float4), the first min(cap, n) of them; returns n.
int nrc_frame_records(Scene& sc, float* recs, float* targets, int cap);
// Per-sample radiance of P paths (a whole number of samples per pixel, from sample0 on) into the
// render target's radiance buffer, slot = sample offset * W * H + pixel, by the wavefront, without
// touching the accumulated image.  Invalidates the stored radiance of the last render.  With
// `records`, also the vertex records of the first VREC_CAP vertices of every path into adj_vrec
// when they fit VREC_BUDGET; returns whether it wrote them.
bool wavefront_radiance(Scene& sc, const CamView& cam, const RenderParams& p, uint32_t sample0, int P,
                        bool records = false);
// Copies the accumulated result: linear mean (float W*H*3), FP64 sums, gamma-2.2 RGB8.
void render_fetch(Scene& sc, float* rgb, double* sum, double* sum2, uint8_t* rgb8);

}  // namespace nr
