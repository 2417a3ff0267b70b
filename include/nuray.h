/* nuray.h -- C ABI of the MI355X ptnee path (libnuray.so).
 *
 * The reference (mollnn/nuRay) has no C ABI: its boundary is the C++ virtual class
 * Renderer { prepare(triangles); render(camera, triangles, img, config, ...); setEnvmap(env); }
 * (src/renderer/renderer.h:15-35) plus the Triangle / Material / Texture / Camera / Config value
 * types.  Each entry point below names the reference interface it replaces.  Host C++ (the
 * nuray::Renderer facade, include/nuray.hpp) and Python tests bind to these functions.
 *
 * Conventions: every function returns NR_OK (0) or a negative error code; nr_last_error()
 * returns the message of the last failure on the calling thread.  No exceptions cross the ABI.
 * Buffers are caller-owned host memory unless a name ends in _device.  One scene is used from
 * one host thread at a time; `cancel` may be written from any thread.
 */
#ifndef NURAY_H
#define NURAY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NR_OK 0
#define NR_CANCELLED 1
#define NR_ERROR (-1)

typedef struct nr_scene nr_scene;

/* Material rows passed to nr_scene_create: 8 floats {type, r, g, b, alpha, ior, texture, 0}.
 * type: NR_MAT_LAMBERT (MatLambert, Kd), NR_MAT_LIGHT (MatLight, Ke), NR_MAT_GGXREFL (MatGGXRefl,
 * F0 + roughness alpha), NR_MAT_GGX (MatGGX translucent dielectric, F0 + alpha + ior),
 * NR_MAT_BLINNPHONG (MatBlinnPhong, Kd + exponent Ns in `alpha`; Ks via nr_scene_set_material_ext),
 * NR_MAT_GLASS (MatGlass, F0 + ior).
 * texture: index of an albedo texture (nr_scene_set_texture_rgba8) or -1. */
#define NR_MAT_LAMBERT 0
#define NR_MAT_LIGHT 1
#define NR_MAT_GGXREFL 2
#define NR_MAT_GGX 3
#define NR_MAT_BLINNPHONG 4
#define NR_MAT_GLASS 5

/* Camera(fov, asp, pos, yaw, pitch, roll), src/scene/camera.h:18 and camera.cpp:8-22,88-105. */
typedef struct nr_camera {
    float pos[3];
    float euler[3]; /* yaw, pitch, roll in degrees */
    float fov;      /* vertical field of view in degrees (camera.cpp:28) */
    float aspect;   /* film width / height */
} nr_camera;

/* The Config keys RendererPTLS::render / ::trace read (rendererptls.cpp:65,81,98-100,130,191)
 * plus the library's own seed / maxdepth / rrdepth.  maxdepth = 0 and rrdepth = 0 are the reference's
 * semantics (no depth limit, Russian roulette at every vertex). */
typedef struct nr_params {
    int32_t width, height; /* imgw, imgh */
    int32_t spp;           /* spp */
    int32_t spp_offset;    /* first sample index (progressive / split renders) */
    uint32_t seed;         /* Philox key */
    float prr;             /* prr, survival probability of Russian roulette */
    int32_t mis;           /* mis: 1 = balance heuristic, 0 = fixed 0.5 weights */
    int32_t maxdepth;      /* path vertices with index >= maxdepth are not created (0 = unbounded) */
    int32_t rrdepth;       /* Russian roulette applies at vertex index >= rrdepth */
    int32_t batch_spp;     /* samples per pixel per wavefront batch (0 = automatic) */
    int32_t accumulate;    /* 1: add to the previous render's sums instead of clearing them */
    int32_t integrator;    /* renderer: 0 = ptnee (RendererPTLS), 1 = pt (RendererPT, rendererpt.cpp:13-61) */
    int32_t env_is;        /* 1: also sample the environment map as a light at Lambert and GGX-reflect
                              vertices, weighted against BSDF sampling by the balance heuristic so that
                              the expected image stays the reference's (a variance reduction the
                              reference does not have; ptnee and the radiance cache, not the adjoint) */
    int32_t path_kernel;   /* 1: trace each sample's whole path in one lane (one launch per batch) instead of
                              the wavefront (one launch per bounce); the per-sample radiance and the image
                              are bit-identical.  Pays off for small scenes and for small frames of long
                              paths, e.g. inverse rendering of a closed room; not with env_is or the
                              radiance cache.  2: automatic -- without env_is, one lane per path for
                              scenes of at most 64 BVH leaves and for renders whose samples x triangles
                              stay below 2e12, the wavefront otherwise (the radiance cache always uses
                              the wavefront).  0 (default): the wavefront */
} nr_params;

typedef void (*nr_progress_fn)(float fraction, void* user);

const char* nr_last_error(void);
/* Selects the HIP device for this host thread (and looks up the device address of the scratch
 * array nr_intersect_noalloc uses on it). */
int nr_init(int32_t device);
/* End of a session: waits for all work on the current device and forgets the per-device state the
 * library keeps (the scratch addresses).  Scenes and caches must be destroyed before.  The
 * reference's session ends with its process (cmd.cpp:112-113: save, then return from main); a
 * host that embeds the library calls this before it lets the device go. */
int nr_shutdown(void);
int nr_synchronize(void);

/* Replaces Loader::getTriangles() + Triangle::Triangle (src/scene/loader.h:15-19,
 * triangle.cpp:5-25): triangles in loader order, p = N x 9 floats (p0, p1, p2), n = N x 9 vertex
 * normals or NULL (flat normals when |n0|^2 < 1e-4), uv = N x 6 or NULL, one material row per
 * triangle.  Triangle ids reported by nr_intersect are indices into this order. */
nr_scene* nr_scene_create(const float* p, const float* n, const float* uv, const int32_t* mat_id,
                          uint32_t n_tris, const float* mats, uint32_t n_mats);
int nr_scene_destroy(nr_scene* s);
/* Texture::load result (texture.cpp:5-31): RGBA8 texels, row-major, first row at v = 1.
 * Textures, material rows and the environment map may change after nr_bvh_build_sah: the texture
 * pool and material table are re-uploaded, the BVH stays. */
int nr_scene_set_texture_rgba8(nr_scene* s, int32_t id, int32_t w, int32_t h, const uint8_t* rgba);
/* MatBlinnPhong's specular colour (matblinnphong.h:16-17): ext = {Ks r, g, b, Ks texture or -1}
 * for material row `id` (the loader's map_Ks). */
int nr_scene_set_material_ext(nr_scene* s, int32_t id, const float* ext);
/* Linear RGB float albedo texture (3 floats per texel, first row at v = 1): the form to optimise
 * with nr_adjoint_ptnee_tex.  Setting a texture of the same size and format on a built scene
 * rewrites it in place (no rebuild).  No reference counterpart (the reference's textures are 8-bit). */
int nr_scene_set_texture_rgb32f(nr_scene* s, int32_t id, int32_t w, int32_t h, const float* rgb);
/* Renderer::setEnvmap / Envmap(texture, gain) (renderer.h:30, envmap.cpp:17-24): linear RGB float
 * texels, row-major, looked up with the reference's bilinear rule.  Allowed before or after
 * nr_bvh_build_sah (replacing the map rebuilds the sky sampler of nr_params.env_is). */
int nr_scene_set_envmap(nr_scene* s, int32_t w, int32_t h, const float* rgb, float gain);

/* Renderer::prepare (renderer.cpp:30-39): LightSampler::setPrimitives + BVH::build with the
 * reference's SAH rule (bvh.cpp:5-97), flattened and uploaded to HBM.  nthreads <= 0: all cores. */
int nr_bvh_build_sah(nr_scene* s, int32_t nthreads);
int nr_bvh_info(nr_scene* s, int32_t* n_nodes, int32_t* depth, double* build_ms);
/* Topology dump in preorder over all nodes: children (-1 for leaves), boxes (min xyz, max xyz),
 * leaf primitive ranges into prim_order (loader ids in DFS leaf order). */
int nr_bvh_export(nr_scene* s, int32_t* left, int32_t* right, float* box, int32_t* leaf_first,
                  int32_t* leaf_count, int32_t* prim_order);

/* BVH::build alone, on the host, without a device (same rule and export format as above).
 * Arrays sized for 2 * n_tris - 1 nodes and n_tris primitives; returns the node count or < 0. */
int nr_bvh_build_host(const float* p, uint32_t n_tris, int32_t nthreads, int32_t* depth, int32_t* left,
                      int32_t* right, float* box, int32_t* leaf_first, int32_t* leaf_count, int32_t* prim_order);

/* BVH::intersection (bvh.cpp:129-137) on a ray batch: o, d = n x 3 floats.
 * any_hit = 0: closest hit with the reference's tie order; tri = loader id or -1, t = 1e18 on a miss.
 * any_hit = 1: the shadow-ray test of rendererptls.cpp:70-71 with segment length tmax[i]: the ray
 * is blocked iff the reference's closest hit satisfies (double)t + 2e-3 <= tmax[i]; tri = a blocking
 * triangle or -1 when the segment is clear. */
int nr_intersect(nr_scene* s, const float* o, const float* d, uint32_t n, int32_t any_hit, const float* tmax,
                 int32_t* tri, float* t, float* b1, float* b2);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): nr_intersect with every
 * array in device memory and no allocation at all -- the ray records, hit records, retry list and
 * counters the traversal needs live in one static per-GPU scratch array (13 MiB, held until the
 * process ends).  At most NR_INTERSECT_NOALLOC_MAX rays per call (larger n: NR_ERROR).  Results are
 * those of nr_intersect, bit for bit.  The call runs on the scene's stream and returns after the
 * results are written.  Calls that overlap on one GPU (several threads or scenes at once) race on
 * the shared scratch array. */
#define NR_INTERSECT_NOALLOC_MAX 262144
int nr_intersect_noalloc(nr_scene* s, const float* o_device, const float* d_device, uint32_t n, int32_t any_hit,
                         const float* tmax_device, int32_t* tri_device, float* t_device, float* b1_device,
                         float* b2_device);

/* RendererPTLS::render (rendererptls.cpp:96-236) into the scene's device accumulators.
 * Outputs may be NULL (then fetch later with nr_render_fetch): rgb_linear = mean radiance
 * (W*H*3 floats), sum / sum2 = FP64 per-pixel sums of the samples and their squares, rgb8 = the
 * reference's display transform.  progress is called after each wavefront batch; a non-zero
 * *cancel stops the render (returns NR_CANCELLED; the partial sums stay readable). */
int nr_render_ptnee(nr_scene* s, const nr_camera* cam, const nr_params* params, float* rgb_linear, double* sum,
                    double* sum2, uint8_t* rgb8, nr_progress_fn progress, void* user, volatile int32_t* cancel);
int nr_render_fetch(nr_scene* s, float* rgb_linear, double* sum, double* sum2, uint8_t* rgb8);
/* Statistics of the last render: [0] wall ms, [1] samples, [2] wavefront iterations,
 * [3] closest-hit kernel ms, [4] closest-hit launches, [5] rays traced (closest),
 * [6] node fetches (closest; 4-wide nodes of 128 B), [7] triangle tests (closest), [8] shadow kernel ms,
 * [9] shadow rays, [10] shadow node fetches, [11] shadow triangle tests, [12] shade kernel ms,
 * [13] shade launches, [14] shadow launches, [15] rays traced again with the full-depth stack
 * (profiling level 2).
 * Kernel times are filled when the render ran with timing on (nr_set_profiling). */
int nr_render_stats(nr_scene* s, double* out16);
/* 1: time the traversal launches with HIP events; 2: also count node fetches / triangle tests.
 * With either on, every launch of a wavefront render runs on one stream (otherwise a bounce's
 * shadow launch runs beside the next bounce's closest-hit launch), so that each kernel's time and
 * counts are its own. */
int nr_set_profiling(nr_scene* s, int32_t mode);
/* Traversal tuning (no counterpart in the reference, whose traversal recurses on the CPU stack):
 * live LDS stack entries per lane of the first persistent trace launch (default 16, at least 1).
 * A ray that needs more is traced again from its start by a launch with the full-depth stack, so
 * results do not depend on it; smaller values raise occupancy, larger ones retry fewer rays. */
int nr_scene_set_short_stack(nr_scene* s, int32_t entries);
/* Traversal tuning (no counterpart in the reference): the order in which a wavefront render's
 * trace launches take their queued rays.  0: queue order, in which neighbouring rays come from
 * neighbouring pixels; 1: secondary and shadow rays grouped by direction octant within chunks of
 * 1024 queue positions; 2: by 64 direction bins within chunks of 4096; 3 (default): shadow rays
 * alone by the 64 bins within chunks of 256, so that a wave's shadow rays leave a few neighbouring
 * pixels towards the same light; 4: the same within chunks of 1024.  Results do not depend on it
 * (ties are broken by leaf order, not by visit order). */
int nr_scene_set_trace_order(nr_scene* s, int32_t mode);

/* ---- Differentiable ptnee (no counterpart in the reference, which has no adjoint).
 * theta: one row {r, g, b, alpha} per material -- albedo / F0 and GGX roughness; rows of
 * emitters are ignored.  Replaces the materials' Kd and alpha_ (matlambert.h, matggxrefl.h). */
int nr_scene_set_params(nr_scene* s, const float* theta, uint32_t n_mats);
/* Mean squared error of the last render's mean image against target (W*H*3 floats); dLoss/dImage
 * stays on the device for nr_adjoint_ptnee and is also copied to dL_dimg when non-NULL. */
int nr_loss_grad_l2(nr_scene* s, const float* target, double* loss, float* dL_dimg);
/* Path-replay backpropagation: grad (n_mats x 4 doubles, same layout as theta) of the loss whose
 * image gradient is dL_dimg (W*H*3 floats, NULL = the one nr_loss_grad_l2 left on the device),
 * for the samples of cam / params (the last render's when they match).  When the last
 * nr_render_ptnee ran those samples as one wavefront batch its per-sample radiance is reused;
 * otherwise each sample is replayed forward first (recompute-only: no per-path storage, any
 * sample count), one lane per path for scenes of at most 64 BVH leaves and as params->path_kernel
 * chooses for the others (0: the wavefront; 1: one lane per path; 2: one lane per path while a
 * chunk's samples x triangles stay below 2e12).  Gradients for Lambert and GGX-reflect materials, with or without albedo textures
 * (a textured material's Kd row gets none); paths through translucent GGX dielectrics are
 * replayed, but the dielectrics' own rows get no gradient. */
int nr_adjoint_ptnee(nr_scene* s, const nr_camera* cam, const nr_params* params, const float* dL_dimg,
                     double* grad);
/* The same, plus the gradient with respect to the linear texel values of albedo texture tex_id
 * (w * h * 3 doubles, texel order of the upload), scattered with the bilinear weights of
 * Texture::pixelBi (texture.cpp:54-72); texels outside the image get none. */
int nr_adjoint_ptnee_tex(nr_scene* s, const nr_camera* cam, const nr_params* params, const float* dL_dimg,
                         double* grad, int32_t tex_id, double* grad_texels);

/* ---- Neural radiance cache.  Replaces NeuralRadianceCache (renderernrc.h:17-43) and its NNNode
 * network (nnnode.cpp:1-208): here a 44 -> 5 x 64 (ReLU) -> 3 MLP on BF16 MFMA with FP32
 * accumulation, trained with Adam on the relative L2 loss mean_s sum_c (y - t)^2 / (y^2 + 0.01).
 * Weights: NR_NRC_NPARAM floats, W^T per layer (row = output unit, 64 inputs per row); layers 0..4
 * are 64 x 64, layer 5 is 16 x 64 of which rows 0..2 are used; input 44 is a constant 1.
 * Query rows: 16 floats {pos x, y, z, roughness alpha, wo theta/pi, wo phi/(2pi) + 1/2,
 * n theta/pi, n phi/(2pi) + 1/2, albedo r, g, b, 0, 0, 0, 0, 0}; theta is measured from +y and
 * phi = atan2(z, x).  Encoding: 12-frequency triangle waves of x, y, z (renderernrc.cpp:50-64,
 * 108-110), the four angles, exp(-alpha) (renderernrc.cpp:115), albedo. */
#define NR_NRC_NPARAM 21504
typedef struct nr_nrc nr_nrc;
/* He-uniform initial weights from `seed`; lr of Adam (beta1 0.9, beta2 0.99, eps 1e-8);
 * max_batch training samples per step (<= 0: 65536).  Inference (nr_nrc_infer and rendering)
 * uses the bias-corrected exponential moving average (decay 0.99) of the trained weights. */
nr_nrc* nr_nrc_create(uint32_t seed, float lr, int32_t max_batch);
int nr_nrc_destroy(nr_nrc* c);
int nr_nrc_set_weights(nr_nrc* c, const float* w, uint32_t n);  /* also resets Adam */
int nr_nrc_get_weights(nr_nrc* c, float* w, uint32_t n);
/* NeuralRadianceCache::eval without the clamp: rgb = n x 3 raw network outputs. */
int nr_nrc_infer(nr_nrc* c, const float* queries, uint32_t n, float* rgb);
/* NeuralRadianceCache::train over a batch: one Adam step; *loss (if non-NULL) = mean loss before
 * the step. */
int nr_nrc_train_step(nr_nrc* c, const float* queries, const float* targets, uint32_t n, double* loss);
/* Gradient of the mean loss (NR_NRC_NPARAM floats) without an update. */
int nr_nrc_gradient(nr_nrc* c, const float* queries, const float* targets, uint32_t n, float* grad, double* loss);

/* No reference counterpart: times the cache kernels on n synthetic device-resident queries,
 * `reps` launches each (HIP events): *infer_ms per inference launch over n queries, *train_ms
 * per training step (forward + backward + weight gradient + Adam) on min(n, max_batch) samples.
 * Changes the weights. */
int nr_nrc_bench(nr_nrc* c, uint32_t n, int32_t reps, double* infer_ms, double* train_ms);

/* RendererNRC::render (renderernrc.cpp:185-281), as the radiance-cache method: the frame's
 * p.spp samples per pixel follow ptnee until the path spread a(x1..xn) exceeds spread * a0
 * (a0 = the primary vertex's footprint), where the cache's answer beta * cache(x, wo) ends the path;
 * train_paths extra paths from random pixels record up to train_maxlen vertices (the last answered
 * by the cache) whose observed radiance trains the cache once per frame (batch <= the cache's
 * max_batch).  use_cache = 0 renders plain ptnee; train = 0 skips the training step.  Results
 * accumulate like nr_render_ptnee (params->accumulate) and are read with nr_render_fetch.
 * stats (optional, 4 doubles): training loss, cache queries, recorded vertices, vertices trained. */
typedef struct nr_nrc_params {
    int32_t train_paths;
    int32_t train_maxlen;
    float spread;
    int32_t use_cache;
    int32_t train;
    int32_t train_steps; /* Adam steps per frame over disjoint parts of the recorded vertices */
    int32_t min_depth;   /* render paths query the cache only from this vertex index on (>= 1) */
} nr_nrc_params;
int nr_render_nrc(nr_scene* s, nr_nrc* c, const nr_camera* cam, const nr_params* params, const nr_nrc_params* np,
                  double* stats);
/* No reference counterpart (the reference keeps its training records inside RendererNRC,
 * renderernrc.cpp:128-133, 175-182): the training records of the scene's last radiance-cache frame
 * and their targets, for checking the cache's data path.  recs: cap rows of 16 floats in the query
 * layout above, except float 11 = the path slot (int bits) and floats 12..14 = the path throughput
 * beta on arrival at the vertex; targets: cap rows of 4 floats, xyz = (L_path - L_before) / beta
 * (L_before: the radiance the path had collected before the vertex; 0 where beta <= 1e-8).
 * Records of one path are in vertex order.  *n = the number of records with targets (at most the
 * cache's max_batch), also when it exceeds cap. */
int nr_nrc_frame_records(nr_scene* s, float* recs, float* targets, int32_t cap, int32_t* n);

/* No reference counterpart: a plain device-to-device copy of `bytes` bytes (16-byte vector loads
 * and stores), run `reps` times; *gbs = bytes read + written per second.  Calibrates the
 * hardware byte counters (a known number of HBM bytes per dispatch) and the achievable HBM rate. */
int nr_bench_copy(uint64_t bytes, int32_t reps, double* gbs);

/* No reference counterpart: the rate (*mops, million per second) of device-scope atomic adds in
 * the wavefront's append pattern -- 5 blocks of 256 threads per CU, each block iteration one add
 * by one thread, waited for by the block -- with block b adding to counter b % counters, the
 * counters stride_ints ints apart, iters iterations per block.  Measures how such atomics share a
 * cache line (the layout of the wavefront's counters follows from it). */
int nr_bench_atomics(int32_t counters, int32_t stride_ints, int32_t iters, double* mops);

#ifdef __cplusplus
}
#endif
#endif
