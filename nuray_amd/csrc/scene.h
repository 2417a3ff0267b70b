// Host scene: owns the loader-order arrays, the SAH BVH and the device copies.
#pragma once
#include <cstdint>
#include <string>
#include <vector>
#include "bvh_build.h"
#include "nr_device.h"

namespace nr {

void check_hip(hipError_t e, const char* what);  // throws std::runtime_error on failure

// A device allocation that frees itself.
struct DevBuf {
    void* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf();
    void ensure(size_t n);          // grow (contents not kept)
    void upload(const void* src, size_t n, hipStream_t s);
    template <class T>
    T* as() const { return static_cast<T*>(ptr); }
};

struct Texture {
    int w = 0, h = 0, fmt = 0;  // 0 RGBA8, 1 RGB32F
    std::vector<uint8_t> bytes;
};

struct RenderTarget;  // defined in render.h

class Scene {
public:
    // loader-order inputs (triangle.h:8-36 fields)
    std::vector<float> p, n, uv;  // N*9, N*9, N*6
    std::vector<int32_t> mat;
    std::vector<float> mats;      // M*8
    std::vector<float> mats_ext;  // M*4: Blinn-Phong Ks rgb, Ks texture (-1 none)
    std::vector<Texture> textures;
    std::vector<size_t> tex_offsets;  // byte offset of each texture in the device pool (after build)
    int env_tex = -1;
    float env_gain = 1.0f;
    unsigned features = 0;        // material features present: 1 = GGX dielectric, 2 = albedo textures
    int short_stack = 16;         // LDS stack entries of the first trace launch
    int trace_order = 3;          // ray order of the wavefront's trace launches (sort_by_direction)
    // Bumped by every change that can alter a render (materials, parameters, textures, sky,
    // build): per-sample results kept from a render serve the adjoint only at the same value.
    uint64_t generation = 0;

    // derived host data
    BvhBuild bvh;
    bool built = false;
    std::vector<LightRec> lights;
    std::vector<float> light_cdf;
    float light_sum = 0.0f;

    // device mirror
    DevBuf d_wide, d_flat, d_tris, d_shade, d_mats, d_lights, d_cdf, d_tex, d_texpool, d_lut, d_env;
    std::vector<float4> env_tab;  // environment sampler cells (SceneView::env_tab)
    SceneView view{};
    hipStream_t stream = nullptr;
    // a wavefront bounce's shadow launch runs here, beside the next bounce's closest-hit launch
    // on `stream`; shadow_ready / shadow_done order it after the shade and before the next shade
    hipStream_t stream2 = nullptr;
    hipEvent_t shadow_ready = nullptr, shadow_done = nullptr;
    int device = 0;
    RenderTarget* target = nullptr;  // lazily created render buffers

    Scene(const float* p, const float* n, const float* uv, const int32_t* mat, uint32_t ntri,
          const float* mats, uint32_t nmat);
    ~Scene();
    int n_tris() const { return int(mat.size()); }
    int add_texture(int id, int w, int h, int fmt, const void* data);
    void set_material_ext(int id, const float* ext);
    std::vector<MatRec> mat_records() const;  // device material table
    // Host SAH build (reference rule), light sampler, upload of every device array.
    void build(int nthreads);
    // Material table / texture pool (and sky sampler) uploads; on a built scene a texture or
    // material change re-uploads only these, the BVH stays.
    void check_materials();
    void upload_materials(bool check = true);
    void upload_textures(bool check = true);
    // Alias table over the environment map's cells (empty without a map or with a black one);
    // uploads it and points view.env_tab at it.
    void build_env_sampler();
};

}  // namespace nr
