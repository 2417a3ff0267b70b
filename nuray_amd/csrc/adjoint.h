// Differentiable ptnee (path-replay backpropagation) host entry points.
#pragma once
#include "render.h"

namespace nr {

// Material parameter rows {r, g, b, alpha} for every material (emitters keep their Ke).
void set_material_params(Scene& sc, const float* theta, int n_mats);
// Mean squared error of the last render's mean image against `target` (W*H*3); keeps
// dLoss/dImage on the device (and copies it to dldi_host if non-null).
double loss_l2(Scene& sc, const float* target, float* dldi_host);
// Gradient of the loss with respect to every material's {r, g, b, alpha} (materials x 4 doubles)
// and, for tex_id >= 0, with respect to the linear texel values of that albedo texture
// (w * h * 3 doubles), by replaying the samples described by cam / p.
void adjoint_ptnee(Scene& sc, const CamView& cam, const RenderParams& p, const float* dldi_host, double* grad,
                   int tex_id = -1, double* gtex = nullptr);
// Forward-only render of the P samples from sample0 on, one lane per path (RenderParams::path_kernel):
// radiance[slot] for slot = sample * W * H + pixel, exactly the wavefront's per-sample radiance.
// With vrec, also the first vrec_cap vertex records of every sample (vrec[slot * vrec_cap + depth]).
void render_paths(Scene& sc, const CamView& cam, const RenderParams& p, uint32_t sample0, int P, float4* radiance,
                  int* err, HitRec* vrec = nullptr, int vrec_cap = 0);

}  // namespace nr
