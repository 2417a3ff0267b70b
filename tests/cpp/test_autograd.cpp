// RenderFunction (include/nuray.hpp), the autograd-style pair, on the GPU: Cornell box from the
// OBJ/MTL fixture, loss L = sum(image^2) / 2, backward(ctx, dL/dimage = image) against central
// differences of L with common random numbers (albedo enters each sample polynomially, so the
// difference is exact up to O(h^2)).  Then two forwards (two parameter sets) before the first
// backward: backward(ctx1) must return the gradient at ctx1's parameters, not the last forward's.
#include <cmath>
#include <cstdio>
#include "../../include/nuray.hpp"

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    try {
        nuray::Loader l;
        l.fromSceneDescription(std::string(argv[1]) + " -s 100");
        nuray::Renderer r("ptnee");
        r.prepare(l);
        nuray::Config cfg("imgw=48 imgh=36 spp=4 seed=7");
        nuray::Camera cam(40.0f, 48.0f / 36.0f, nuray::vec3(0.0f, 100.0f, 350.0f), 0.0f, 0.0f, 0.0f);
        const auto& mats = l.materials();
        std::vector<float> theta;
        for (const auto& m : mats) theta.insert(theta.end(), {m.kd[0], m.kd[1], m.kd[2], m.alpha});
        nuray::RenderContext ctx;
        std::vector<float> img = nuray::RenderFunction::forward(ctx, r, cam, cfg, theta);
        std::vector<double> g = nuray::RenderFunction::backward(ctx, img);
        auto loss = [&](const std::vector<float>& th) {
            nuray::RenderContext c2;
            std::vector<float> im = nuray::RenderFunction::forward(c2, r, cam, cfg, th);
            double s = 0;
            for (float v : im) s += 0.5 * double(v) * double(v);
            return s;
        };
        int bad = 0, checked = 0;
        for (size_t m = 0; m < mats.size(); ++m) {
            if (mats[m].type != NR_MAT_LAMBERT) continue;
            for (int c = 0; c < 3; ++c) {
                std::vector<float> tp = theta, tm = theta;
                tp[4 * m + c] += 1e-3f;
                tm[4 * m + c] -= 1e-3f;
                const double fd = (loss(tp) - loss(tm)) / (double(tp[4 * m + c]) - double(tm[4 * m + c]));
                const double ad = g[4 * m + c];
                const bool ok = std::fabs(fd - ad) <= 2e-3 * std::fabs(fd) + 1e-9;
                std::printf("material %zu channel %d: backward %.6e central difference %.6e %s\n", m, c, ad, fd,
                            ok ? "" : "MISMATCH");
                bad += !ok;
                ++checked;
            }
        }
        // interleaved: forward(ctx1, theta), forward(ctx2, theta2), backward(ctx1) == backward of a
        // fresh forward at theta (to the ~1e-7 of the FP32 per-block gradient sums)
        std::vector<float> theta2 = theta;
        for (size_t i = 0; i < theta2.size(); i += 4)
            for (int c = 0; c < 3; ++c) theta2[i + c] *= 0.6f;
        nuray::RenderContext c1, c2;
        std::vector<float> im1 = nuray::RenderFunction::forward(c1, r, cam, cfg, theta);
        std::vector<float> im2 = nuray::RenderFunction::forward(c2, r, cam, cfg, theta2);
        std::vector<double> g1 = nuray::RenderFunction::backward(c1, im1);
        std::vector<double> g2 = nuray::RenderFunction::backward(c2, im2);
        double gmax = 0, d1 = 0, d12 = 0;
        for (size_t i = 0; i < g.size(); ++i) {
            gmax = std::fmax(gmax, std::fabs(g[i]));
            d1 = std::fmax(d1, std::fabs(g1[i] - g[i]));
            d12 = std::fmax(d12, std::fabs(g2[i] - g[i]));
        }
        const bool inter_ok = d1 <= 1e-5 * gmax && d12 > 1e-2 * gmax;
        std::printf("interleaved forwards: backward(ctx1) differs from the fresh gradient by %.3e (max |g| %.3e), "
                    "backward(ctx2) by %.3e %s\n", d1, gmax, d12, inter_ok ? "" : "MISMATCH");
        bad += !inter_ok;
        std::printf("%s %d checked\n", bad || checked == 0 ? "FAILED" : "OK", checked);
        return bad || checked == 0 ? 1 : 0;
    } catch (const std::exception& e) {
        std::printf("FAILED %s\n", e.what());
        return 1;
    }
}
