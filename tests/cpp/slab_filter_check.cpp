// Host check of the traversal kernels' filtered slab test (nr_device.h: slab_ray / slab_filter)
// against the reference's slab test as the oracle restates it (orc::slab_hit<float>,
// bvhnode.cpp:35-48): whenever the filter reports its decision as proven ("clear"), the decision
// must equal the reference's.  The filter is restated here step for step in binary32 (std::fmaf
// is the GPU's fused multiply-add, min / max / med3 are exact), with the band's two constants as
// parameters: the relative factor on |t_in| + |t_out| and the absolute factor on
// A = max_k |o_k r_k|.
//
// Usage: slab_filter_check <cases> <seed> <rel> <abs> -> prints "cases clear violations".
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include "../../oracle/orc_common.h"

namespace {

struct Ray {
    float o[3], d[3], r[3], c[3], a2;
};

float med3(float a, float b, float c) { return std::fmax(std::fmin(a, b), std::fmin(std::fmax(a, b), c)); }

Ray slab_ray(const float o[3], const float d[3], float abs_factor) {
    Ray s;
    float amax = 0.0f;
    bool finite = true;
    for (int k = 0; k < 3; ++k) {
        s.o[k] = o[k];
        s.d[k] = d[k];
        const float dd = d[k] + 1e-9f;
        finite = finite && dd != 0.0f;
        s.r[k] = 1.0f / dd;
        s.c[k] = o[k] * s.r[k];
        amax = std::fmax(amax, std::fabs(s.c[k]));
    }
    const float a20 = std::fmax(amax * abs_factor, 1e-30f);
    s.a2 = finite ? 2.0f * a20 : INFINITY;
    return s;
}

// returns the fast decision; `clear` when it is proven
bool slab_filter(const float mn[3], const float mx[3], const Ray& s, float rel, bool& clear) {
    const float ax = std::fmaf(mn[0], s.r[0], -s.c[0]), bx = std::fmaf(mx[0], s.r[0], -s.c[0]);
    const float ay = std::fmaf(mn[1], s.r[1], -s.c[1]), by = std::fmaf(mx[1], s.r[1], -s.c[1]);
    const float az = std::fmaf(mn[2], s.r[2], -s.c[2]), bz = std::fmaf(mx[2], s.r[2], -s.c[2]);
    const float t_in = std::fmax(std::fmax(std::fmin(ax, bx), std::fmin(ay, by)), med3(az, bz, -1e18f));
    const float t_out = std::fmin(std::fmin(std::fmax(ax, bx), std::fmax(ay, by)), med3(az, bz, 1e18f));
    const float diff = t_out - t_in;
    const float band = std::fmaf(rel, std::fabs(t_in) + std::fabs(t_out), s.a2);
    clear = std::fabs(t_out) > s.a2 && std::fabs(diff) > band;
    return t_out >= 0.0f && diff > 0.0f;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 5) {
        std::fprintf(stderr, "usage: %s cases seed rel abs\n", argv[0]);
        return 2;
    }
    const long long cases = std::atoll(argv[1]);
    std::mt19937_64 rng(std::strtoull(argv[2], nullptr, 10));
    const float rel = std::strtof(argv[3], nullptr), abs_factor = std::strtof(argv[4], nullptr);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    auto u = [&]() { return U(rng); };
    const float scales[] = {1.0f, 30.0f, 1000.0f, 20000.0f};
    const float tiny[] = {0.0f, 1e-9f, -1e-9f, 1e-12f, -1e-12f, 1e-7f, -1e-7f, 1e-5f, -1e-5f, 2e-9f};
    long long clear_n = 0, viol = 0;
    for (long long i = 0; i < cases; ++i) {
        const float S = scales[rng() % 4];
        float o[3], d[3];
        for (int k = 0; k < 3; ++k) o[k] = float((2 * u() - 1) * S);
        // direction: random unit vector, some components replaced by tiny or zero values
        double v[3], n2 = 0;
        for (int k = 0; k < 3; ++k) {
            v[k] = 2 * u() - 1;
            n2 += v[k] * v[k];
        }
        for (int k = 0; k < 3; ++k) d[k] = float(v[k] / std::sqrt(n2 + 1e-300));
        const int mode = int(rng() % 8);
        if (mode == 1) d[rng() % 3] = tiny[rng() % 10];
        if (mode == 2) {
            d[0] = tiny[rng() % 10];
            d[1] = tiny[rng() % 10];
        }
        const Ray s = slab_ray(o, d, abs_factor);
        // a box: around a point of the line at t0, with random extents; then adversarial edits
        const double t0 = (u() < 0.3 ? (u() - 0.2) * 1e-2 : (u() - 0.1) * 3.0) * S;
        float p[3], mn[3], mx[3];
        for (int k = 0; k < 3; ++k) {
            p[k] = float(double(o[k]) + double(d[k]) * t0);
            const double e = u() < 0.2 ? 0.0 : u() * (u() < 0.5 ? 1e-3 : 1.0) * S;
            const double f = u() < 0.2 ? 0.0 : u() * (u() < 0.5 ? 1e-3 : 1.0) * S;
            mn[k] = float(p[k] - e);
            mx[k] = float(p[k] + f);
        }
        const int edit = int(rng() % 6);
        const int ax = int(rng() % 3);
        if (edit == 1) mx[ax] = p[ax];                        // the line touches a face from inside
        if (edit == 2) mn[ax] = p[ax];
        if (edit == 3) mn[ax] = mx[ax] = p[ax];               // a flat box through the point
        if (edit == 4) { mn[ax] = o[ax]; }                    // a face through the ray's origin
        if (edit == 5) {                                      // shifted by one ulp off the line
            mn[ax] = std::nextafter(p[ax], (rng() & 1) ? INFINITY : -INFINITY);
            mx[ax] = std::fmax(mx[ax], mn[ax]);
        }
        const bool exact = orc::slab_hit<float>(mn, mx, orc::V3<float>(o[0], o[1], o[2]),
                                                orc::V3<float>(d[0], d[1], d[2]));
        bool clear = false;
        const bool fast = slab_filter(mn, mx, s, rel, clear);
        if (clear) {
            ++clear_n;
            if (fast != exact) {
                if (viol < 5)
                    std::fprintf(stderr, "violation: o %.9g %.9g %.9g d %.9g %.9g %.9g box %.9g %.9g %.9g / %.9g %.9g %.9g "
                                 "exact %d fast %d\n", o[0], o[1], o[2], d[0], d[1], d[2], mn[0], mn[1], mn[2],
                                 mx[0], mx[1], mx[2], int(exact), int(fast));
                ++viol;
            }
        }
    }
    std::printf("%lld %lld %lld\n", cases, clear_n, viol);
    return 0;
}
