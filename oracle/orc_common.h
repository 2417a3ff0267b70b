// ORACLE -- test infrastructure only.
//
// A CPU restatement of mollnn/nuRay's `ptnee` hot path (BVH build + traversal, triangle
// test, camera, materials, light sampler, textures, the path-tracing estimator), written
// from the reference's behaviour.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg may load it, and only as the checker: the product (libnuray) never links,
// calls or falls back to anything here.
//
// Build: oracle/Makefile -> oracle/_build/liboracle.so, compiled with -ffp-contract=off and
// no fast-math so that float arithmetic follows IEEE binary32 step by step, as the
// reference's x86-64 SSE build does (projects/nuRay.pro:13: -O3, no -mfma).
//
// The checks that pin each function against the reference's behaviour live in
// tests/test_oracle_pins.py.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <vector>

namespace orc {

// ---------------------------------------------------------------- vec3 semantics
// Follows src/math/vec3.cpp: component-wise ops, dot left-to-right (:54-57), cross (:58-61),
// normalized() = v / (|v| + 1e-8) (:70-73), min/max through std::min/std::max (:103-111).
template <class T>
struct V3 {
    T x, y, z;
    V3() : x(0), y(0), z(0) {}
    explicit V3(T s) : x(s), y(s), z(s) {}
    V3(T a, T b, T c) : x(a), y(b), z(c) {}
    template <class U>
    explicit V3(const V3<U>& o) : x(T(o.x)), y(T(o.y)), z(T(o.z)) {}
    T operator[](int i) const { return i == 0 ? x : (i == 1 ? y : z); }
    T& operator[](int i) { return i == 0 ? x : (i == 1 ? y : z); }
    V3 operator+(const V3& r) const { return {x + r.x, y + r.y, z + r.z}; }
    V3 operator-(const V3& r) const { return {x - r.x, y - r.y, z - r.z}; }
    V3 operator*(const V3& r) const { return {x * r.x, y * r.y, z * r.z}; }
    V3 operator*(T s) const { return {x * s, y * s, z * s}; }
    V3 operator/(T s) const { return {x / s, y / s, z / s}; }
    V3 operator-() const { return {-x, -y, -z}; }
    V3& operator+=(const V3& r) { x = x + r.x; y = y + r.y; z = z + r.z; return *this; }
    T dot(const V3& r) const { return x * r.x + y * r.y + z * r.z; }
    V3 cross(const V3& r) const {
        return {y * r.z - z * r.y, z * r.x - x * r.z, x * r.y - y * r.x};
    }
    T norm2() const { return x * x + y * y + z * z; }
    T norm() const { return std::sqrt(norm2()); }
    V3 normalized() const { return (*this) / (norm() + T(1e-8f)); }
};
template <class T>
inline V3<T> operator*(T s, const V3<T>& v) { return v * s; }
template <class T>
inline V3<T> vmin(const V3<T>& a, const V3<T>& b) {
    return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)};
}
template <class T>
inline V3<T> vmax(const V3<T>& a, const V3<T>& b) {
    return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)};
}
using V3f = V3<float>;
using V3d = V3<double>;

// ---------------------------------------------------------------- Philox4x32-10
// Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11.
// Replaces the reference's shared glibc rand() (src/sampler/samplerstd.cpp:7-10).
// Pinned by the Random123 known-answer vectors (tests/test_oracle_rng.py).
inline void philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = uint64_t(0xD2511F53u) * c0;
        uint64_t p1 = uint64_t(0xCD9E8D57u) * c2;
        uint32_t n0 = uint32_t(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = uint32_t(p1);
        uint32_t n2 = uint32_t(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = uint32_t(p0);
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
inline float u01(uint32_t v) { return float(v >> 8) * (1.0f / 16777216.0f); }

// Random-number layout of one camera sample (shared contract with the GPU path):
// key = (seed, 0x6E755261); counter = (pixel, sample, c2, c3) where
//   c2 = 0          : camera jitter   -> u[0], u[1]
//   c2 = 1 + depth  : vertex `depth`  -> c3 = 0: light pick, light r1, light r2, RR
//                                        c3 = 1: bsdf r1, bsdf r2, bsdf lobe, spare
struct Stream {
    uint32_t key[2];
    uint32_t pixel, sample;
    Stream(uint32_t seed, uint32_t px, uint32_t s) : pixel(px), sample(s) {
        key[0] = seed; key[1] = 0x6E755261u;
    }
    void draw(uint32_t c2, uint32_t c3, float u[4]) const {
        uint32_t ctr[4] = {pixel, sample, c2, c3}, out[4];
        philox4x32_10(ctr, key, out);
        for (int i = 0; i < 4; ++i) u[i] = u01(out[i]);
    }
};

// ---------------------------------------------------------------- triangle test
// src/scene/triangle.cpp:27-33.  q = s1.e1 + 1e-6 is evaluated in double (the literal 1e-6 is
// a double) and rounded to float on assignment.  Dot/cross follow vec3.cpp operand order.
inline void tri_test_f(const V3f& p0, const V3f& e1, const V3f& e2, const V3f& o, const V3f& d,
                       float& t, float& b1, float& b2) {
    V3f s = o - p0, s1 = d.cross(e2), s2 = s.cross(e1);
    float tn = s2.dot(e2), b1n = s1.dot(s), b2n = s2.dot(d);
    float q = float(double(s1.dot(e1)) + 1e-6);
    t = tn / q; b1 = b1n / q; b2 = b2n / q;
}
// Same test at double precision for the FP64 oracle mode.
inline void tri_test_d(const V3d& p0, const V3d& e1, const V3d& e2, const V3d& o, const V3d& d,
                       double& t, double& b1, double& b2) {
    V3d s = o - p0, s1 = d.cross(e2), s2 = s.cross(e1);
    double q = s1.dot(e1) + 1e-6;
    t = s2.dot(e2) / q; b1 = s1.dot(s) / q; b2 = s2.dot(d) / q;
}
// Acceptance rule of the leaf loop, src/hierarchy/bvhnode.cpp:57.
template <class T>
inline bool tri_accept(T t, T b1, T b2, T best) {
    return t < best && t > 0 && b1 >= 0 && b2 >= 0 && b1 + b2 <= 1;
}

// ---------------------------------------------------------------- slab test
// src/hierarchy/bvhnode.cpp:35-48: t = (bound - o) / (d + 1e-9f), swap if t1 > t2, then
// std::max / std::min in that operand order (NaN in the second operand is ignored).
template <class T>
inline bool slab_hit(const T mn[3], const T mx[3], const V3<T>& o, const V3<T>& d) {
    T t_in = T(-1e18), t_out = T(1e18);
    for (int i = 0; i < 3; ++i) {
        T dd = d[i] + T(1e-9f);
        T t1 = (mn[i] - o[i]) / dd, t2 = (mx[i] - o[i]) / dd;
        if (t1 > t2) std::swap(t1, t2);
        t_in = std::max(t_in, t1);
        t_out = std::min(t_out, t2);
    }
    return t_in <= t_out && t_out >= 0;
}

}  // namespace orc
