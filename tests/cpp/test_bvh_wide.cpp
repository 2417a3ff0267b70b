// The 4-wide form of the SAH tree and the flat leaf list (nuray_amd/csrc/bvh_build.cpp) against the
// binary tree they are made from, on the host:
//   * every binary leaf is the child of exactly one Node4, with the same reference and box, and
//     every other Node4 child box is the box of a binary internal node;
//   * a walk of the 4-wide tree with the reference slab test (bvhnode.cpp:35-48) reaches exactly
//     the leaves the binary recursion (bvh.cpp:99-127) reaches, for random, axis-aligned, nearly
//     axis-aligned and vertex-aimed rays;
//   * the walk's stack never holds more than 3 entries per Node4 level + 1;
//   * the flat list of small trees holds every leaf in DFS order with its box.
// Usage: test_bvh_wide  (exit status 0 when every check passes)
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "bvh_build.h"

using nr::BvhBuild;

static int g_fail = 0;
#define CHECK(c, ...)                                                       \
    do {                                                                    \
        if (!(c)) {                                                         \
            if (g_fail < 20) std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
            if (g_fail < 20) std::printf(__VA_ARGS__);                      \
            if (g_fail < 20) std::printf("\n");                             \
            ++g_fail;                                                       \
        }                                                                   \
    } while (0)

// BVHNode::intersect, bvhnode.cpp:35-48 (std::max / std::min in the reference's operand order)
static bool slab(const float mn[3], const float mx[3], const float o[3], const float d[3]) {
    float t_in = -1e18f, t_out = 1e18f;
    for (int k = 0; k < 3; ++k) {
        const float dd = d[k] + 1e-9f;
        float t1 = (mn[k] - o[k]) / dd, t2 = (mx[k] - o[k]) / dd;
        if (t1 > t2) std::swap(t1, t2);
        t_in = (t_in < t1) ? t1 : t_in;
        t_out = (t2 < t_out) ? t2 : t_out;
    }
    return t_in <= t_out && t_out >= 0.0f;
}

static float comp(const float4& v, int k) { return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w; }
static int comp(const int4& v, int k) { return k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w; }

// leaves reached by the binary recursion: first primitive of each
static void reach_binary(const BvhBuild& b, int v, const float o[3], const float d[3], std::vector<int>& out) {
    if (!slab(&b.box[6 * v], &b.box[6 * v + 3], o, d)) return;
    if (b.left[v] < 0) {
        out.push_back(b.leaf_first[v]);
        return;
    }
    reach_binary(b, b.left[v], o, d, out);
    reach_binary(b, b.right[v], o, d, out);
}

// leaves reached by the 4-wide walk (root box first, as the kernels do); returns the largest stack
static int reach_wide(const BvhBuild& b, const float o[3], const float d[3], std::vector<int>& out) {
    if (b.box.empty() || !slab(&b.box[0], &b.box[3], o, d)) return 0;
    std::vector<int32_t> st{b.wide_root};
    int max_sp = 1;
    while (!st.empty()) {
        const int32_t e = st.back();
        st.pop_back();
        if (nr::ref_is_leaf(e)) {
            if (nr::ref_count(e) > 0) out.push_back(nr::ref_first(e));
            continue;
        }
        const nr::Node4& nd = b.wide[e];
        for (int k = 0; k < 4; ++k) {
            const float mn[3] = {comp(nd.lox, k), comp(nd.loy, k), comp(nd.loz, k)};
            const float mx[3] = {comp(nd.hix, k), comp(nd.hiy, k), comp(nd.hiz, k)};
            if (slab(mn, mx, o, d)) st.push_back(comp(nd.ref, k));
        }
        max_sp = std::max(max_sp, int(st.size()));
    }
    return max_sp;
}

static void check_structure(const BvhBuild& b, const char* name) {
    const int nn = int(b.left.size());
    int n_leaves = 0;
    for (int v = 0; v < nn; ++v) n_leaves += b.left[v] < 0;
    if (nn == 1) {  // a single leaf: no Node4, the root is that leaf
        CHECK(b.wide.empty() && nr::ref_is_leaf(b.wide_root), "%s: single-leaf root", name);
        return;
    }
    // map (box bits, leaf ref) of the binary tree
    std::vector<int> seen(nn, 0);
    int used_slots = 0;
    for (size_t w = 0; w < b.wide.size(); ++w) {
        const nr::Node4& nd = b.wide[w];
        const int k_used = nd.pad.x;
        CHECK(k_used >= 2 && k_used <= 4, "%s: Node4 %zu uses %d slots", name, w, k_used);
        for (int k = 0; k < 4; ++k) {
            const int32_t r = comp(nd.ref, k);
            const float bx[6] = {comp(nd.lox, k), comp(nd.loy, k), comp(nd.loz, k),
                                 comp(nd.hix, k), comp(nd.hiy, k), comp(nd.hiz, k)};
            if (k >= k_used) {
                CHECK(r == nr::make_leaf(0, 0) && bx[0] == 1e30f && bx[5] == 1e30f, "%s: unused slot", name);
                continue;
            }
            ++used_slots;
            // find the binary node with this box (and, for leaves, this leaf range)
            int match = -1;
            for (int v = 0; v < nn && match < 0; ++v) {
                if (std::memcmp(&b.box[6 * v], bx, sizeof(bx)) != 0) continue;
                if (nr::ref_is_leaf(r)) {
                    if (b.left[v] < 0 && b.leaf_first[v] == nr::ref_first(r) && b.leaf_count[v] == nr::ref_count(r))
                        match = v;
                } else if (b.left[v] >= 0 && !seen[v]) {
                    match = v;
                }
            }
            CHECK(match >= 0, "%s: Node4 %zu slot %d matches no binary node", name, w, k);
            if (match >= 0) ++seen[match];
            if (!nr::ref_is_leaf(r)) CHECK(r > 0 && r < int(b.wide.size()), "%s: bad Node4 reference %d", name, r);
        }
    }
    for (int v = 0; v < nn; ++v)
        if (b.left[v] < 0) CHECK(seen[v] == 1, "%s: leaf node %d appears %d times", name, v, seen[v]);
    CHECK(used_slots == int(b.wide.size()) - 1 + n_leaves, "%s: %d used slots for %zu Node4s and %d leaves",
          name, used_slots, b.wide.size(), n_leaves);
    // flat list: small trees only, leaves in DFS (preorder) order with their boxes
    if (n_leaves <= nr::FLAT_LEAVES) {
        CHECK(int(b.flat.size()) == 2 * n_leaves, "%s: flat list has %zu entries for %d leaves", name,
              b.flat.size() / 2, n_leaves);
        int j = 0;
        for (int v = 0; v < nn && 2 * j + 1 < int(b.flat.size()); ++v) {
            if (b.left[v] >= 0) continue;
            const float4 lo = b.flat[2 * j], hi = b.flat[2 * j + 1];
            int32_t r;
            std::memcpy(&r, &lo.w, 4);
            CHECK(nr::ref_first(r) == b.leaf_first[v] && nr::ref_count(r) == b.leaf_count[v], "%s: flat leaf %d", name, j);
            CHECK(lo.x == b.box[6 * v] && lo.y == b.box[6 * v + 1] && lo.z == b.box[6 * v + 2] &&
                      hi.x == b.box[6 * v + 3] && hi.y == b.box[6 * v + 4] && hi.z == b.box[6 * v + 5],
                  "%s: flat box %d", name, j);
            ++j;
        }
    } else {
        CHECK(b.flat.empty(), "%s: flat list for %d leaves", name, n_leaves);
    }
}

static void check_rays(const BvhBuild& b, const std::vector<float>& p, const char* name, int n_rays, uint64_t seed) {
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<float> U(0.0f, 1.0f);
    float lo[3] = {1e30f, 1e30f, 1e30f}, hi[3] = {-1e30f, -1e30f, -1e30f};
    for (size_t i = 0; i < p.size(); ++i) {
        lo[i % 3] = std::min(lo[i % 3], p[i]);
        hi[i % 3] = std::max(hi[i % 3], p[i]);
    }
    const int ntri = int(p.size() / 9);
    int worst = 0;
    for (int r = 0; r < n_rays; ++r) {
        float o[3], d[3];
        const int mode = r % 4;
        for (int k = 0; k < 3; ++k) o[k] = lo[k] + (hi[k] - lo[k]) * (1.4f * U(g) - 0.2f);
        if (mode == 0) {  // random direction
            for (int k = 0; k < 3; ++k) d[k] = 2.0f * U(g) - 1.0f;
        } else if (mode == 1) {  // axis-aligned, exact zeros
            d[0] = d[1] = d[2] = 0.0f;
            d[g() % 3] = (g() & 1) ? 1.0f : -1.0f;
        } else if (mode == 2) {  // nearly axis-aligned: +-1e-9 and -1e-9 (dd = 0) components
            const float tiny[4] = {1e-9f, -1e-9f, 0.0f, 5e-10f};
            for (int k = 0; k < 3; ++k) d[k] = tiny[g() % 4];
            d[g() % 3] = (g() & 1) ? 1.0f : -1.0f;
        } else {  // aimed at a vertex
            const int t = int(g() % uint64_t(ntri)), v = int(g() % 3);
            for (int k = 0; k < 3; ++k) d[k] = p[9 * t + 3 * v + k] - o[k];
        }
        const float n = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        if (mode != 2 && n > 0) for (int k = 0; k < 3; ++k) d[k] /= n;
        std::vector<int> a, w;
        reach_binary(b, 0, o, d, a);
        const int sp = reach_wide(b, o, d, w);
        worst = std::max(worst, sp);
        std::sort(a.begin(), a.end());
        std::sort(w.begin(), w.end());
        CHECK(a == w, "%s: ray %d (mode %d) reaches %zu leaves in the binary tree, %zu in the 4-wide walk", name, r,
              mode, a.size(), w.size());
        CHECK(sp <= 3 * b.wide_depth + 1, "%s: stack %d above the bound %d", name, sp, 3 * b.wide_depth + 1);
    }
    std::printf("%s: %d triangles, %zu Node4s, depth %d (binary %d), %zu flat leaves, largest stack %d\n", name,
                ntri, b.wide.size(), b.wide_depth, b.depth, b.flat.size() / 2, worst);
}

static void run(const char* name, const std::vector<float>& p, int n_rays, uint64_t seed) {
    const BvhBuild b = nr::build_sah(p.data(), int(p.size() / 9), 0);
    check_structure(b, name);
    check_rays(b, p, name, n_rays, seed);
}

int main() {
    std::mt19937_64 g(7);
    std::uniform_real_distribution<float> U(0.0f, 1.0f);
    // random small triangles in a box (many leaves)
    {
        std::vector<float> p;
        for (int t = 0; t < 3000; ++t) {
            const float c[3] = {100 * U(g), 100 * U(g), 100 * U(g)};
            for (int v = 0; v < 3; ++v)
                for (int k = 0; k < 3; ++k) p.push_back(c[k] + 4 * U(g) - 2);
        }
        run("random3000", p, 20000, 1);
    }
    // heightfield grid (axis-aligned structure, shared edges)
    {
        std::vector<float> p;
        const int n = 40;
        auto h = [](int i, int j) { return float(std::sin(0.3 * i) * std::cos(0.2 * j)) * 3.0f; };
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                const float a[3] = {float(i), h(i, j), float(j)}, bq[3] = {float(i + 1), h(i + 1, j), float(j)};
                const float c[3] = {float(i + 1), h(i + 1, j + 1), float(j + 1)}, dq[3] = {float(i), h(i, j + 1), float(j + 1)};
                for (const float* q : {a, bq, c, c, dq, a}) p.insert(p.end(), q, q + 3);
            }
        run("grid40", p, 20000, 2);
    }
    // a chain: coincident triangles (every split costs the same) among a few others
    {
        std::vector<float> p;
        for (int t = 0; t < 80; ++t) p.insert(p.end(), {0, 0, 0, 1, 0, 0, 0, 1, 0});
        for (int t = 0; t < 20; ++t) {
            const float c[3] = {5 * U(g), 5 * U(g), 5 * U(g)};
            for (int v = 0; v < 3; ++v)
                for (int k = 0; k < 3; ++k) p.push_back(c[k] + U(g));
        }
        run("chain", p, 5000, 3);
    }
    // small trees (flat lists) and the single-leaf root
    for (int n : {1, 5, 6, 40, 200, 321}) {
        std::vector<float> p;
        for (int t = 0; t < n; ++t)
            for (int k = 0; k < 9; ++k) p.push_back(10 * U(g));
        char name[32];
        std::snprintf(name, sizeof(name), "small%d", n);
        run(name, p, 3000, 10 + n);
    }
    if (g_fail) std::printf("%d checks failed\n", g_fail);
    else std::printf("all checks passed\n");
    return g_fail ? 1 : 0;
}
