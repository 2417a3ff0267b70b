// SAH BVH construction with the split rule of src/hierarchy/bvh.cpp:5-70.
//
// Reference rule (bvh.cpp:5-58), reproduced step for step because the topology must match:
//   for axis 0, 1, 2 in sequence on the same array: std::sort by pMax[axis] (unstable), suffix
//   surface areas from the right, prefix sweep with cost = SA_pre * (i+1) + SA_suf[i+1] * (n-i),
//   strict '<' argmin (first axis / first i wins); then std::sort once more by the best axis and
//   split at ans_mid.  Leaves hold <= 5 triangles (bvh.cpp:62); left is built before right.
//
// MI355X-host design: the reference copies a pointer vector per node and recomputes pMax() per
// comparison; here each node sorts its own contiguous range of one int32 index array in place
// against precomputed per-triangle pMin/pMax keys.  std::sort's output permutation is a function
// of its comparison outcomes only, so sorting indices against identical keys reproduces the
// reference's order, ties included.  Disjoint subranges are independent, so subtrees above a
// size threshold run as OpenMP tasks; node numbering is fixed afterwards by a preorder pass,
// so the result does not depend on scheduling.  Because the ranges are recursed in place,
// the final index array is already the DFS (left-first) leaf order that GPU traversal uses to
// break exact-t ties the way bvh.cpp:119-126 does.
#include "bvh_build.h"
#include <algorithm>
#include <atomic>
#include <stdexcept>
#include <chrono>
#include <cstring>
#include <cmath>
#include <omp.h>

namespace nr {
namespace {

struct Tmp {
    int lo, hi;  // index range
    int l, r;    // child Tmp ids or -1
};

// std::min / std::max in the reference's operand order (vec3.cpp:103-111).
inline float smin(float a, float b) { return (b < a) ? b : a; }
inline float smax(float a, float b) { return (a < b) ? b : a; }

struct Builder {
    const float* p;
    std::vector<float> pmin, pmax;  // 3 per triangle, Triangle::pMin/pMax (triangle.cpp:68-76)
    std::vector<int32_t> idx;
    std::vector<Tmp> nodes;
    std::atomic<int> n_nodes{0};
    // set when no split had a cost below the initial 1e38 (surface areas that overflow FP32, or
    // NaN coordinates): the reference would recurse on an empty part forever; this builder stops
    // and reports it
    std::atomic<bool> degenerate{false};

    int alloc(int lo, int hi, int l, int r) {
        int id = n_nodes.fetch_add(1);
        nodes[id] = Tmp{lo, hi, l, r};
        return id;
    }

    static float sa(const float lo[3], const float hi[3]) {
        float d0 = hi[0] - lo[0], d1 = hi[1] - lo[1], d2 = hi[2] - lo[2];
        return (d0 * d1 + d1 * d2 + d2 * d0) * 2;
    }

    // BVH::_divide on idx[lo, hi): returns ans_mid (size of the left part).
    int divide(int lo, int hi, std::vector<float>& suffix) {
        int32_t* a = idx.data() + lo;
        const int n = hi - lo;
        int ans_axis = 0, ans_mid = 0;
        float ans_cost = float(1e38);
        suffix.resize(n);
        for (int axis = 0; axis < 3; ++axis) {
            const float* key = pmax.data() + axis;
            std::sort(a, a + n, [key](int32_t x, int32_t y) { return key[3 * x] < key[3 * y]; });
            float bl[3] = {1e18f, 1e18f, 1e18f}, bh[3] = {-1e18f, -1e18f, -1e18f};
            for (int i = n - 1; i >= 0; --i) {
                const float* mn = &pmin[3 * a[i]];
                const float* mx = &pmax[3 * a[i]];
                for (int k = 0; k < 3; ++k) { bl[k] = smin(bl[k], mn[k]); bh[k] = smax(bh[k], mx[k]); }
                suffix[i] = sa(bl, bh);
            }
            for (int k = 0; k < 3; ++k) { bl[k] = 1e18f; bh[k] = -1e18f; }
            for (int i = 0; i < n - 1; ++i) {
                const float* mn = &pmin[3 * a[i]];
                const float* mx = &pmax[3 * a[i]];
                for (int k = 0; k < 3; ++k) { bl[k] = smin(bl[k], mn[k]); bh[k] = smax(bh[k], mx[k]); }
                float pre = sa(bl, bh);
                float cost = pre * float(i + 1) + suffix[i + 1] * float(size_t(n) - size_t(i));
                if (cost < ans_cost) { ans_cost = cost; ans_mid = i + 1; ans_axis = axis; }
            }
        }
        const float* key = pmax.data() + ans_axis;
        std::sort(a, a + n, [key](int32_t x, int32_t y) { return key[3 * x] < key[3 * y]; });
        return ans_mid;
    }

    int rec(int lo, int hi) {
        if (hi - lo <= 5) return alloc(lo, hi, -1, -1);
        thread_local std::vector<float> suffix;
        int mid = lo + divide(lo, hi, suffix);
        if (mid == lo || mid == hi) {
            degenerate = true;
            return alloc(lo, hi, -1, -1);
        }
        int l, r;
        if (hi - lo > 4096) {
#pragma omp task shared(l) firstprivate(lo, mid)
            l = rec(lo, mid);
            r = rec(mid, hi);
#pragma omp taskwait
        } else {
            l = rec(lo, mid);
            r = rec(mid, hi);
        }
        return alloc(lo, hi, l, r);
    }
};

struct Box {
    float mn[3], mx[3];
};

inline void set4(float4& v, int k, float x) { (k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w) = x; }
inline void set4(int4& v, int k, int32_t x) { (k == 0 ? v.x : k == 1 ? v.y : k == 2 ? v.z : v.w) = x; }

// The 4-wide form of the tree for the persistent trace kernels.  Each Node4 starts from one binary
// node's two children and opens the internal child with the largest surface area until it has four
// (or only leaves are left).  Leaves stay the binary tree's leaves, with the same references, and
// every box is the binary node's own box, so a leaf is reached iff its exact box passes the slab
// test (its ancestors' boxes contain it) and the result of the walk is the reference's.  Node4s are
// numbered in preorder so that a subtree sits together in memory.
template <class RefOf>
void collapse_wide(const std::vector<Tmp>& nodes, const std::vector<Box>& bx, int root, RefOf ref_of, BvhBuild& out) {
    out.wide.clear();
    out.wide_depth = 0;
    if (nodes[root].l < 0) {
        out.wide_root = ref_of(root);
        return;
    }
    auto area = [&](int v) {
        const Box& B = bx[v];
        const double dx = double(B.mx[0]) - B.mn[0], dy = double(B.mx[1]) - B.mn[1], dz = double(B.mx[2]) - B.mn[2];
        return dx * dy + dy * dz + dz * dx;
    };
    struct Frame {
        int v, depth, parent_slot;  // parent_slot = parent Node4 * 4 + slot, -1 for the root
    };
    std::vector<Frame> st{{root, 1, -1}};
    while (!st.empty()) {
        const Frame f = st.back();
        st.pop_back();
        const int me = int(out.wide.size());
        out.wide.emplace_back();
        if (f.parent_slot >= 0) set4(out.wide[f.parent_slot / 4].ref, f.parent_slot % 4, me);
        out.wide_depth = std::max(out.wide_depth, f.depth);
        int c[4] = {nodes[f.v].l, nodes[f.v].r, -1, -1};
        int k = 2;
        while (k < 4) {
            int best = -1;
            double ba = -1.0;
            for (int j = 0; j < k; ++j)
                if (nodes[c[j]].l >= 0 && area(c[j]) > ba) { ba = area(c[j]); best = j; }
            if (best < 0) break;
            const int x = c[best];
            c[best] = nodes[x].l;
            c[k++] = nodes[x].r;
        }
        Node4& nd = out.wide[me];
        nd.pad = make_int4(k, 0, 0, 0);
        for (int j = 0; j < 4; ++j) {
            const bool used = j < k;
            const Box B = used ? bx[c[j]] : Box{{1e30f, 1e30f, 1e30f}, {1e30f, 1e30f, 1e30f}};
            set4(nd.lox, j, B.mn[0]); set4(nd.loy, j, B.mn[1]); set4(nd.loz, j, B.mn[2]);
            set4(nd.hix, j, B.mx[0]); set4(nd.hiy, j, B.mx[1]); set4(nd.hiz, j, B.mx[2]);
            set4(nd.ref, j, used && nodes[c[j]].l < 0 ? ref_of(c[j]) : make_leaf(0, 0));
        }
        for (int j = k - 1; j >= 0; --j)  // internal children, first slot on top of the stack
            if (nodes[c[j]].l >= 0) st.push_back({c[j], f.depth + 1, me * 4 + j});
    }
    out.wide_root = 0;
}

}  // namespace

BvhBuild build_sah(const float* p, int n, int nthreads) {
    auto t0 = std::chrono::steady_clock::now();
    BvhBuild out;
    Builder b;
    b.p = p;
    b.pmin.resize(3 * size_t(n));
    b.pmax.resize(3 * size_t(n));
    b.idx.resize(n);
    for (int i = 0; i < n; ++i) {
        const float* q = p + 9 * size_t(i);
        for (int k = 0; k < 3; ++k) {
            // min(p[0], min(p[1], p[2])) with std::min semantics
            b.pmin[3 * i + k] = smin(q[k], smin(q[3 + k], q[6 + k]));
            b.pmax[3 * i + k] = smax(q[k], smax(q[3 + k], q[6 + k]));
        }
        b.idx[i] = i;
    }
    if (n == 0) {
        for (int j = 0; j < 3; ++j) { out.root_min[j] = 1e18f; out.root_max[j] = -1e18f; }
        out.wide_root = make_leaf(0, 0);
        return out;
    }
    b.nodes.resize(2 * size_t(n));
    int root;
    [[maybe_unused]] const int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel num_threads(nt)
    {
#pragma omp single
        root = b.rec(0, n);
    }
    if (b.degenerate)
        throw std::runtime_error("SAH build: no split has a finite cost (non-finite coordinates, or a scene "
                                 "extent whose FP32 surface areas overflow)");
    const int nn = b.n_nodes.load();

    // Preorder numbering over all nodes, leaf / internal bounds (bvhnode.cpp:4-33) in a post-order pass.
    std::vector<int> pre(nn), order_nodes;
    order_nodes.reserve(nn);
    std::vector<int> depth_of(nn, 0);
    {
        std::vector<int> st{root};
        depth_of[root] = 1;
        while (!st.empty()) {
            int v = st.back();
            st.pop_back();
            pre[v] = int(order_nodes.size());
            order_nodes.push_back(v);
            out.depth = std::max(out.depth, depth_of[v]);
            if (b.nodes[v].l >= 0) {
                depth_of[b.nodes[v].l] = depth_of[b.nodes[v].r] = depth_of[v] + 1;
                st.push_back(b.nodes[v].r);
                st.push_back(b.nodes[v].l);
            }
        }
    }
    std::vector<Box> bx(nn);
    for (int k = nn - 1; k >= 0; --k) {  // reverse preorder: children before parents
        int v = order_nodes[k];
        const Tmp& t = b.nodes[v];
        Box& B = bx[v];
        if (t.l < 0) {
            for (int j = 0; j < 3; ++j) { B.mn[j] = 1e18f; B.mx[j] = -1e18f; }
            for (int s = t.lo; s < t.hi; ++s) {
                const float* q = p + 9 * size_t(b.idx[s]);
                for (int vtx = 0; vtx < 3; ++vtx)
                    for (int j = 0; j < 3; ++j) {
                        B.mn[j] = smin(B.mn[j], q[3 * vtx + j]);
                        B.mx[j] = smax(B.mx[j], q[3 * vtx + j]);
                    }
            }
        } else {
            const Box &L = bx[t.l], &R = bx[t.r];
            for (int j = 0; j < 3; ++j) { B.mn[j] = smin(L.mn[j], R.mn[j]); B.mx[j] = smax(L.mx[j], R.mx[j]); }
        }
    }
    auto ref_of = [&](int v) -> int32_t {  // leaves only
        const Tmp& t = b.nodes[v];
        return make_leaf(t.lo, t.hi - t.lo);
    };
    for (int j = 0; j < 3; ++j) { out.root_min[j] = bx[root].mn[j]; out.root_max[j] = bx[root].mx[j]; }
    out.order = b.idx;
    collapse_wide(b.nodes, bx, root, ref_of, out);
    // small trees: the leaves' own boxes in DFS order for the per-lane walks (a leaf is reached by
    // the tree walk iff its box passes the slab test, so testing every leaf box gives the same hits)
    out.flat.clear();
    if ((nn + 1) / 2 <= FLAT_LEAVES) {
        for (int v : order_nodes) {
            const Tmp& t = b.nodes[v];
            if (t.l >= 0) continue;
            const int32_t r = ref_of(v);
            float rf;
            std::memcpy(&rf, &r, 4);
            out.flat.push_back(make_float4(bx[v].mn[0], bx[v].mn[1], bx[v].mn[2], rf));
            out.flat.push_back(make_float4(bx[v].mx[0], bx[v].mx[1], bx[v].mx[2], 0.0f));
        }
    }

    out.left.resize(nn); out.right.resize(nn);
    out.leaf_first.resize(nn); out.leaf_count.resize(nn);
    out.box.resize(6 * size_t(nn));
    for (int k = 0; k < nn; ++k) {
        int v = order_nodes[k];
        const Tmp& t = b.nodes[v];
        for (int j = 0; j < 3; ++j) { out.box[6 * k + j] = bx[v].mn[j]; out.box[6 * k + 3 + j] = bx[v].mx[j]; }
        if (t.l < 0) {
            out.left[k] = out.right[k] = -1;
            out.leaf_first[k] = t.lo;
            out.leaf_count[k] = t.hi - t.lo;
        } else {
            out.left[k] = pre[t.l];
            out.right[k] = pre[t.r];
            out.leaf_first[k] = -1;
            out.leaf_count[k] = 0;
        }
    }
    out.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return out;
}

}  // namespace nr
