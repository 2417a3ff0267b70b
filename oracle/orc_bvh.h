// ORACLE -- test infrastructure only (see orc_common.h).
//
// Binary SAH BVH restated from src/hierarchy/bvh.cpp and bvhnode.cpp: same split rule, same
// sort sequence (std::sort, unstable), same leaf rule, same recursive closest-hit traversal
// with no distance culling.
#pragma once
#include <atomic>
#include "orc_scene.h"
#include <tuple>

namespace orc {

struct Hit {
    float t = 1e18f, b1 = 0, b2 = 0;
    int tri = -1;
};

class BVH {
public:
    struct Node {
        V3f mn, mx;
        int l = -1, r = -1;     // children (node indices) or -1 for a leaf
        std::vector<int> prims; // leaf primitives, reference order
    };
    std::vector<Node> nodes;
    int root = -1;
    const Scene* sc = nullptr;
    mutable std::atomic<bool> degenerate{false};   // a split with no finite cost was met

    // BVH::build(const std::vector<Triangle>&), bvh.cpp:91-97 (scene order in, recursion :60-70).
    // Subtrees are independent, so they are built as OpenMP tasks (above 4096 primitives) and
    // then flattened children-first, as the sequential recursion would store them.
    void build(const Scene& s) {
        sc = &s;
        degenerate = false;
        nodes.clear();
        std::vector<int> all(s.tris.size());
        for (size_t i = 0; i < all.size(); ++i) all[i] = int(i);
        if (all.empty()) {
            root = -1;
            return;
        }
        TNode* t = nullptr;
#pragma omp parallel
#pragma omp single
        t = build_tree(all);
        root = flatten(t);
    }

    // BVH::_traversal, bvh.cpp:99-127 (recursive, left then right, right wins ties).
    Hit intersect(const V3f& o, const V3f& d) const {
        if (root < 0) return Hit();
        return trav(root, o, d);
    }

    // Renderer::intersect's commented-out brute-force loop, renderer.cpp:12-25.
    Hit brute(const V3f& o, const V3f& d) const {
        Hit h;
        for (size_t i = 0; i < sc->tris.size(); ++i) {
            const Tri& T = sc->tris[i];
            float t, b1, b2;
            tri_test_f(T.p[0], T.p[1] - T.p[0], T.p[2] - T.p[0], o, d, t, b1, b2);
            if (tri_accept(t, b1, b2, h.t)) { h.t = t; h.b1 = b1; h.b2 = b2; h.tri = int(i); }
        }
        return h;
    }

    // FP64 closest hit over the same tree (same quirks, double arithmetic), for the FP64 mode.
    bool intersect_d(const V3d& o, const V3d& d, double& t, double& b1, double& b2, int& tri) const {
        t = 1e18; b1 = b2 = 0; tri = -1;
        if (root < 0) return false;
        trav_d(root, o, d, t, b1, b2, tri);
        return tri >= 0;
    }

    int depth() const { return root < 0 ? 0 : depth_rec(root); }

private:
    // BVH::_divide, bvh.cpp:5-58.
    std::pair<std::vector<int>, std::vector<int>> divide(const std::vector<int>& in) const {
        int ans_axis = 0, ans_mid = 0;
        float ans_cost = float(1e38);
        std::vector<int> prims = in;
        auto sa = [](const V3f& a, const V3f& b) {
            V3f d = b - a;
            return (d[0] * d[1] + d[1] * d[2] + d[2] * d[0]) * 2;
        };
        const size_t n = prims.size();
        for (int axis = 0; axis < 3; ++axis) {
            std::sort(prims.begin(), prims.end(), [&](int a, int b) {
                return sc->tris[a].pmax()[axis] < sc->tris[b].pmax()[axis];
            });
            std::vector<float> suffix(n);
            V3f lo(1e18f), hi(-1e18f);
            for (int i = int(n) - 1; i >= 0; --i) {
                lo = vmin(lo, sc->tris[prims[i]].pmin());
                hi = vmax(hi, sc->tris[prims[i]].pmax());
                suffix[i] = sa(lo, hi);
            }
            lo = V3f(1e18f); hi = V3f(-1e18f);
            for (int i = 0; i < int(n) - 1; ++i) {
                lo = vmin(lo, sc->tris[prims[i]].pmin());
                hi = vmax(hi, sc->tris[prims[i]].pmax());
                float pre = sa(lo, hi);
                float cost = pre * float(i + 1) + suffix[i + 1] * float(n - size_t(i));
                if (cost < ans_cost) { ans_cost = cost; ans_mid = i + 1; ans_axis = axis; }
            }
        }
        std::sort(prims.begin(), prims.end(), [&](int a, int b) {
            return sc->tris[a].pmax()[ans_axis] < sc->tris[b].pmax()[ans_axis];
        });
        return {std::vector<int>(prims.begin(), prims.begin() + ans_mid),
                std::vector<int>(prims.begin() + ans_mid, prims.end())};
    }

    struct TNode {
        V3f mn, mx;
        TNode* l = nullptr;
        TNode* r = nullptr;
        std::vector<int> prims;
    };

    // BVH::_build, bvh.cpp:60-70, with BVHNode constructors bvhnode.cpp:4-33.
    TNode* build_tree(const std::vector<int>& prims) const {
        TNode* nd = new TNode;
        if (prims.size() <= 5) {
            nd->mn = V3f(1e18f); nd->mx = V3f(-1e18f);
            for (int t : prims)
                for (int k = 0; k < 3; ++k) {
                    nd->mn = vmin(nd->mn, sc->tris[t].p[k]);
                    nd->mx = vmax(nd->mx, sc->tris[t].p[k]);
                }
            nd->prims = prims;
            return nd;
        }
        auto parts = divide(prims);
        if (parts.first.empty() || parts.second.empty()) {  // no finite-cost split: the reference
            degenerate = true;                              // would recurse forever; stop, report
            nd->mn = V3f(1e18f); nd->mx = V3f(-1e18f);
            nd->prims = prims;
            return nd;
        }
        TNode* L = nullptr;
        TNode* R = nullptr;
        if (prims.size() >= 4096) {
#pragma omp task shared(parts, L)
            L = build_tree(parts.first);
            R = build_tree(parts.second);
#pragma omp taskwait
        } else {
            L = build_tree(parts.first);
            R = build_tree(parts.second);
        }
        nd->l = L; nd->r = R;
        nd->mn = vmin(L->mn, R->mn);
        nd->mx = vmax(L->mx, R->mx);
        return nd;
    }

    int flatten(TNode* t) {
        Node nd;
        nd.mn = t->mn; nd.mx = t->mx;
        if (t->l) {
            nd.l = flatten(t->l);
            nd.r = flatten(t->r);
        } else {
            nd.prims = std::move(t->prims);
        }
        delete t;
        nodes.push_back(std::move(nd));
        return int(nodes.size()) - 1;
    }

    bool box(int i, const V3f& o, const V3f& d) const {
        const Node& nd = nodes[i];
        float mn[3] = {nd.mn.x, nd.mn.y, nd.mn.z}, mx[3] = {nd.mx.x, nd.mx.y, nd.mx.z};
        return slab_hit<float>(mn, mx, o, d);
    }

    Hit trav(int i, const V3f& o, const V3f& d) const {
        if (!box(i, o, d)) return Hit();
        const Node& nd = nodes[i];
        if (nd.l < 0) {  // BVHNode::getIntersection, bvhnode.cpp:50-66
            Hit h;
            for (int t : nd.prims) {
                const Tri& T = sc->tris[t];
                float tt, b1, b2;
                tri_test_f(T.p[0], T.p[1] - T.p[0], T.p[2] - T.p[0], o, d, tt, b1, b2);
                if (tri_accept(tt, b1, b2, h.t)) { h.t = tt; h.b1 = b1; h.b2 = b2; h.tri = t; }
            }
            return h;
        }
        Hit L = trav(nd.l, o, d), R = trav(nd.r, o, d);
        if (L.t > 0 && L.t < R.t) return L;
        return R;
    }

    void trav_d(int i, const V3d& o, const V3d& d, double& bt, double& bb1, double& bb2, int& btri) const {
        const Node& nd = nodes[i];
        double mn[3] = {nd.mn.x, nd.mn.y, nd.mn.z}, mx[3] = {nd.mx.x, nd.mx.y, nd.mx.z};
        if (!slab_hit<double>(mn, mx, o, d)) return;
        if (nd.l < 0) {
            for (int t : nd.prims) {
                const Tri& T = sc->tris[t];
                V3d p0(T.p[0]), p1(T.p[1]), p2(T.p[2]);
                double tt, b1, b2;
                tri_test_d(p0, p1 - p0, p2 - p0, o, d, tt, b1, b2);
                if (tri_accept(tt, b1, b2, bt)) { bt = tt; bb1 = b1; bb2 = b2; btri = t; }
            }
            return;
        }
        trav_d(nd.l, o, d, bt, bb1, bb2, btri);
        trav_d(nd.r, o, d, bt, bb1, bb2, btri);
    }

    int depth_rec(int i) const {
        if (nodes[i].l < 0) return 1;
        return 1 + std::max(depth_rec(nodes[i].l), depth_rec(nodes[i].r));
    }
};

}  // namespace orc
