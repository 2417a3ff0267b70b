// Host SAH BVH builder with the reference's split rule, flattened for the GPU.
#pragma once
#include <cstdint>
#include <vector>
#include "nr_device.h"

namespace nr {

struct BvhBuild {
    // GPU layout
    std::vector<int32_t> order;    // loader triangle ids in DFS leaf order
    float root_min[3], root_max[3];
    int depth = 0;                 // nodes on the longest root-to-leaf path
    std::vector<Node4> wide;       // the tree collapsed 4 wide (Node4 index 0 = root)
    int32_t wide_root = 0;         // Node4 index 0 or a leaf reference
    int wide_depth = 0;            // Node4s on the longest root-to-leaf path
    std::vector<float4> flat;      // small trees: 2 float4 per leaf (see SceneView::flat), else empty
    // canonical preorder export over all nodes (same format as the test oracle's export)
    std::vector<int32_t> left, right, leaf_first, leaf_count;
    std::vector<float> box;        // 6 per node: min xyz, max xyz
    double build_ms = 0;
};

// p: N x 9 floats (p0, p1, p2 per triangle, loader order).  nthreads <= 0 uses all cores.
// The resulting topology depends only on the input, never on the thread count.
BvhBuild build_sah(const float* p, int n, int nthreads);

}  // namespace nr
