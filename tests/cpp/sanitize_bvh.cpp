// Drives the host SAH builder (nuray_amd/csrc/bvh_build.cpp) under the sanitizers: random scenes
// of 1 to 20,000 triangles built with 1 and 8 OpenMP threads must give the same tree.
#include <cstdio>
#include <random>
#include <vector>
#include "../../nuray_amd/csrc/bvh_build.h"

int main() {
    std::mt19937 g(1);
    std::uniform_real_distribution<float> u(0.0f, 100.0f), e(-3.0f, 3.0f);
    for (int n : {1, 2, 6, 7, 300, 20000}) {
        std::vector<float> p(9 * size_t(n));
        for (int i = 0; i < n; ++i) {
            const float c[3] = {u(g), u(g), u(g)};
            for (int k = 0; k < 9; ++k) p[9 * size_t(i) + k] = c[k % 3] + e(g);
        }
        const nr::BvhBuild a = nr::build_sah(p.data(), n, 1), b = nr::build_sah(p.data(), n, 8);
        if (a.order != b.order || a.left != b.left || a.right != b.right || a.box != b.box) {
            std::printf("MISMATCH at %d triangles\n", n);
            return 1;
        }
    }
    std::printf("OK\n");
    return 0;
}
