// ORACLE -- test infrastructure only (see orc_common.h).  The scene handle behind liboracle's C API.
#pragma once
#include "orc_render.h"

namespace orc {
struct OScene {
    Scene sc;
    BVH bvh;
    bool built = false;
};
}  // namespace orc
