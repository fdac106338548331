#pragma once

#include <stdint.h>
#include <vector>

#include "rl_device.h"

namespace rl {

struct Bvh4Build {
    std::vector<Bvh4Node> nodes;  // nodes[0] is the root
    std::vector<uint32_t> order;  // leaf-order position -> scene-order triangle id
    int stack_bound = 0;          // worst-case traversal stack entries (near-first order)
    int depth = 0;
};

// tri9: num_tris * 9 floats (v0, v1, v2 xyz). Returns 0 on success.
int build_bvh4(const float* tri9, uint32_t num_tris, Bvh4Build& out);

}  // namespace rl
