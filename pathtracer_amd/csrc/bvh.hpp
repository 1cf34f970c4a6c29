// bvh.hpp — host-built wide bounding volume hierarchy (2-, 4- or 8-wide).
//
// Built as a binary tree by binned surface-area heuristic (16 bins per axis,
// Wald 2007, "On fast construction of SAH-based bounding volume hierarchies"),
// then collapsed top-down into W-wide nodes by repeatedly opening the child
// with the largest surface area (the usual BVH2 -> BVHn collapse).
//
// Device node layout, 8*W 32-bit words per node (W = 4: 128 bytes):
//   float lo_x[W], lo_y[W], lo_z[W], hi_x[W], hi_y[W], hi_z[W];
//   int   child[W];   internal: node index; leaf: first slot in prim_refs
//   int   count[W];   0 = internal child, > 0 = leaf primitive count, -1 = empty
// Empty slots carry an inverted box (lo = +inf, hi = -inf) that no ray hits.
#pragma once

#include <cstdint>
#include <vector>

namespace pt {

struct Aabb {
    float lo[3], hi[3];
};

struct WideBvh {
    int width = 4;
    std::vector<uint32_t> words;      // node records, 8 * width words each
    std::vector<uint32_t> prim_refs;  // primitive references in leaf order
    int n_nodes = 0;
    int depth = 0;       // levels of wide nodes
    int max_stack = 0;   // exact worst-case traversal stack (far children pushed)
    bool root_leaf = false;
};

// prims[i] bounds primitive i; refs[i] is the 32-bit reference stored in the
// leaves for it. Returns the tree (an empty input gives one empty node).
WideBvh build_wide_bvh(const std::vector<Aabb>& prims, const std::vector<uint32_t>& refs, int width,
                       int max_leaf = 4);

}  // namespace pt
