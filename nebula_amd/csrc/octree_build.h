// octree_build.h -- host-side octree construction for the device layout.
#pragma once

#include <cstdint>
#include <vector>

#include "../../include/nbl_hip.h"
#include "nbl_device.h"

namespace nbl {

struct built_octree {
    std::vector<uint64_t> nodes;     // 8 child words per internal node, BFS order
    std::vector<tri_rec> recs;       // leaf records, BFS leaf order
    std::vector<leaf_info> leaves;   // per leaf, in breadth-first order (index in its child word)
    std::vector<leaf_rope> ropes;    // 6 per leaf
    float center[3] = { 0, 0, 0 }, halfsize[3] = { 0, 0, 0 };
    float max_extent = 0;
    uint32_t n_unique = 0, n_leaves = 0, max_depth = 0;
};

// Builds with the reference's rules (geometry/octree/octree_builder.cpp:93-352):
// insertion in input order, 1.01x inflated separating-axis overlap test in
// double, split above 16 triangles while depth < 21, duplicates (perfect
// overlaps) dropped per leaf, root = triangle AABB + 1 on each side.
// Throws std::runtime_error on invalid input.
built_octree build_octree(const nbl_triangle* tris, uint32_t n);

// The reference's linearized form (octree_builder.cpp:252-352) rebuilt from
// the device layout: octree_data and, per triangle_data position, the input
// index.
// First record of the leaf with child word w (0 for an empty leaf).
inline uint32_t leaf_first(const built_octree& o, uint64_t w)
{
    return ((w >> 32) & 0x7FFFFFFFu) ? o.leaves[(uint32_t)w].first : 0u;
}

void export_linearized(const built_octree& o, std::vector<int32_t>& octree_data,
    std::vector<uint32_t>& triangle_order);

}  // namespace nbl
