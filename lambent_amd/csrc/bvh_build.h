// bvh_build.h - host construction of the compressed 8-wide BVH.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "lmb_layout.h"

namespace lmb {

struct BvhBuildResult {
    std::vector<Bvh8Node> nodes;  // node 0 is the root
    std::vector<LeafTri> tris;    // triangles in leaf order
    uint32_t max_depth = 0;       // number of node levels
    double sah_cost = 0.0;        // sum of node areas relative to the root
    double build_ms = 0.0;
};

// vertices: 4 floats per vertex (xyz used); indices: 3 per triangle or NULL
// for non-indexed geometry.  Returns 0, or -1 with *err set.
int build_bvh8(const float* vertices, uint64_t num_vertices, const uint32_t* indices,
               uint64_t num_triangles, BvhBuildResult* out, std::string* err);

// Slot (0-7) of each of n <= 8 children of a node covering [lo, hi]: greedy
// maximum alignment of each child's centroid offset from the node centre with
// the slot's octant direction (bit a of the slot set = +axis a), so that
// traversal in `slot ^ octant(ray)` order meets near children first.
void assign_octant_slots(const float lo[3], const float hi[3], const float (*child_lo)[3],
                         const float (*child_hi)[3], int n, int slot_of[8]);

}  // namespace lmb
