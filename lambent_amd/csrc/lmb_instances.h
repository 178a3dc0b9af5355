// lmb_instances.h - moving the instances of an instanced scene.
//
// An instanced scene's BVH is one 8-wide tree in the traversal's format
// whose upper levels are a tree over instances and whose subtrees are
// instance-pure: every node below an instance's root holds only that
// instance's triangles, in world space.  Moving instances (the reference
// writes new TlasInstance transforms and rebuilds the TLAS,
// examples/cornell_box/src/main.rs:560-563) is then:
//   1. on the host, a new top tree over the instances' world boxes (small);
//   2. k_instance_place: each triangle's world-space vertices from its
//      object-space ones and its instance's new transform;
//   3. k_refit, once per node height from the leaves up: every node's child
//      boxes from its triangles or from its children's exact boxes,
//      re-quantized conservatively.
// The subtrees keep the topology built for the creation pose.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "lmb_layout.h"

namespace lmb {

// One node to recompute: its position (or kRootOfPart | instance part, whose
// position the top tree decides), and its topology.
struct RefitEntry {
    uint32_t dst;
    uint32_t child_base;
    uint32_t tri_base;
    uint8_t imask;
    uint8_t pad[3];
    uint8_t meta[8];
};
static_assert(sizeof(RefitEntry) == 24, "RefitEntry must be 24 bytes");

constexpr uint32_t kRootOfPart = 0x80000000u;

// Object-to-world transform of one instance part (3x4 row-major) and whether
// it mirrors (determinant < 0).
struct PartTransform {
    float m[12];
    uint32_t mirrored;
    uint32_t pad[3];
};
static_assert(sizeof(PartTransform) == 64, "PartTransform must be 64 bytes");

// tris[first[p] .. first[p + 1]) belong to part p (leaf order); obj holds
// three float4 per primitive (object space), indexed by the primitive id in
// each triangle's v0.w.  Writes world-space xyz of tris and the mirror flag
// into obj[3 * prim].w.
hipError_t launch_instance_place(float4* obj, LeafTri* tris, uint32_t num_tris,
                                 const uint32_t* first, uint32_t num_parts,
                                 const PartTransform* xf, hipStream_t stream);

// Recomputes nodes a[0..na) and b[0..nb) (one height level).  root_pos[p] is
// the node position of part p's root.  box: two float4 (lo, hi) per node
// position, the exact box of each node's contents.  error: set to 1 when a
// quantized box would not enclose its child (never expected).
hipError_t launch_refit(const RefitEntry* a, uint32_t na, const RefitEntry* b, uint32_t nb,
                        const uint32_t* root_pos, const LeafTri* tris, Bvh8Node* nodes,
                        float4* box, uint32_t* error, hipStream_t stream);

}  // namespace lmb
