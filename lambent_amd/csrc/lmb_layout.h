// lmb_layout.h - data layouts in HBM shared by the host library and the
// gfx950 kernels.  Plain structs and constants only.
#pragma once

#include <stddef.h>
#include <stdint.h>

namespace lmb {

// ---------------------------------------------------------------------------
// Compressed 8-wide BVH node, 80 bytes (five 16-byte loads).
//
// Child boxes are quantized to 8 bits per plane against a per-node origin
// `p` and per-axis power-of-two scale 2^(e-127).  The builder verifies on the
// host that fmaf(q, scale, p) - the exact expression the kernel evaluates -
// encloses every child box, so the dequantized boxes are conservative.
//
// Child slots are assigned by direction octant: slot bit k set means the
// child lies on the + side of axis k.  A ray whose direction sign mask is
// `oct` (bit k = dir[k] < 0) visits slots in increasing (slot ^ oct) order,
// an approximate front-to-back order.
//
// meta[slot]: 0x00 empty; 0xFF internal child (node index child_base +
// popcount(imask & ((1 << slot) - 1))); otherwise a leaf holding
// popcount(meta >> 5) (1..3, unary) triangles starting at
// tri_base + (meta & 31).
// ---------------------------------------------------------------------------
struct alignas(16) Bvh8Node {
    float p[3];
    uint8_t e[3];
    uint8_t imask;
    uint32_t child_base;
    uint32_t tri_base;
    uint8_t meta[8];
    uint8_t qlo[3][8];
    uint8_t qhi[3][8];
};
static_assert(sizeof(Bvh8Node) == 80, "Bvh8Node must be 80 bytes");

constexpr uint8_t kMetaEmpty = 0x00;
constexpr uint8_t kMetaInternal = 0xFF;

// Traversal stack: kStackLds entries per lane live in LDS, deeper entries
// spill to a per-thread global-memory area of (kStackMax - kStackLds)
// entries.  Stack depth never exceeds the BVH depth, which the builder
// checks against kStackMax.
constexpr int kStackLds = 8;
constexpr int kStackMax = 48;

// Work distribution of the persistent traversal: a queue of rays is cut into
// kFetchSegments contiguous segments, each with its own fetch counter on a
// 256-byte line (kFetchStride words apart).  Blocks are dispatched to the 8
// XCDs in turn, so block b starts on segment b % 8 and takes the rest of the
// queue only once that segment is drained: 8 atomic streams instead of one
// queue of atomics on a single address, and each XCD's L2 sees a contiguous
// part of the queue.
#ifndef LMB_FETCH_SEGMENTS
#define LMB_FETCH_SEGMENTS 8
#endif
constexpr uint32_t kFetchSegments = LMB_FETCH_SEGMENTS;
constexpr uint32_t kFetchStride = 64;
constexpr uint32_t kFetchWords = kFetchSegments * kFetchStride;

// Triangle in BVH leaf order: three float4; v0.w holds the original
// primitive index bits, v1.w its material index (bindings.wesl:59-61).
// 48 bytes.
struct alignas(16) LeafTri {
    float v0[4];
    float v1[4];
    float v2[4];
};
static_assert(sizeof(LeafTri) == 48, "LeafTri must be 48 bytes");

// Material record, 32 bytes (lib.rs:85-94, bindings.wesl:48-57).
struct Material {
    uint32_t tex_pos_1, tex_pos_2, tex_pos_3, tex_pos_recolour;
    uint32_t tex_idx_diffuse_emission, tex_idx_attributes_ty;
    uint32_t emission_scale, refractive_index;
};
static_assert(sizeof(Material) == 32, "Material must be 32 bytes");

// Texture in the texel pool (RGBA8 texels, row-major).
struct TexDesc {
    uint32_t offset;  // first texel in the pool
    uint32_t width;
    uint32_t height;
    uint32_t pad;
};

// camera.rs:8-11 Camera: inverse projection and inverse view, column-major.
struct CameraDev {
    float projection_inverse[16];
    float view_inverse[16];
};
static_assert(sizeof(CameraDev) == 128, "CameraDev must be 128 bytes");

constexpr uint32_t kMatDiffuse = 0u;
constexpr uint32_t kMatMetallic = 1u;
constexpr uint32_t kMatTransparent = 2u;
constexpr uint32_t kMatSky = 0xFFu;
constexpr uint32_t kNoTexture = 0xFFFFu;
constexpr uint32_t kMissPrim = 0xFFFFFFFFu;

// ---------------------------------------------------------------------------
// Persistent per-frame state, WGSL storage layouts (bindings.wesl).
// ---------------------------------------------------------------------------
// bindings.wesl:3-9 MarcovChainState, 32 bytes (vec3 aligned to 16).
struct MarkovChain {
    float light_source[3];
    float mean_cosine;
    float weight_sum;
    uint32_t num_samples;
    float score;
    uint32_t pad;
};
static_assert(sizeof(MarkovChain) == 32, "MarkovChain must be 32 bytes");

// bindings.wesl:33-40 WorldMarcovChainState, 56 bytes.
struct WorldMarkovChain {
    uint32_t secondary_hash;
    uint32_t num_samples;
    uint32_t luminance[3];  // f32 bits
    uint32_t lock;
    uint32_t chain_light_source[3];  // f32 bits
    uint32_t chain_mean_cosine;
    uint32_t chain_weight_sum;
    uint32_t chain_num_samples;
    uint32_t chain_score;
    uint32_t timeout;
};
static_assert(sizeof(WorldMarkovChain) == 56, "WorldMarkovChain must be 56 bytes");

// bindings.wesl:158-169 Reservoir, 64 bytes.
struct Reservoir {
    float visible_point[3];
    uint32_t visible_normal;
    float sample_point[3];
    uint32_t sample_normal;
    float out_radiance[3];
    uint32_t ty;
    float roughness;
    float w;
    uint32_t packed_confidence_valid;
    float W;
};
static_assert(sizeof(Reservoir) == 64, "Reservoir must be 64 bytes");

// bindings.wesl:63-65 CompressedInfo, 36 bytes: cam_loc, emission, albedo.
struct CompressedInfo {
    float arr[9];
};
static_assert(sizeof(CompressedInfo) == 36, "CompressedInfo must be 36 bytes");

}  // namespace lmb
