// dev_anyhit.h - the any-hit test of non-opaque geometry for the queue
// traversal (general.wesl:1057-1066): a candidate triangle counts only when
// its diffuse texel's alpha at the hit's texture coordinates is not near
// zero, or (path-tracing shaders) when its material is TRANSPARENT; the
// debugging shaders test the alpha alone (front_face.wesl:82-91).  The
// texture is sampled at level 0 with the default sampler (nearest, clamp):
// the reference samples at level t, and its textures have one level.
#pragma once

#include "dev_shade.h"
#include "dev_traverse.h"

namespace lmb {

struct AlphaScene : OneScene {
    static constexpr bool kAlphaTest = true;
    const Material* materials;
    const TexDesc* tex;
    const uint32_t* texels;
    uint32_t diffuse_base;
    bool transparent_passes;
    // b: the triangle's second leaf vertex (w = material index); u, v: its
    // barycentric weights of vertices 1 and 2.
    __device__ __forceinline__ bool accept(const float4& b, float u, float v) const {
        const Material m = materials[__float_as_uint(b.w)];
        if (transparent_passes && (m.tex_idx_attributes_ty >> 16) == kMatTransparent) return true;
        float tu, tv;
        hit_tex_coords(m, u, v, &tu, &tv);
        const uint32_t texel =
            tex_texel(tex, texels, diffuse_base + (m.tex_idx_diffuse_emission & 0xFFFFu), tu, tv);
        return !near_zero((float)(texel >> 24) / 255.0f);
    }
};

__host__ __device__ inline AlphaScene alpha_scene(const SceneDev& s, bool transparent_passes) {
    AlphaScene a;
    a.nodes_ = s.nodes;
    a.tris_ = s.tris;
    a.num_nodes_ = s.num_nodes;
    a.materials = s.materials;
    a.tex = s.tex;
    a.texels = s.texels;
    a.diffuse_base = s.diffuse_base;
    a.transparent_passes = transparent_passes;
    return a;
}

// Queue traversal of one scene with its any-hit rule (uniform branch).
template <int BLOCK, int REFILL, class Source>
__device__ __forceinline__ void traverse_scene(const SceneDev& s, uint64_t count, uint32_t* fetch,
                                               ShortStack<BLOCK>& st, const TriShare& ts,
                                               Source& src, bool transparent_passes) {
    if (s.alpha_test) {
        AlphaScene sc = alpha_scene(s, transparent_passes);
        UntaggedSource<Source> us{src};
        traverse_queue_t<BLOCK>(sc, count, fetch, (uint32_t)REFILL, st, ts, us);
    } else {
        traverse_queue<BLOCK, REFILL>(s.nodes, s.tris, s.num_nodes, count, fetch, st, ts, src);
    }
}

}  // namespace lmb
