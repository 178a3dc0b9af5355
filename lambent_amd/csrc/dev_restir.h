// dev_restir.h - device restatement of the reference's reservoir algebra
// (importance_sampling/shared.wesl:63-168) and the octahedral normal
// packing (general.wesl:1293-1304, shared.wesl:233-262).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev_common.h"
#include "dev_shade.h"
#include "lmb_layout.h"

namespace lmb {

constexpr uint32_t kMaxConfidence = 0x7FFFFFFFu;

// shared.wesl:255-262 copy_sign: |dst| with the sign of src (src < 0).
__device__ __forceinline__ float copy_sign(float src, float dst) {
    float a = fabsf(dst);
    return src < 0.0f ? -a : a;
}

// WGSL pack2x16snorm: floor(0.5 + 32767 * clamp(e, -1, 1)) per component.
__device__ __forceinline__ uint32_t pack2x16snorm(float a, float b) {
    int32_t ia = (int32_t)floorf(0.5f + 32767.0f * fminf(fmaxf(a, -1.0f), 1.0f));
    int32_t ib = (int32_t)floorf(0.5f + 32767.0f * fminf(fmaxf(b, -1.0f), 1.0f));
    return ((uint32_t)ia & 0xFFFFu) | (((uint32_t)ib & 0xFFFFu) << 16);
}
// WGSL unpack2x16snorm: max(f32(i16) / 32767, -1).
__device__ __forceinline__ void unpack2x16snorm(uint32_t u, float* a, float* b) {
    *a = fmaxf((float)(int16_t)(u & 0xFFFFu) / 32767.0f, -1.0f);
    *b = fmaxf((float)(int16_t)(u >> 16) / 32767.0f, -1.0f);
}

// general.wesl:1293-1304 normalised_to_u32 (octahedral encoding).
__device__ __forceinline__ uint32_t normalised_to_u32(V3 n) {
    V3 v = n / fmaxf((fabsf(n.x) + fabsf(n.y)) + fabsf(n.z), 0.01f);
    if (v.y >= 0.0f) return pack2x16snorm(v.x, v.z);
    return pack2x16snorm(copy_sign(v.x, 1.0f - fabsf(v.z)), copy_sign(v.z, 1.0f - fabsf(v.x)));
}

// shared.wesl:233-245 u32_to_normalised
__device__ __forceinline__ V3 u32_to_normalised(uint32_t u) {
    float x, z;
    unpack2x16snorm(u, &x, &z);
    V3 v = v3(x, 1.0f - (fabsf(x) + fabsf(z)), z);
    if (v.y < 0.0f) {
        float ox = v.x;
        v.x = copy_sign(ox, 1.0f - fabsf(v.z));
        v.z = copy_sign(v.z, 1.0f - fabsf(ox));
    }
    return normalize3(v);
}

__device__ __forceinline__ float len_sqrd(V3 v) {
    return fmaf(v.x, v.x, fmaf(v.y, v.y, v.z * v.z));
}

// shared.wesl:294-312
__device__ __forceinline__ float safe_div(float n, float d) {
    return near_zero(d) ? 0.0f : n / d;
}
__device__ __forceinline__ V3 safe_div_v3(V3 n, float d) {
    return near_zero(d) ? v3s(0.0f) : n / d;
}

__device__ __forceinline__ V3 res_out_radiance(const Reservoir& r) {
    return v3(r.out_radiance[0], r.out_radiance[1], r.out_radiance[2]);
}
// shared.wesl:160-168 p_hat = max(|out_radiance|, 1e-5)
__device__ __forceinline__ float p_hat(const Reservoir& r) {
    return fmaxf(length3(res_out_radiance(r)), 0.00001f);
}
__device__ __forceinline__ uint32_t confidence_of(uint32_t packed) {
    return packed & kMaxConfidence;
}

// shared.wesl:117-128 assign_sam_to_res: sample fields + valid bit.
__device__ __forceinline__ void assign_sample(Reservoir& dst, const Reservoir& src) {
    for (int k = 0; k < 3; ++k) {
        dst.visible_point[k] = src.visible_point[k];
        dst.sample_point[k] = src.sample_point[k];
        dst.out_radiance[k] = src.out_radiance[k];
    }
    dst.visible_normal = src.visible_normal;
    dst.sample_normal = src.sample_normal;
    dst.ty = src.ty;
    dst.roughness = src.roughness;
    dst.packed_confidence_valid = (dst.packed_confidence_valid & kMaxConfidence) | (1u << 31);
}

// shared.wesl:99-115 merge(reservoir, r, p, seed); returns picked_r.
__device__ __forceinline__ bool merge(Reservoir& res, const Reservoir& r, float p, uint32_t seed) {
    const uint32_t oc = confidence_of(r.packed_confidence_valid);
    const float w_new = p * r.W * (float)oc;
    const uint32_t conf = (confidence_of(res.packed_confidence_valid) + oc) & kMaxConfidence;
    res.packed_confidence_valid = conf | (res.packed_confidence_valid & ~kMaxConfidence);
    res.w = res.w + w_new;
    if (rand_f32(seed) * res.w < w_new) {
        assign_sample(res, r);
        return true;
    }
    return false;
}

}  // namespace lmb
