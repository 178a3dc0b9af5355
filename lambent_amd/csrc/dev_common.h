// dev_common.h - small device helpers shared by the gfx950 kernels:
// the reference's PCG hash RNG and the primary-ray construction.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lmb_layout.h"

namespace lmb {

// importance_sampling/shared.wesl:266-270 rand_u32 (PCG hash,
// Jarzynski & Olano 2020).
__device__ __forceinline__ uint32_t rand_u32(uint32_t s) {
    uint32_t state = s * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

// shared.wesl:272-284 rand_f32: f32(hash) * 2^-32, then fract (hashes that
// round to 2^32 give 0).
__device__ __forceinline__ float rand_f32(uint32_t s) {
    float f = (float)rand_u32(s) * 2.3283064365386963e-10f;
    return f - floorf(f);
}

// mat4 (column-major) times vec4, summed column by column.
__device__ __forceinline__ void mat4_mul(const float* m, float x, float y, float z,
                                         float w, float out[4]) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        out[r] = ((m[r] * x + m[4 + r] * y) + m[8 + r] * z) + m[12 + r] * w;
    }
}

// general.wesl:1158-1173 create_ray: pixel corner (no +0.5, no jitter),
// y flipped, the 4-component vector P^-1 (x, y, 1, 1) normalized as a vec4
// and its xyz rotated by the view inverse.
__device__ __forceinline__ void create_ray(const CameraDev& cam, uint32_t px, uint32_t py,
                                           uint32_t width, uint32_t height, float o[3],
                                           float d[3]) {
    float hx = (float)width / 2.0f;
    float hy = (float)height / 2.0f;
    float nx = (float)px / hx - 1.0f;
    float ny = (float)py / hy - 1.0f;
    float P[4];
    mat4_mul(cam.projection_inverse, nx, -ny, 1.0f, 1.0f, P);
    float len = sqrtf(((P[0] * P[0] + P[1] * P[1]) + P[2] * P[2]) + P[3] * P[3]);
    float q0 = P[0] / len, q1 = P[1] / len, q2 = P[2] / len;
    float D[4];
    mat4_mul(cam.view_inverse, q0, q1, q2, 0.0f, D);
    float O[4];
    mat4_mul(cam.view_inverse, 0.0f, 0.0f, 0.0f, 1.0f, O);
    d[0] = D[0];
    d[1] = D[1];
    d[2] = D[2];
    o[0] = O[0];
    o[1] = O[1];
    o[2] = O[2];
}

}  // namespace lmb
