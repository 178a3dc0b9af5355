// debug.hip - the debugging shaders (src/debug/mod.rs:1-58): the camera ray's
// closest hit only, one value per pixel written straight to the colour
// output (normal and albedo zero, w = 1), no samples, bounces or guiding.
//   front face  (front_face.wesl:52-63): front (0, 1/(t+1), 0), back
//               (1/(t+1), 0, 0), miss 0; t is the parametric hit distance.
//   reflectance (reflectance.wesl:49-57, 92-96): the hit material's IOR pair
//               unpacked from f16, (lo / 5, hi / 5, 0); miss 0.
//   tangent     (tangent.wesl:52-72, 111-122): correct_tangent of
//               normalize(v0 - v1), mapped by (d + 1) / 2; a miss maps the
//               zero vector, so it reads (0.5, 0.5, 0.5).
// The camera rays go through the wavefront's generation and traversal
// kernels (k_generate, k_extend), then one shading pass over bounce 0.
#include <hip/hip_runtime.h>

#include "dev_common.h"
#include "dev_shade.h"
#include "../../include/lambent.h"
#include "lmb_frame_dev.h"

namespace lmb {

constexpr int kDebugBlock = 256;

template <uint32_t K>
__global__ __launch_bounds__(kDebugBlock) void k_debug(SceneDev s, FrameDev f) {
    const uint64_t stride = (uint64_t)gridDim.x * kDebugBlock;
    const uint32_t count = f.counters[0];
    const float4* __restrict__ q = f.queue[0];
    for (uint64_t slot = (uint64_t)blockIdx.x * kDebugBlock + threadIdx.x; slot < count;
         slot += stride) {
        const float4 o4 = q[2 * slot];
        const uint32_t pid = __float_as_uint(o4.w);
        const uint4 h = f.hits[slot];
        V3 col = v3s(0.0f);
        if (h.x != kMissPrim) {
            const float4* tp = reinterpret_cast<const float4*>(s.tris + h.x);
            const float4 a = tp[0], b = tp[1], c = tp[2];
            V3 v0 = v3(a.x, a.y, a.z), v1 = v3(b.x, b.y, b.z), v2 = v3(c.x, c.y, c.z);
            bool mirrored = false;
            if (s.obj_tris) {  // instanced: object-space vertices, as get_vertices
                const float4* o = s.obj_tris + 3 * (uint64_t)__float_as_uint(a.w);
                const float4 oa = o[0], ob = o[1], oc = o[2];
                mirrored = oa.w != 0.0f;
                if constexpr (K == LMB_QUALITY_DEBUG_TANGENT) {
                    v0 = v3(oa.x, oa.y, oa.z);
                    v1 = v3(ob.x, ob.y, ob.z);
                }
            }
            if constexpr (K == LMB_QUALITY_DEBUG_FRONT_FACE) {
                const float4 d4 = q[2 * slot + 1];
                const V3 dir = v3(d4.x, d4.y, d4.z);
                const bool front = (dot3(dir, cross3(v1 - v0, v2 - v0)) < 0.0f) != mirrored;
                const float shade = 1.0f / (__uint_as_float(h.y) + 1.0f);
                col = front ? v3(0.0f, shade, 0.0f) : v3(shade, 0.0f, 0.0f);
            } else if constexpr (K == LMB_QUALITY_DEBUG_REFLECTANCE) {
                const Material m = s.materials[__float_as_uint(b.w)];
                col = v3(f16_to_f32(m.refractive_index & 0xFFFFu) / 5.0f,
                         f16_to_f32(m.refractive_index >> 16) / 5.0f, 0.0f);
            } else {
                const V3 t = correct_tangent(normalize3(v0 - v1));
                col = v3((t.x + 1.0f) / 2.0f, (t.y + 1.0f) / 2.0f, (t.z + 1.0f) / 2.0f);
            }
        } else if constexpr (K == LMB_QUALITY_DEBUG_TANGENT) {
            col = v3s(0.5f);  // (0 + 1) / 2
        }
        f.out_color[pid] = make_float4(col.x, col.y, col.z, 1.0f);
        f.out_normal[pid] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
        f.out_albedo[pid] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);
    }
}

hipError_t launch_debug(const SceneDev& s, const CameraDev& cam, const FrameDev& f,
                        const TraceParams& p, hipStream_t stream, KernelTimer* tm) {
    if (f.npix == 0) return hipSuccess;
    hipError_t e = launch_generate(cam, f, p, 0, stream);
    if (e != hipSuccess) return e;
    if (tm) tm->begin(kKernelExtend, stream);
    e = launch_extend(s, f, p, 0, stream);
    if (tm) tm->end(kKernelExtend, stream);
    if (e != hipSuccess) return e;
    const dim3 grid((uint32_t)(frame_threads(f.npix) / kDebugBlock)), block(kDebugBlock);
    if (tm) tm->begin(kKernelShade, stream);
    switch (p.quality) {
        case LMB_QUALITY_DEBUG_FRONT_FACE:
            hipLaunchKernelGGL(k_debug<LMB_QUALITY_DEBUG_FRONT_FACE>, grid, block, 0, stream, s, f);
            break;
        case LMB_QUALITY_DEBUG_REFLECTANCE:
            hipLaunchKernelGGL(k_debug<LMB_QUALITY_DEBUG_REFLECTANCE>, grid, block, 0, stream, s, f);
            break;
        default:
            hipLaunchKernelGGL(k_debug<LMB_QUALITY_DEBUG_TANGENT>, grid, block, 0, stream, s, f);
            break;
    }
    if (tm) tm->end(kKernelShade, stream);
    return hipGetLastError();
}

}  // namespace lmb
