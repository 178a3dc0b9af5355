// restir.hip - ReSTIR GI passes and end-of-frame kernels.
//
//   k_temporal      importance_sampling/temporal_resampling.wesl:37-64
//   k_spatial       importance_sampling/spatial_resampling.wesl:35-194
//   k_eof_gi        data_buffer/end_of_frame.wesl:14-31  (confidence cap)
//   k_eof_world     end_of_frame.wesl:36-99 (timeout / radiance decay)
//   k_fill_zero     buffer clears (a kernel, so frames can be graph-captured)
//
// The spatial pass keeps no 129-entry candidate array (spatial:69): the only
// use of it is Z, the summed confidence of the centre and accepted
// neighbours, which is accumulated on the fly.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dev_restir.h"
#include "lmb_frame_dev.h"
#include "lmb_restir.h"

namespace lmb {

constexpr int kRsBlock = 256;

static uint32_t rs_blocks(uint64_t n) {
    uint64_t b = (n + kRsBlock - 1) / kRsBlock;
    uint64_t cap = (uint64_t)device_cu_count() * 16;
    return (uint32_t)std::max<uint64_t>(1, std::min(b, cap));
}

__global__ __launch_bounds__(kRsBlock) void k_temporal(FrameDev f, uint32_t seed_offset) {
    const uint64_t stride = (uint64_t)gridDim.x * kRsBlock;
    if (f.npix > (uint64_t)f.gi_count) return;  // reservoir buffers hold one pixel each
    for (uint64_t pix = (uint64_t)blockIdx.x * kRsBlock + threadIdx.x; pix < f.npix;
         pix += stride) {
        const uint32_t x = (uint32_t)pix % f.width, y = (uint32_t)pix / f.width;
        const uint32_t own_seed = rand_u32(x * y) + rand_u32(seed_offset + y);
        Reservoir Rs = f.gi[pix];
        Reservoir R = f.gi_old[pix];
        if (dot3(u32_to_normalised(R.sample_normal), u32_to_normalised(Rs.sample_normal)) < 0.95f)
            R = Reservoir{};
        if (Rs.ty == kMatSky) continue;
        const float w = p_hat(R);
        // merge(Rs, R, w): R is merged into a copy of Rs
        Reservoir M = Rs;
        merge(M, R, w, own_seed);
        const uint32_t conf = confidence_of(M.packed_confidence_valid);
        M.W = safe_div(M.w, (float)conf * p_hat(M));
        f.gi[pix] = M;
    }
}

// spatial_resampling.wesl:35-42 jacobian_determinant(q, r)
__device__ __forceinline__ float jacobian_determinant(const Reservoir& q, const Reservoir& r) {
    const V3 qv = v3(q.visible_point[0], q.visible_point[1], q.visible_point[2]);
    const V3 q1_to_q2 = v3(q.sample_point[0], q.sample_point[1], q.sample_point[2]) - qv;
    const V3 r1_to_q2 = v3(r.sample_point[0], r.sample_point[1], r.sample_point[2]) - qv;
    const V3 vn = u32_to_normalised(q.visible_normal);
    const float cos_r = dot3(normalize3(r1_to_q2), vn);
    const float cos_q = dot3(normalize3(q1_to_q2), vn);
    return safe_div(cos_r * len_sqrd(q1_to_q2), cos_q * len_sqrd(r1_to_q2));
}

__global__ __launch_bounds__(kRsBlock) void k_spatial(FrameDev f, uint32_t is_samples,
                                                      uint32_t is_space, uint32_t seed_offset) {
    const uint64_t stride = (uint64_t)gridDim.x * kRsBlock;
    if (f.npix > (uint64_t)f.gi_count) return;
    const uint32_t max_val = is_space * is_space - 1u;
    const uint32_t min_space = is_space / 2u;
    const uint32_t self_idx = min_space + min_space * is_space;
    for (uint64_t pix = (uint64_t)blockIdx.x * kRsBlock + threadIdx.x; pix < f.npix;
         pix += stride) {
        const uint32_t x = (uint32_t)pix % f.width, y = (uint32_t)pix / f.width;
        uint32_t own_seed = rand_u32(x * y) + rand_u32(seed_offset) + y;
        Reservoir Rs = f.gi[pix];
        if (Rs.ty == kMatSky) continue;
        V3 Rs_normal = u32_to_normalised(Rs.sample_normal);
        uint32_t Z = confidence_of(Rs.packed_confidence_valid);
        uint32_t accepted = 0;
        for (uint32_t i = 0; i < is_samples; ++i) {
            uint32_t rnd = rand_u32(own_seed) % max_val;
            if (rnd >= self_idx) rnd += 1u;
            const int32_t nx = (int32_t)(rnd % is_space) - (int32_t)min_space;
            const int32_t ny = (int32_t)(rnd / is_space) - (int32_t)min_space;
            own_seed = rand_u32(own_seed);
            const int32_t ox = (int32_t)x + nx, oy = (int32_t)y + ny;
            if (ox < 0 || oy < 0 || (nx == 0 && ny == 0)) continue;
            if ((uint32_t)ox >= f.width || (uint32_t)oy >= f.height) continue;
            const Reservoir Rn = f.gi[(uint64_t)ox + (uint64_t)oy * f.width];
            const float similarity = dot3(u32_to_normalised(Rn.sample_normal), Rs_normal);
            if (similarity < 0.9f || Rn.ty != Rs.ty || fabsf(Rs.roughness - Rn.roughness) > 0.1f)
                continue;
            const float jac = jacobian_determinant(Rn, Rs);
            if (Rs.ty != kMatDiffuse) continue;
            const float p_hat_dashed = p_hat(Rn) * jac;
            const bool picked = merge(Rs, Rn, p_hat_dashed, own_seed);
            own_seed = rand_u32(own_seed);
            if (picked) Rs_normal = u32_to_normalised(Rs.sample_normal);
            // p_hat() >= 1e-5 > 0, so every accepted neighbour counts toward Z
            Z += confidence_of(Rn.packed_confidence_valid);
            own_seed = rand_u32(own_seed);
            ++accepted;
        }
        const float z_p_hat = (float)Z * p_hat(Rs);
        Rs.W = safe_div(Rs.w, z_p_hat);
        const CompressedInfo info = f.info[pix];
        const V3 L = res_out_radiance(Rs) * Rs.W;
        const V3 pix_is = fma3(L, v3(info.arr[6], info.arr[7], info.arr[8]),
                               v3(info.arr[3], info.arr[4], info.arr[5]));
        // the last of the five stores at spatial:188-193 is the one that stays
        f.out_color[pix] = make_float4(pix_is.x, pix_is.y, pix_is.z, 1.0f);
    }
}

__global__ __launch_bounds__(kRsBlock) void k_eof_gi(Reservoir* __restrict__ res, uint64_t n,
                                                     uint32_t cap) {
    const uint64_t stride = (uint64_t)gridDim.x * kRsBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kRsBlock + threadIdx.x; i < n; i += stride) {
        Reservoir r = res[i];
        const uint32_t conf = confidence_of(r.packed_confidence_valid);
        if (conf > cap) {
            const float mult = (float)cap / (float)conf;
            r.w *= mult;
            r.packed_confidence_valid = (cap & kMaxConfidence) | (r.packed_confidence_valid & ~kMaxConfidence);
            res[i] = r;
        }
    }
}

// end_of_frame.wesl:36-99 with SAMPLE_MULTIPLIER 8, SAMPLE_DIVISOR 9.
__global__ __launch_bounds__(kRsBlock) void k_eof_world(WorldMarkovChain* __restrict__ w,
                                                        uint64_t n, uint32_t mul, uint32_t div) {
    const uint64_t stride = (uint64_t)gridDim.x * kRsBlock;
    const uint32_t sam_mul = mul < div ? mul : div;
    for (uint64_t i = (uint64_t)blockIdx.x * kRsBlock + threadIdx.x; i < n; i += stride) {
        WorldMarkovChain c = w[i];
        if (c.timeout == 0u) {
            w[i].secondary_hash = 0u;
            w[i].num_samples = 0u;
            w[i].luminance[0] = 0u;
            w[i].luminance[1] = 0u;
            w[i].luminance[2] = 0u;
            w[i].lock = 0u;
            continue;
        }
        const uint32_t old_n = c.num_samples;
        if (old_n != 0u) {
            const uint32_t prod = old_n * sam_mul;
            const uint32_t new_n = prod / div + ((prod % div) != 0u ? 1u : 0u);
            const float mult = (float)new_n / (float)old_n;
            for (int k = 0; k < 3; ++k)
                w[i].luminance[k] = __float_as_uint(__uint_as_float(c.luminance[k]) * mult);
            w[i].num_samples = new_n;
        }
        w[i].timeout = c.timeout - 1u;
    }
}

__global__ __launch_bounds__(kRsBlock) void k_fill_zero(uint32_t* __restrict__ p, uint64_t n4) {
    const uint64_t stride = (uint64_t)gridDim.x * kRsBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kRsBlock + threadIdx.x; i < n4; i += stride) p[i] = 0u;
}

hipError_t launch_temporal(const FrameDev& f, uint32_t seed_offset, hipStream_t stream) {
    hipLaunchKernelGGL(k_temporal, dim3(rs_blocks(f.npix)), dim3(kRsBlock), 0, stream, f,
                       seed_offset);
    return hipGetLastError();
}

hipError_t launch_spatial(const FrameDev& f, uint32_t is_samples, uint32_t is_space,
                          uint32_t seed_offset, hipStream_t stream) {
    hipLaunchKernelGGL(k_spatial, dim3(rs_blocks(f.npix)), dim3(kRsBlock), 0, stream, f,
                       is_samples, is_space, seed_offset);
    return hipGetLastError();
}

hipError_t launch_eof_gi(Reservoir* res, uint64_t n, uint32_t cap, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_eof_gi, dim3(rs_blocks(n)), dim3(kRsBlock), 0, stream, res, n, cap);
    return hipGetLastError();
}

hipError_t launch_eof_world(WorldMarkovChain* w, uint64_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_eof_world, dim3(rs_blocks(n)), dim3(kRsBlock), 0, stream, w, n, 8u, 9u);
    return hipGetLastError();
}

hipError_t launch_fill_zero(void* p, uint64_t bytes, hipStream_t stream) {
    if (bytes == 0) return hipSuccess;
    if (bytes % 4 != 0 || ((uintptr_t)p % 4) != 0) return hipErrorInvalidValue;
    const uint64_t n4 = bytes / 4;
    hipLaunchKernelGGL(k_fill_zero, dim3(rs_blocks(n4)), dim3(kRsBlock), 0, stream, (uint32_t*)p, n4);
    return hipGetLastError();
}

}  // namespace lmb
