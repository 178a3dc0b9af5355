// frame_ring.hip - the example application's frame averaging
// (examples/cornell_box/src/average_to_buf.wgsl:13-22): the colour outputs
// of the last `num` frames, each divided by num and summed in slot order,
// written as packed RGB f32 (3 floats per pixel, the staging layout of its
// denoiser buffers, main.rs:797-963).  Streaming, HBM bound.
#include <hip/hip_runtime.h>

#include "lmb_frame_dev.h"

namespace lmb {

constexpr int kRingBlock = 256;

__global__ __launch_bounds__(kRingBlock) void k_ring_average(const float4* __restrict__ frames,
                                                             uint32_t num, uint64_t npix,
                                                             float* __restrict__ rgb) {
    const uint64_t stride = (uint64_t)gridDim.x * kRingBlock;
    const float n = (float)num;
    for (uint64_t pix = (uint64_t)blockIdx.x * kRingBlock + threadIdx.x; pix < npix;
         pix += stride) {
        float x = 0.0f, y = 0.0f, z = 0.0f;
        for (uint32_t i = 0; i < num; ++i) {
            const float4 c = frames[(uint64_t)i * npix + pix];
            x = x + c.x / n;
            y = y + c.y / n;
            z = z + c.z / n;
        }
        rgb[3 * pix] = x;
        rgb[3 * pix + 1] = y;
        rgb[3 * pix + 2] = z;
    }
}

// One rgba32f image as packed RGB (the denoiser's colour / normal / albedo
// inputs).
__global__ __launch_bounds__(kRingBlock) void k_pack_rgb(const float4* __restrict__ src,
                                                         uint64_t npix, float* __restrict__ rgb) {
    const uint64_t stride = (uint64_t)gridDim.x * kRingBlock;
    for (uint64_t pix = (uint64_t)blockIdx.x * kRingBlock + threadIdx.x; pix < npix;
         pix += stride) {
        const float4 c = src[pix];
        rgb[3 * pix] = c.x;
        rgb[3 * pix + 1] = c.y;
        rgb[3 * pix + 2] = c.z;
    }
}

static dim3 ring_grid(uint64_t npix) {
    const uint64_t blocks = (npix + kRingBlock - 1) / kRingBlock;
    const uint64_t cap = (uint64_t)device_cu_count() * 8;
    return dim3((uint32_t)(blocks < cap ? (blocks ? blocks : 1) : cap));
}

hipError_t launch_ring_average(const float4* frames, uint32_t num, uint64_t npix, float* rgb,
                               hipStream_t stream) {
    if (npix == 0) return hipSuccess;
    hipLaunchKernelGGL(k_ring_average, ring_grid(npix), dim3(kRingBlock), 0, stream, frames, num,
                       npix, rgb);
    return hipGetLastError();
}

hipError_t launch_pack_rgb(const float4* src, uint64_t npix, float* rgb, hipStream_t stream) {
    if (npix == 0) return hipSuccess;
    hipLaunchKernelGGL(k_pack_rgb, ring_grid(npix), dim3(kRingBlock), 0, stream, src, npix, rgb);
    return hipGetLastError();
}

}  // namespace lmb
