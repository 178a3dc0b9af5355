// api_ring.cpp - the example application's frame ring and its average
// (examples/cornell_box/src/main.rs:377-413, 699-748; average_to_buf.wgsl):
// the colour outputs of the last `frames` traces are kept in a ring of
// device images and averaged into packed RGB f32 for display or a denoiser.
#include <hip/hip_runtime.h>

#include <cstring>

#include "../../include/lambent.h"
#include "lmb_frame_dev.h"
#include "lmb_internal.h"

using namespace lmb;

struct lmb_frame_ring_s {
    uint32_t width = 0, height = 0, frames = 0;
    uint32_t next = 0, count = 0;  // the next slot written; slots holding a frame
    DeviceBuffer ring;             // frames x width x height rgba32f
    DeviceBuffer rgb;              // width x height x 3 f32
};

extern "C" {

lmb_status lmb_frame_ring_create(uint32_t width, uint32_t height, uint32_t frames,
                                 lmb_frame_ring* out) {
    if (!out) return fail(LMB_ERR_INVALID_ARGUMENT, "out is NULL");
    if (width == 0 || height == 0 || frames == 0)
        return fail(LMB_ERR_INVALID_ARGUMENT, "width, height and frames must be non-zero");
    const uint64_t npix = (uint64_t)width * height;
    lmb_frame_ring r = new lmb_frame_ring_s();
    r->width = width;
    r->height = height;
    r->frames = frames;
    lmb_status st = r->ring.alloc(npix * 16 * frames);
    if (!st) st = r->rgb.alloc(npix * 12);
    if (!st) {
        const hipError_t e = hipMemset(r->rgb.ptr, 0, npix * 12);
        if (e != hipSuccess) st = fail_hip(e, "hipMemset");
    }
    if (st) {
        delete r;
        return st;
    }
    *out = r;
    return LMB_OK;
}

lmb_status lmb_frame_ring_destroy(lmb_frame_ring ring) {
    delete ring;
    return LMB_OK;
}

lmb_status lmb_frame_ring_push(lmb_frame_ring ring, lmb_data_buffers bufs, void* stream) {
    if (!ring || !bufs) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    void* src = nullptr;
    uint64_t bytes = 0;
    lmb_status st = lmb_data_buffers_output(bufs, LMB_OUTPUT_COLOR, &src, &bytes);
    if (st) return st;
    const uint64_t frame_bytes = (uint64_t)ring->width * ring->height * 16;
    if (bytes != frame_bytes)
        return fail(LMB_ERR_INVALID_ARGUMENT, "the buffers' size differs from the ring's");
    char* dst = static_cast<char*>(ring->ring.ptr) + (uint64_t)ring->next * frame_bytes;
    const hipError_t e =
        hipMemcpyAsync(dst, src, frame_bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpyAsync");
    ring->next = ring->next + 1 == ring->frames ? 0 : ring->next + 1;
    if (ring->count < ring->frames) ++ring->count;
    return LMB_OK;
}

lmb_status lmb_frame_ring_count(lmb_frame_ring ring, uint32_t* count) {
    if (!ring || !count) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    *count = ring->count;
    return LMB_OK;
}

lmb_status lmb_frame_ring_average(lmb_frame_ring ring, void* stream) {
    if (!ring) return fail(LMB_ERR_INVALID_ARGUMENT, "ring is NULL");
    if (ring->count == 0) return fail(LMB_ERR_INVALID_ARGUMENT, "no frame has been pushed");
    const hipError_t e = launch_ring_average(ring->ring.as<float4>(), ring->count,
                                             (uint64_t)ring->width * ring->height,
                                             ring->rgb.as<float>(), (hipStream_t)stream);
    if (e != hipSuccess) return fail_hip(e, "ring average launch");
    return LMB_OK;
}

lmb_status lmb_frame_ring_output(lmb_frame_ring ring, void** rgb_dev, uint64_t* bytes) {
    if (!ring || !rgb_dev) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    *rgb_dev = ring->rgb.ptr;
    if (bytes) *bytes = ring->rgb.bytes;
    return LMB_OK;
}

lmb_status lmb_frame_ring_read_average(lmb_frame_ring ring, float* host) {
    if (!ring || !host) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    const hipError_t e = hipMemcpy(host, ring->rgb.ptr, ring->rgb.bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy average");
    return LMB_OK;
}

lmb_status lmb_data_buffers_pack_rgb(lmb_data_buffers bufs, uint32_t which, void* rgb_dev,
                                     void* stream) {
    if (!bufs || !rgb_dev) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    void* src = nullptr;
    uint64_t bytes = 0;
    lmb_status st = lmb_data_buffers_output(bufs, which, &src, &bytes);
    if (st) return st;
    const hipError_t e = launch_pack_rgb(static_cast<const float4*>(src), bytes / 16,
                                         static_cast<float*>(rgb_dev), (hipStream_t)stream);
    if (e != hipSuccess) return fail_hip(e, "pack launch");
    return LMB_OK;
}

}  // extern "C"
