// lmb_frame_dev.h - device view of the per-frame buffers (DataBuffers) and the
// launch entry points of the wavefront path tracer.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "lmb_layout.h"
#include "lmb_scene_dev.h"

namespace lmb {

constexpr uint32_t kMaxBounces = 64;
// counters[]: [0, kMaxBounces] queue length per bounce;
// [kCounterVertices] (64-bit, two words) path vertices of the current trace;
// [kCounterVerticesTotal] (64-bit) path vertices since the last stats read.
constexpr uint32_t kCounterVertices = kMaxBounces + 2;
constexpr uint32_t kCounterVerticesTotal = kMaxBounces + 4;
// [kCounterFetch + s * kFetchWords]: the fetch counters of the traversal of
// bounce b (s = b & 1; the other set is zeroed by that launch for the next
// bounce), 256-byte aligned.
constexpr uint32_t kCounterFetch = 128;
constexpr uint32_t kNumCounters = kCounterFetch + 2 * kFetchWords;
static_assert(kCounterVerticesTotal + 2 <= kCounterFetch, "counter layout");

__host__ __device__ inline uint32_t* fetch_set(uint32_t* counters, uint32_t bounce) {
    return counters + kCounterFetch + (bounce & 1u) * kFetchWords;
}

struct FrameDev {
    uint32_t width, height;
    uint64_t npix;
    // outputs: rgba32float images (bindings.wesl:117-122)
    float4* out_color;
    float4* out_normal;
    float4* out_albedo;
    // per pixel
    uint32_t* own_seed;
    float4* pix_sum;
    // per path (path id = pixel index)
    uint32_t* ps_seed;
    float4* ps_through;  // xyz throughput, w picked refractive-index parameter (<0: none)
    float4* ps_accum;    // xyz sample colour
    // wavefront queues: slot -> 2 float4 (origin | path id bits, direction)
    float4* queue[2];
    uint4* hits;  // slot -> (leaf, t, u, v)
    // Bounce-0 hits kept over the samples of one trace (null: not kept).
    // Primary rays carry no jitter, so every sample of a pixel has the same
    // camera ray and the same closest hit: it is traced once per trace.
    uint4* hits_primary;
    uint32_t* counters;
    uint2* spill;
    // persistent state (data_buffer/mod.rs:103-114)
    Reservoir* gi;
    Reservoir* gi_old;
    CompressedInfo* info;
    MarkovChain* mc;
    MarkovChain* mc_old;
    WorldMarkovChain* wmc;
    WorldMarkovChain* wmc_old;
    uint64_t gi_count;   // entries of gi / gi_old (1 unless spatial buffers exist)
    uint64_t mc_count;   // entries of mc / mc_old
    uint32_t world_size;
    uint32_t flags;
    // Paths that size the persistent grids of this frame's launches (0: npix);
    // a pass over a small subset of the pixels runs on a smaller grid.
    uint64_t grid_paths;
    // Samples of the plain tracer traced together (path id pixel * group +
    // j); 0 or 1: one sample at a time, path id = pixel.
    uint32_t group;
};

inline uint64_t grid_paths(const FrameDev& f) { return f.grid_paths ? f.grid_paths : f.npix; }

// Pixel of camera-ray queue slot `slot`: bands of 8 rows (the last band the
// rows left), each band column by column, so a wave's 64 rays cover an 8 x 8
// block of pixels (more coherent traversal than a 64 x 1 strip).
__host__ __device__ inline void camera_slot_pixel(uint64_t slot, uint32_t width, uint32_t height,
                                                  uint32_t& x, uint32_t& y) {
    const uint64_t band_px = 8ull * width;
    const uint32_t band = (uint32_t)(slot / band_px);
    const uint32_t r = (uint32_t)(slot - (uint64_t)band * band_px);
    const uint32_t rows = height - 8u * band < 8u ? height - 8u * band : 8u;
    x = r / rows;
    y = 8u * band + r % rows;
}

struct TraceParams {
    uint32_t quality;
    uint32_t samples;
    float tmin, tmax;
    uint32_t restir;
    uint32_t bounces;
    uint32_t rr_start;
    uint32_t guiding;
    uint32_t seed_offset;
};

// Threads of the persistent wavefront kernels for a frame of npix paths.
// The frame as bounce 0 of a trace sees it: hits in hits_primary when the
// trace has several samples and that buffer exists (the first sample traces
// the camera rays into it, later samples shade from it without tracing).
FrameDev primary_view(const FrameDev& f, const TraceParams& p);
uint64_t frame_threads(uint64_t npix);   // generation / shading grid
uint64_t extend_threads(uint64_t npix);  // persistent traversal grid (sizes the stack spill)

// Optional per-launch timing hooks (HIP events recorded on the launch stream).
enum KernelKind {
    kKernelExtend = 0,
    kKernelShade = 1,
    kKernelOther = 2,
    kKernelCacheInfer = 3,
    kKernelCacheTrain = 4,
    kNumKernelKinds = 5
};
struct KernelTimer {
    virtual void begin(int kind, hipStream_t stream) = 0;
    virtual void end(int kind, hipStream_t stream) = 0;
    virtual ~KernelTimer() = default;
};

// Single wavefront stages of the plain tracer, shared by the other modes:
// per-pixel seeds and primary rays of `sample` into queue 0, closest hits of
// the bounce-`bounce` queue, and colour = max(pix_sum / samples, 0).
hipError_t launch_generate(const CameraDev& cam, const FrameDev& f, const TraceParams& p,
                           uint32_t sample, hipStream_t stream);
hipError_t launch_extend(const SceneDev& s, const FrameDev& f, const TraceParams& p,
                         uint32_t bounce, hipStream_t stream);
hipError_t launch_resolve(const FrameDev& f, const TraceParams& p, hipStream_t stream);

// One full trace (all samples, all bounces) on `stream`.
// The example's frame averaging and RGB packing (frame_ring.hip).
hipError_t launch_ring_average(const float4* frames, uint32_t num, uint64_t npix, float* rgb,
                               hipStream_t stream);
hipError_t launch_pack_rgb(const float4* src, uint64_t npix, float* rgb, hipStream_t stream);

// The debugging shaders (debug.hip; p.quality = LMB_QUALITY_DEBUG_*).
hipError_t launch_debug(const SceneDev& s, const CameraDev& cam, const FrameDev& f,
                        const TraceParams& p, hipStream_t stream, KernelTimer* timer);
hipError_t launch_trace(const SceneDev& s, const CameraDev& cam, const FrameDev& f,
                        const TraceParams& p, hipStream_t stream, KernelTimer* timer);
// The plain tracer with `group` samples traced together: fg is the frame
// with the group's per-path state (ps_seed, ps_through, ps_accum, queues,
// hits and spill for npix * group paths) and pix_sum = each path's sample
// colour; the image, the camera queue and hits_primary are f's.  Images are
// those of launch_trace: a pixel's sample colours are summed in sample order.
hipError_t launch_trace_grouped(const SceneDev& s, const CameraDev& cam, const FrameDev& f,
                                const FrameDev& fg, uint32_t group, const TraceParams& p,
                                hipStream_t stream, KernelTimer* timer);

}  // namespace lmb
