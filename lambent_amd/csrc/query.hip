// query.hip - ray-query kernels: arbitrary rays and camera primary rays, both
// through the persistent queue traversal (traverse_queue).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "dev_common.h"
#include "dev_anyhit.h"
#include "dev_traverse.h"
#include "lmb_scene_dev.h"

namespace lmb {

constexpr int kQueryBlock = 256;
constexpr int kQueryRefill = 16;

int device_cu_count() {
    static int cached_dev = -1;
    static int cached_cu = 0;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    if (dev != cached_dev) {
        int cu = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) !=
                hipSuccess ||
            cu <= 0)
            cu = 256;
        cached_cu = cu;
        cached_dev = dev;
    }
    return cached_cu;
}

uint64_t query_threads(uint64_t n) {
    uint64_t blocks = (n + kQueryBlock - 1) / kQueryBlock;
    uint64_t cap = (uint64_t)device_cu_count() * 8;
    blocks = std::max<uint64_t>(1, std::min(blocks, cap));
    return blocks * kQueryBlock;
}

struct RaySource {
    const Ray* __restrict__ rays;
    uint4* __restrict__ hits;
    __device__ __forceinline__ Ray load(uint64_t i) const { return rays[i]; }
    __device__ __forceinline__ void store(uint64_t i, const Hit& h) const {
        hits[i] = make_uint4(h.prim, __float_as_uint(h.t), __float_as_uint(h.u),
                             __float_as_uint(h.v));
    }
};

struct CameraSource {
    CameraDev cam;
    uint32_t width, height;
    float tmin, tmax;
    uint32_t* __restrict__ prim_out;
    float* __restrict__ t_out;
    __device__ __forceinline__ Ray load(uint64_t i) const {
        const uint32_t x = (uint32_t)(i % width), y = (uint32_t)(i / width);
        float o[3], d[3];
        create_ray(cam, x, y, width, height, o, d);
        Ray r;
        r.ox = o[0];
        r.oy = o[1];
        r.oz = o[2];
        r.tmin = tmin;
        r.dx = d[0];
        r.dy = d[1];
        r.dz = d[2];
        r.tmax = tmax;
        return r;
    }
    __device__ __forceinline__ void store(uint64_t i, const Hit& h) const {
        prim_out[i] = h.prim;
        t_out[i] = h.t;
    }
};

// spill_stride: stack entries per thread in `spill` (at least the BVH's
// node levels minus kStackLds).
__global__ __launch_bounds__(kQueryBlock) void k_intersect(SceneDev s, const Ray* __restrict__ rays,
                                                           uint64_t n, uint4* __restrict__ hits,
                                                           uint2* __restrict__ spill,
                                                           uint32_t spill_stride,
                                                           uint32_t* __restrict__ fetch) {
    __shared__ uint2 lds_stack[kStackLds * kQueryBlock];
    const uint64_t gtid = (uint64_t)blockIdx.x * kQueryBlock + threadIdx.x;
    ShortStack<kQueryBlock> st;
    st.lds = lds_stack + threadIdx.x;
    st.spill = spill + gtid * spill_stride;
    LMB_TRI_SHARE(kQueryBlock, ts);
    RaySource src{rays, hits};
    traverse_scene<kQueryBlock, kQueryRefill>(s, n, fetch, st, ts, src, true);
}

__global__ __launch_bounds__(kQueryBlock) void k_primary(SceneDev s, CameraDev cam,
                                                         uint32_t width, uint32_t height,
                                                         float tmin, float tmax,
                                                         uint32_t* __restrict__ prim_out,
                                                         float* __restrict__ t_out,
                                                         uint2* __restrict__ spill,
                                                         uint32_t* __restrict__ fetch) {
    __shared__ uint2 lds_stack[kStackLds * kQueryBlock];
    const uint64_t gtid = (uint64_t)blockIdx.x * kQueryBlock + threadIdx.x;
    ShortStack<kQueryBlock> st;
    st.lds = lds_stack + threadIdx.x;
    st.spill = spill + gtid * (kStackMax - kStackLds);
    LMB_TRI_SHARE(kQueryBlock, ts);
    CameraSource src{cam, width, height, tmin, tmax, prim_out, t_out};
    traverse_scene<kQueryBlock, kQueryRefill>(s, (uint64_t)width * height, fetch, st, ts, src, true);
}

hipError_t launch_intersect(const SceneDev& s, const void* rays, uint64_t n, void* hits,
                            uint2* spill, uint32_t* fetch, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipError_t e = hipMemsetAsync(fetch, 0, kFetchWords * sizeof(uint32_t), stream);
    if (e != hipSuccess) return e;
    uint64_t threads = query_threads(n);
    hipLaunchKernelGGL(k_intersect, dim3((uint32_t)(threads / kQueryBlock)), dim3(kQueryBlock),
                       0, stream, s, (const Ray*)rays, n, (uint4*)hits, spill,
                       (uint32_t)(kStackMax - kStackLds), fetch);
    return hipGetLastError();
}

// ------------------------------------------------------------------
// Allocation-free intersection.  The traversal's scratch - the fetch
// counters (first kNoAllocSpillOffset bytes) and the per-thread stack spill
// (the rest) - lives in one static device array per GPU, for callers that
// cannot allocate.  Every call on a GPU shares it, so calls that overlap on
// one GPU race.  Host code uses the array's device address from
// hipGetSymbolAddress, never its name.
constexpr uint64_t kNoAllocScratchBytes = 16ull << 20;
constexpr uint64_t kNoAllocSpillOffset = 4096;
static_assert(kFetchWords * sizeof(uint32_t) <= kNoAllocSpillOffset, "fetch counters fit");
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;

hipError_t noalloc_scratch(int device, uint8_t** out) {
    static std::mutex mu;
    static uint8_t* ptrs[kMaxDevices] = {};
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!ptrs[device]) {
        // the device must be current: each GPU holds its own copy of the array
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        ptrs[device] = static_cast<uint8_t*>(p);
    }
    *out = ptrs[device];
    return hipSuccess;
}

uint32_t noalloc_spill_stride(uint32_t bvh_levels) {
    return bvh_levels > (uint32_t)kStackLds ? bvh_levels - (uint32_t)kStackLds : 1u;
}

uint64_t noalloc_max_blocks(uint32_t bvh_levels) {
    const uint64_t per_block = (uint64_t)kQueryBlock * noalloc_spill_stride(bvh_levels) * sizeof(uint2);
    return (kNoAllocScratchBytes - kNoAllocSpillOffset) / per_block;
}

uint64_t noalloc_scratch_bytes() { return kNoAllocScratchBytes; }

__global__ void k_zero_words(uint32_t* __restrict__ p, uint32_t n) {
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0u;
}

hipError_t launch_intersect_noalloc(const SceneDev& s, const void* rays, uint64_t n, void* hits,
                                    uint8_t* scratch_dev, uint32_t bvh_levels,
                                    hipStream_t stream) {
    if (n == 0) return hipSuccess;
    const uint64_t fit = noalloc_max_blocks(bvh_levels);
    if (fit == 0) return hipErrorInvalidValue;
    const uint64_t blocks = std::min<uint64_t>(query_threads(n) / kQueryBlock, fit);
    uint32_t* fetch = reinterpret_cast<uint32_t*>(scratch_dev);
    uint2* spill = reinterpret_cast<uint2*>(scratch_dev + kNoAllocSpillOffset);
    hipLaunchKernelGGL(k_zero_words, dim3(1), dim3(256), 0, stream, fetch, kFetchWords);
    hipLaunchKernelGGL(k_intersect, dim3((uint32_t)blocks), dim3(kQueryBlock), 0, stream, s,
                       (const Ray*)rays, n, (uint4*)hits, spill, noalloc_spill_stride(bvh_levels),
                       fetch);
    return hipGetLastError();
}

hipError_t launch_primary(const SceneDev& s, const CameraDev& cam, uint32_t width,
                          uint32_t height, float tmin, float tmax, uint32_t* prim, float* t,
                          uint2* spill, uint32_t* fetch, hipStream_t stream) {
    uint64_t n = (uint64_t)width * height;
    if (n == 0) return hipSuccess;
    hipError_t e = hipMemsetAsync(fetch, 0, kFetchWords * sizeof(uint32_t), stream);
    if (e != hipSuccess) return e;
    uint64_t threads = query_threads(n);
    hipLaunchKernelGGL(k_primary, dim3((uint32_t)(threads / kQueryBlock)), dim3(kQueryBlock), 0,
                       stream, s, cam, width, height, tmin, tmax, prim, t, spill, fetch);
    return hipGetLastError();
}

}  // namespace lmb
