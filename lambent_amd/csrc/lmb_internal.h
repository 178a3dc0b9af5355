// lmb_internal.h - host-side objects behind the opaque C-ABI handles.
#pragma once

#include <hip/hip_runtime_api.h>

#include <string>
#include <vector>

#include "../../include/lambent.h"
#include "bvh_build.h"
#include "lmb_layout.h"
#include "lmb_scene_dev.h"

namespace lmb {

lmb_status fail(lmb_status code, const std::string& msg);
lmb_status fail_hip(hipError_t e, const char* what);

// Owning device allocation.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    lmb_status alloc(size_t n);  // frees the previous allocation first
    lmb_status upload(const void* host, size_t n);
    void release();
    template <class T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

// Per-thread traversal-stack overflow area of the persistent ray kernels.
struct SpillBuffer {
    DeviceBuffer buf;
    uint64_t threads = 0;
    uint2* ptr = nullptr;
};
lmb_status ensure_spill(SpillBuffer* sb, uint64_t threads);

struct HostImage {
    uint32_t width = 0, height = 0;
    std::vector<uint32_t> texels;  // RGBA8, r in the low byte
};

}  // namespace lmb

struct lmb_textures_s {
    std::vector<lmb::HostImage> diffuse, emission, attributes;
    lmb::HostImage recolour;
    bool has_recolour = false;
    lmb::HostImage background[6];
    bool has_background = false;
};

struct lmb_scene_s {
    int device = 0;
    lmb::SceneDev dev{};
    lmb::DeviceBuffer nodes, tris, prim_material, materials, texels, tex;
    lmb::DeviceBuffer fetch;  // ray counter of the query kernels
    lmb::SpillBuffer spill;
    lmb_bvh_stats stats{};
    uint64_t num_prims = 0;
};

namespace lmb {
lmb_status scene_upload(lmb_scene s, const lmb_scene_desc* desc, uint64_t ntri,
                        const BvhBuildResult& bvh);
void scene_free(lmb_scene s);
}  // namespace lmb
