// lmb_scene_dev.h - device-pointer view of a scene, passed by value to the
// kernels, and the launch entry points of the ray-query kernels.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "lmb_layout.h"

namespace lmb {

struct SceneDev {
    const Bvh8Node* nodes;
    const LeafTri* tris;
    uint32_t num_nodes;
    uint32_t num_materials;
    const uint32_t* prim_material;  // material index per original primitive
    const Material* materials;
    const uint32_t* texels;         // RGBA8 texel pool
    const TexDesc* tex;             // all textures
    uint32_t diffuse_base, num_diffuse;
    uint32_t emission_base, num_emission;
    uint32_t attr_base, num_attr;
    uint32_t recolour;              // TexDesc index of the recolour texture
    uint32_t bg_base;               // six TexDesc entries: +x -x +y -y +z -z
    uint32_t alpha_test;            // non-opaque geometry: any-hit alpha test
    // Instanced scenes with a non-identity transform: per primitive the
    // object-space triangle (3 float4, x of the fourth component of the first
    // = 1 when the instance transform mirrors), else NULL.
    const float4* obj_tris;
};

// Number of compute units of the current device (cached).
int device_cu_count();

// Threads launched by the persistent ray-query kernels for n rays.
uint64_t query_threads(uint64_t n);

// rays: n x 32 bytes (origin, tmin, direction, tmax); hits: n x 16 bytes
// (prim, t, u, v).  spill: query_threads(n) * (kStackMax - kStackLds) uint2.
// fetch: a device u32 the kernel uses as its ray counter (zeroed here).
hipError_t launch_intersect(const SceneDev& s, const void* rays, uint64_t n, void* hits,
                            uint2* spill, uint32_t* fetch, hipStream_t stream);

// Allocation-free intersection (query.hip): the fetch counters and stack
// spill live in a static device array per GPU.  noalloc_scratch returns its
// device address on `device`, which must be current (looked up once per
// device); noalloc_max_blocks is the number of query blocks whose spill fits
// for a BVH of `bvh_levels` node levels.
hipError_t noalloc_scratch(int device, uint8_t** out);
uint64_t noalloc_max_blocks(uint32_t bvh_levels);
uint64_t noalloc_scratch_bytes();
int device_cu_count();
hipError_t launch_intersect_noalloc(const SceneDev& s, const void* rays, uint64_t n, void* hits,
                                    uint8_t* scratch_dev, uint32_t bvh_levels,
                                    hipStream_t stream);
hipError_t launch_primary(const SceneDev& s, const CameraDev& cam, uint32_t width,
                          uint32_t height, float tmin, float tmax, uint32_t* prim,
                          float* t, uint2* spill, uint32_t* fetch, hipStream_t stream);

}  // namespace lmb
