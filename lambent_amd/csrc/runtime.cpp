// runtime.cpp - device allocations and the scene upload (geometry, BVH,
// materials and the texel pool).
#include <hip/hip_runtime_api.h>

#include <cstring>
#include <string>
#include <vector>

#include "lmb_internal.h"

namespace lmb {

lmb_status DeviceBuffer::alloc(size_t n) {
    release();
    if (n == 0) return LMB_OK;
    hipError_t e = hipMalloc(&ptr, n);
    if (e != hipSuccess) {
        ptr = nullptr;
        (void)hipGetLastError();
        if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
            return fail(LMB_ERR_OUT_OF_MEMORY,
                        "device allocation of " + std::to_string(n) + " bytes failed");
        return fail_hip(e, "hipMalloc");
    }
    bytes = n;
    return LMB_OK;
}

lmb_status DeviceBuffer::upload(const void* host, size_t n) {
    lmb_status st = alloc(n);
    if (st || n == 0) return st;
    hipError_t e = hipMemcpy(ptr, host, n, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy");
    return LMB_OK;
}

void DeviceBuffer::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
}

lmb_status ensure_spill(SpillBuffer* sb, uint64_t threads) {
    if (sb->threads >= threads && sb->ptr) return LMB_OK;
    size_t bytes = (size_t)threads * (kStackMax - kStackLds) * sizeof(uint2);
    lmb_status st = sb->buf.alloc(bytes);
    if (st) {
        sb->threads = 0;
        sb->ptr = nullptr;
        return st;
    }
    sb->threads = threads;
    sb->ptr = sb->buf.as<uint2>();
    return LMB_OK;
}

namespace {

void append_image(const HostImage& img, std::vector<uint32_t>& pool,
                  std::vector<TexDesc>& table) {
    TexDesc d;
    d.offset = (uint32_t)pool.size();
    d.width = img.width;
    d.height = img.height;
    d.pad = 0;
    pool.insert(pool.end(), img.texels.begin(), img.texels.end());
    table.push_back(d);
}

HostImage solid(uint32_t rgba) {
    HostImage img;
    img.width = img.height = 1;
    img.texels.assign(1, rgba);
    return img;
}

}  // namespace

lmb_status scene_upload(lmb_scene s, const lmb_scene_desc* desc, uint64_t ntri,
                        const BvhBuildResult& bvh) {
    int dev = 0;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return fail(LMB_ERR_DEVICE, "no HIP device available");
    }
    e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail_hip(e, "hipGetDevice");
    s->device = dev;
    s->num_prims = ntri;
    if (desc->num_vertices > 0) {
        for (int k = 0; k < 3; ++k) s->bounds_lo[k] = s->bounds_hi[k] = desc->vertices[k];
        for (uint64_t v = 0; v < desc->num_vertices; ++v)
            for (int k = 0; k < 3; ++k) {
                const float c = desc->vertices[4 * v + k];  // vec4 per vertex
                s->bounds_lo[k] = c < s->bounds_lo[k] ? c : s->bounds_lo[k];
                s->bounds_hi[k] = c > s->bounds_hi[k] ? c : s->bounds_hi[k];
            }
    }

    lmb_status st = s->nodes.upload(bvh.nodes.data(), bvh.nodes.size() * sizeof(Bvh8Node));
    if (st) return st;
    st = s->tris.upload(bvh.tris.data(), bvh.tris.size() * sizeof(LeafTri));
    if (st) return st;
    st = s->prim_material.upload(desc->material_idx, ntri * sizeof(uint32_t));
    if (st) return st;
    st = s->materials.upload(desc->materials, desc->num_materials * sizeof(Material));
    if (st) return st;

    // Texel pool: diffuse[], emission[], attributes[], recolour, 6 sky faces.
    std::vector<uint32_t> pool;
    std::vector<TexDesc> table;
    const lmb_textures_s* tx = desc->textures;
    lmb_textures_s empty;
    if (!tx) tx = &empty;
    s->dev.diffuse_base = (uint32_t)table.size();
    for (const HostImage& im : tx->diffuse) append_image(im, pool, table);
    s->dev.num_diffuse = (uint32_t)tx->diffuse.size();
    s->dev.emission_base = (uint32_t)table.size();
    for (const HostImage& im : tx->emission) append_image(im, pool, table);
    s->dev.num_emission = (uint32_t)tx->emission.size();
    s->dev.attr_base = (uint32_t)table.size();
    for (const HostImage& im : tx->attributes) append_image(im, pool, table);
    s->dev.num_attr = (uint32_t)tx->attributes.size();
    s->dev.recolour = (uint32_t)table.size();
    // textures.rs:122-149: the default recolour texture is 1x1 white.
    append_image(tx->has_recolour ? tx->recolour : solid(0xFFFFFFFFu), pool, table);
    s->dev.bg_base = (uint32_t)table.size();
    for (int f = 0; f < 6; ++f)
        append_image(tx->has_background ? tx->background[f] : solid(0xFF000000u), pool, table);
    if (pool.size() >= 0xFFFFFFFFull) return fail(LMB_ERR_TEXTURE, "texel pool too large");
    st = s->texels.upload(pool.data(), pool.size() * sizeof(uint32_t));
    if (st) return st;
    st = s->tex.upload(table.data(), table.size() * sizeof(TexDesc));
    if (st) return st;

    s->dev.nodes = s->nodes.as<Bvh8Node>();
    s->dev.tris = s->tris.as<LeafTri>();
    s->dev.num_nodes = (uint32_t)bvh.nodes.size();
    s->dev.num_materials = (uint32_t)desc->num_materials;
    s->dev.prim_material = s->prim_material.as<uint32_t>();
    s->dev.materials = s->materials.as<Material>();
    s->dev.texels = s->texels.as<uint32_t>();
    s->dev.tex = s->tex.as<TexDesc>();

    s->stats.num_triangles = ntri;
    s->stats.num_nodes = bvh.nodes.size();
    s->stats.max_depth = bvh.max_depth;
    s->stats.sah_cost = bvh.sah_cost;
    s->stats.build_ms = bvh.build_ms;
    s->stats.device_bytes = s->nodes.bytes + s->tris.bytes + s->prim_material.bytes +
                            s->materials.bytes + s->texels.bytes + s->tex.bytes;
    return LMB_OK;
}

void scene_free(lmb_scene s) {
    s->nodes.release();
    s->tris.release();
    s->prim_material.release();
    s->materials.release();
    s->texels.release();
    s->tex.release();
    s->obj_tris.release();
    s->inst.reset();
    s->spill.buf.release();
    s->spill.ptr = nullptr;
    s->spill.threads = 0;
}

}  // namespace lmb
