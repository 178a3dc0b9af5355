// api_core.cpp - C ABI: errors, POD packers (materials, camera), textures,
// scenes and ray queries.  Reference interfaces are cited per function.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/lambent.h"
#include "bvh_build.h"
#include "lmb_internal.h"

namespace lmb {

static thread_local std::string g_last_error;

lmb_status fail(lmb_status code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

lmb_status fail_hip(hipError_t e, const char* what) {
    return fail(LMB_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

void clear_error() { g_last_error.clear(); }

// IEEE binary16 from binary32, round to nearest even (what half::f16::from_f32
// does, half 2.6.0, used at lib.rs:172-176).
uint16_t f32_to_f16_bits(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    uint32_t sign = (x >> 16) & 0x8000u;
    uint32_t exp = (x >> 23) & 0xFFu;
    uint32_t man = x & 0x7FFFFFu;
    if (exp == 0xFFu) {  // Inf / NaN (NaN stays a quiet NaN)
        if (man == 0) return (uint16_t)(sign | 0x7C00u);
        return (uint16_t)(sign | 0x7E00u | (man >> 13));
    }
    int e = (int)exp - 127 + 15;
    if (e >= 0x1F) return (uint16_t)(sign | 0x7C00u);  // overflow -> Inf
    if (e <= 0) {
        // subnormal half (or zero): shift the full significand
        if (e < -10) return (uint16_t)sign;
        uint32_t m = man | 0x800000u;
        int shift = 14 - e;  // 24-bit significand -> 10-bit field
        uint32_t half_m = m >> shift;
        uint32_t rem = m & ((1u << shift) - 1u);
        uint32_t halfway = 1u << (shift - 1);
        if (rem > halfway || (rem == halfway && (half_m & 1u))) half_m++;
        return (uint16_t)(sign | half_m);
    }
    uint32_t half_m = man >> 13;
    uint32_t rem = man & 0x1FFFu;
    uint32_t out = sign | ((uint32_t)e << 10) | half_m;
    if (rem > 0x1000u || (rem == 0x1000u && (half_m & 1u))) out++;  // may carry into exp
    return (uint16_t)out;
}

// bfloat16 from binary32, round to nearest even; NaN keeps its top bits and
// is made quiet (half::bf16::from_f32, used at lib.rs:103-121).
uint16_t f32_to_bf16_bits(float f) {
    uint32_t x;
    std::memcpy(&x, &f, 4);
    if ((x & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((x >> 16) | 0x0040u);
    uint32_t lsb = (x >> 16) & 1u;
    uint32_t rounded = x + 0x7FFFu + lsb;
    return (uint16_t)(rounded >> 16);
}

}  // namespace lmb

using namespace lmb;

extern "C" {

const char* lmb_last_error(void) { return g_last_error.c_str(); }

const char* lmb_version(void) { return "lambent-mi355x 0.1.0 (gfx950)"; }

lmb_status lmb_emission_scale_pack(float emissive_texture_scale, float diffuse_texture_scale,
                                   uint32_t* packed) {
    if (!packed) return fail(LMB_ERR_INVALID_ARGUMENT, "packed is NULL");
    *packed = (uint32_t)f32_to_bf16_bits(emissive_texture_scale) |
              ((uint32_t)f32_to_bf16_bits(diffuse_texture_scale) << 16);
    return LMB_OK;
}

static inline uint32_t pack2xf16(const float v[2]) {
    return (uint32_t)f32_to_f16_bits(v[0]) | ((uint32_t)f32_to_f16_bits(v[1]) << 16);
}

lmb_status lmb_material_new(const float tex_pos_1[2], const float tex_pos_2[2],
                            const float tex_pos_3[2], const uint16_t tex_pos_recolour[2],
                            uint16_t tex_idx_diffuse, int32_t tex_idx_emission,
                            int32_t tex_idx_attributes, uint32_t emission_scale_packed,
                            const float* refractive_index, uint32_t ty, lmb_material* out) {
    if (!tex_pos_1 || !tex_pos_2 || !tex_pos_3 || !tex_pos_recolour || !out)
        return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument to lmb_material_new");
    if (ty > LMB_MATERIAL_TRANSPARENT)
        return fail(LMB_ERR_INVALID_ARGUMENT, "material type must be 0, 1 or 2");
    // lib.rs:165-170 map_optional_idx: None -> u16::MAX; Some(u16::MAX) panics.
    auto map_idx = [](int32_t idx, uint16_t* o) -> lmb_status {
        if (idx < 0) {
            *o = 0xFFFFu;
            return LMB_OK;
        }
        if (idx >= 0xFFFF)
            return fail(LMB_ERR_RESERVED_INDEX,
                        "texture index " + std::to_string(idx) +
                            " is reserved or out of range (u16::MAX means None)");
        *o = (uint16_t)idx;
        return LMB_OK;
    };
    uint16_t em, at;
    lmb_status st = map_idx(tex_idx_emission, &em);
    if (st) return st;
    st = map_idx(tex_idx_attributes, &at);
    if (st) return st;
    const float vacuum[2] = {1.0f, 1.0f};
    const float* ior = refractive_index ? refractive_index : vacuum;
    out->tex_pos_1 = pack2xf16(tex_pos_1);
    out->tex_pos_2 = pack2xf16(tex_pos_2);
    out->tex_pos_3 = pack2xf16(tex_pos_3);
    out->tex_pos_recolour = (uint32_t)tex_pos_recolour[0] | ((uint32_t)tex_pos_recolour[1] << 16);
    out->tex_idx_diffuse_emission = (uint32_t)tex_idx_diffuse | ((uint32_t)em << 16);
    out->tex_idx_attributes_ty = (uint32_t)at | (ty << 16);
    out->emission_scale = emission_scale_packed;
    out->refractive_index = pack2xf16(ior);
    return LMB_OK;
}

// Inverse of a column-major 4x4 matrix by cofactors in double precision,
// rounded to f32.  A matrix whose determinant is zero or not finite is
// singular (the reference returns None from cgmath's invert, camera.rs:59-70).
static bool invert4(const float m32[16], float out[16]) {
    double m[16];
    for (int i = 0; i < 16; ++i) m[i] = m32[i];
    double inv[16];
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] +
             m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] -
             m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] +
             m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] -
              m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] -
             m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] +
             m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] -
             m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] +
              m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] +
             m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] -
             m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] +
              m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] -
              m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] -
             m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] +
             m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] -
              m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] +
              m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    if (det == 0.0 || !std::isfinite(det)) return false;
    double r = 1.0 / det;
    for (int i = 0; i < 16; ++i) {
        double v = inv[i] * r;
        if (!std::isfinite(v)) return false;
        out[i] = (float)v;
    }
    return true;
}

lmb_status lmb_camera_from_proj_view(const float proj[16], const float view[16],
                                     lmb_camera* out) {
    if (!proj || !view || !out) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    lmb_camera c;
    if (!invert4(proj, c.projection_inverse))
        return fail(LMB_ERR_SINGULAR_PROJECTION, "projection matrix is not invertible");
    if (!invert4(view, c.view_inverse))
        return fail(LMB_ERR_SINGULAR_VIEW, "view matrix is not invertible");
    *out = c;
    return LMB_OK;
}

lmb_status lmb_camera_update_from_proj_view(lmb_camera* cam, const float* proj,
                                            const float* view) {
    if (!cam) return fail(LMB_ERR_INVALID_ARGUMENT, "camera is NULL");
    // camera.rs:27-42 updates the projection first; a failing view leaves the
    // new projection in place, as the reference does.
    if (proj) {
        float inv[16];
        if (!invert4(proj, inv))
            return fail(LMB_ERR_SINGULAR_PROJECTION, "projection matrix is not invertible");
        std::memcpy(cam->projection_inverse, inv, sizeof inv);
    }
    if (view) {
        float inv[16];
        if (!invert4(view, inv))
            return fail(LMB_ERR_SINGULAR_VIEW, "view matrix is not invertible");
        std::memcpy(cam->view_inverse, inv, sizeof inv);
    }
    return LMB_OK;
}

void lmb_dispatch_size(uint32_t width, uint32_t height, uint32_t* groups_x,
                       uint32_t* groups_y) {
    if (groups_x) *groups_x = (width + 63u) / 64u;
    if (groups_y) *groups_y = height;
}

// ---------------------------------------------------------------- textures
lmb_status lmb_textures_create(lmb_textures* out) {
    if (!out) return fail(LMB_ERR_INVALID_ARGUMENT, "out is NULL");
    try {
        *out = new lmb_textures_s();
    } catch (...) {
        return fail(LMB_ERR_OUT_OF_MEMORY, "host allocation failed");
    }
    return LMB_OK;
}

static lmb_status make_image(uint32_t w, uint32_t h, const uint8_t* rgba, HostImage* img) {
    if (w == 0 || h == 0 || !rgba)
        return fail(LMB_ERR_TEXTURE, "texture must have non-zero size and data");
    if ((uint64_t)w * h > (1ull << 28)) return fail(LMB_ERR_TEXTURE, "texture too large");
    img->width = w;
    img->height = h;
    img->texels.resize((size_t)w * h);
    std::memcpy(img->texels.data(), rgba, (size_t)w * h * 4);
    return LMB_OK;
}

lmb_status lmb_textures_add(lmb_textures tex, uint32_t kind, uint32_t width, uint32_t height,
                            const uint8_t* rgba8, uint32_t* index) {
    if (!tex) return fail(LMB_ERR_INVALID_ARGUMENT, "textures is NULL");
    if (kind > LMB_TEXTURE_ATTRIBUTES) return fail(LMB_ERR_INVALID_ARGUMENT, "bad texture kind");
    HostImage img;
    lmb_status st = make_image(width, height, rgba8, &img);
    if (st) return st;
    std::vector<HostImage>& v = kind == LMB_TEXTURE_DIFFUSE    ? tex->diffuse
                                : kind == LMB_TEXTURE_EMISSION ? tex->emission
                                                               : tex->attributes;
    if (v.size() >= 0xFFFF) return fail(LMB_ERR_TEXTURE, "too many textures");
    v.push_back(std::move(img));
    if (index) *index = (uint32_t)(v.size() - 1);
    return LMB_OK;
}

lmb_status lmb_textures_set_recolour(lmb_textures tex, uint32_t width, uint32_t height,
                                     const uint8_t* rgba8) {
    if (!tex) return fail(LMB_ERR_INVALID_ARGUMENT, "textures is NULL");
    HostImage img;
    lmb_status st = make_image(width, height, rgba8, &img);
    if (st) return st;
    tex->recolour = std::move(img);
    tex->has_recolour = true;
    return LMB_OK;
}

lmb_status lmb_textures_set_background(lmb_textures tex, uint32_t size,
                                       const uint8_t* rgba8_faces) {
    if (!tex) return fail(LMB_ERR_INVALID_ARGUMENT, "textures is NULL");
    if (size == 0 || !rgba8_faces) return fail(LMB_ERR_TEXTURE, "background needs data");
    for (int f = 0; f < 6; ++f) {
        lmb_status st = make_image(size, size, rgba8_faces + (size_t)f * size * size * 4,
                                   &tex->background[f]);
        if (st) return st;
    }
    tex->has_background = true;
    return LMB_OK;
}

lmb_status lmb_textures_destroy(lmb_textures tex) {
    delete tex;
    return LMB_OK;
}

// ------------------------------------------------------------------ scenes
namespace {

lmb_status scene_create(const lmb_scene_desc* desc, lmb_scene* out) {
    if (!desc || !out) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (desc->num_vertices > 0 && !desc->vertices)
        return fail(LMB_ERR_INVALID_ARGUMENT, "vertices is NULL");
    uint64_t ntri = desc->indices ? desc->num_indices / 3 : desc->num_vertices / 3;
    if (desc->indices && desc->num_indices % 3 != 0)
        return fail(LMB_ERR_INVALID_ARGUMENT, "index count is not a multiple of 3");
    if (ntri > 0 && !desc->material_idx)
        return fail(LMB_ERR_INVALID_ARGUMENT, "material_idx is NULL");
    if (ntri >= 0xFFFFFFFFull) return fail(LMB_ERR_INVALID_ARGUMENT, "too many triangles");
    if (desc->num_materials == 0 || !desc->materials)
        return fail(LMB_ERR_INVALID_ARGUMENT, "at least one material is required");
    if (desc->flags & ~LMB_SCENE_ALPHA_TEST) return fail(LMB_ERR_INVALID_ARGUMENT, "unknown scene flags");
    for (uint64_t t = 0; t < ntri; ++t) {
        if (desc->material_idx[t] >= desc->num_materials)
            return fail(LMB_ERR_INVALID_ARGUMENT,
                        "material_idx[" + std::to_string(t) + "] is out of range");
    }
    BvhBuildResult bvh;
    std::string err;
    if (build_bvh8(desc->vertices, desc->num_vertices, desc->indices, ntri, &bvh, &err) != 0)
        return fail(LMB_ERR_INVALID_ARGUMENT, err);
    if (bvh.max_depth > (uint32_t)kStackMax)
        return fail(LMB_ERR_BVH_TOO_DEEP, "BVH depth " + std::to_string(bvh.max_depth) +
                                              " exceeds the traversal stack (" +
                                              std::to_string(kStackMax) + ")");
    // The material index of each triangle rides in its leaf record (v1.w), so
    // shading needs no separate material_idx lookup.
    for (LeafTri& lt : bvh.tris) {
        uint32_t id;
        std::memcpy(&id, &lt.v0[3], 4);
        const uint32_t mat = desc->material_idx[id];
        std::memcpy(&lt.v1[3], &mat, 4);
    }
    lmb_scene s = new lmb_scene_s();
    lmb_status st = scene_upload(s, desc, ntri, bvh);
    if (st) {
        scene_free(s);
        delete s;
        return st;
    }
    s->dev.alpha_test = (desc->flags & LMB_SCENE_ALPHA_TEST) ? 1u : 0u;
    *out = s;
    return LMB_OK;
}

}  // namespace

lmb_status lmb_scene_create(const lmb_scene_desc* desc, lmb_scene* out) {
    return scene_create(desc, out);
}

// TLAS instances (main.rs:560-563, general.wesl:1052-1059): each instance's
// geometry is transformed to world space here and all instances share one
// BVH.  The closest hit over instances is the same as the ray query's
// (rounding aside: the reference transforms the ray into object space
// instead), and traversal stays single-level.  Triangles are numbered
// instance by instance; instances whose mask has no bit in 0xFF (the ray
// query's cull mask) are left out.
lmb_status lmb_scene_destroy(lmb_scene scene) {
    if (!scene) return LMB_OK;
    scene_free(scene);
    delete scene;
    return LMB_OK;
}

lmb_status lmb_scene_bvh_stats(lmb_scene scene, lmb_bvh_stats* stats) {
    if (!scene || !stats) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    *stats = scene->stats;
    return LMB_OK;
}

lmb_status lmb_scene_intersect_device(lmb_scene scene, const void* rays_dev, uint64_t n,
                                      void* hits_dev, void* stream) {
    if (!scene) return fail(LMB_ERR_INVALID_ARGUMENT, "scene is NULL");
    if (n == 0) return LMB_OK;
    if (!rays_dev || !hits_dev) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL buffer");
    hipError_t e = hipSetDevice(scene->device);
    if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
    lmb_status st = ensure_spill(&scene->spill, query_threads(n));
    if (st) return st;
    if (!scene->fetch.ptr && (st = scene->fetch.alloc(kFetchWords * 4))) return st;
    e = launch_intersect(scene->dev, rays_dev, n, hits_dev, scene->spill.ptr,
                         scene->fetch.as<uint32_t>(), (hipStream_t)stream);
    if (e != hipSuccess) return fail_hip(e, "intersect kernel launch");
    return LMB_OK;
}

lmb_status lmb_scene_intersect_device_noalloc(lmb_scene scene, const void* rays_dev, uint64_t n,
                                              void* hits_dev, void* stream) {
    if (!scene) return fail(LMB_ERR_INVALID_ARGUMENT, "scene is NULL");
    if (n == 0) return LMB_OK;
    if (!rays_dev || !hits_dev) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL buffer");
    hipError_t e = hipSetDevice(scene->device);
    if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
    const uint32_t levels = scene->stats.max_depth;
    const uint64_t fit = noalloc_max_blocks(levels);
    if (fit < (uint64_t)device_cu_count())
        return fail(LMB_ERR_INVALID_ARGUMENT,
                    "BVH of " + std::to_string(levels) + " node levels: its stack spill for one "
                    "query block per CU exceeds the " +
                        std::to_string(noalloc_scratch_bytes() >> 20) +
                        " MiB static scratch of the allocation-free query (use "
                        "lmb_scene_intersect_device)");
    uint8_t* scratch_dev = nullptr;
    e = noalloc_scratch(scene->device, &scratch_dev);
    if (e != hipSuccess) return fail_hip(e, "hipGetSymbolAddress of the static scratch");
    e = launch_intersect_noalloc(scene->dev, rays_dev, n, hits_dev, scratch_dev, levels,
                                 (hipStream_t)stream);
    if (e != hipSuccess) return fail_hip(e, "intersect kernel launch");
    return LMB_OK;
}

lmb_status lmb_scene_intersect(lmb_scene scene, const float* rays, uint64_t n, lmb_hit* hits) {
    if (!scene) return fail(LMB_ERR_INVALID_ARGUMENT, "scene is NULL");
    if (n == 0) return LMB_OK;
    if (!rays || !hits) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL buffer");
    hipError_t e = hipSetDevice(scene->device);
    if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
    DeviceBuffer dr, dh;
    lmb_status st = dr.alloc(n * 32);
    if (st) return st;
    st = dh.alloc(n * 16);
    if (st) return st;
    e = hipMemcpy(dr.ptr, rays, n * 32, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy rays");
    st = lmb_scene_intersect_device(scene, dr.ptr, n, dh.ptr, nullptr);
    if (st) return st;
    e = hipMemcpy(hits, dh.ptr, n * 16, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy hits");
    return LMB_OK;
}

lmb_status lmb_primary_hits(lmb_scene scene, const lmb_camera* cam, uint32_t width,
                            uint32_t height, float t_min, float t_max, uint32_t* prim,
                            float* t) {
    if (!scene || !cam || !prim || !t) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    uint64_t n = (uint64_t)width * height;
    if (n == 0) return LMB_OK;
    hipError_t e = hipSetDevice(scene->device);
    if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
    DeviceBuffer dp, dt;
    lmb_status st = dp.alloc(n * 4);
    if (st) return st;
    st = dt.alloc(n * 4);
    if (st) return st;
    st = ensure_spill(&scene->spill, query_threads(n));
    if (st) return st;
    if (!scene->fetch.ptr && (st = scene->fetch.alloc(kFetchWords * 4))) return st;
    CameraDev cd;
    std::memcpy(&cd, cam, sizeof cd);
    e = launch_primary(scene->dev, cd, width, height, t_min, t_max, (uint32_t*)dp.ptr,
                       (float*)dt.ptr, scene->spill.ptr, scene->fetch.as<uint32_t>(), nullptr);
    if (e != hipSuccess) return fail_hip(e, "primary kernel launch");
    e = hipMemcpy(prim, dp.ptr, n * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy prim");
    e = hipMemcpy(t, dt.ptr, n * 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy t");
    return LMB_OK;
}

lmb_status lmb_sync(void* stream) {
    hipError_t e = stream ? hipStreamSynchronize((hipStream_t)stream) : hipDeviceSynchronize();
    if (e != hipSuccess) return fail_hip(e, "synchronize");
    return LMB_OK;
}

}  // extern "C"
