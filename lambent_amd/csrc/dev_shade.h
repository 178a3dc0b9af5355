// dev_shade.h - gfx950 device restatement of the reference's shading math:
// vector built-ins with WGSL semantics, texture fetches, the committed-hit
// record (ray_hit), BSDF sampling and the Low/Medium/High quality hooks.
//
// Compiled with -ffp-contract=off; fma() in the WGSL is fmaf() here.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev_common.h"
#include "lmb_layout.h"
#include "lmb_scene_dev.h"

namespace lmb {

constexpr float kPi = 3.14159265358979323846f;
constexpr float kHalfPi = kPi / 2.0f;
constexpr float kTau = 2.0f * kPi;

constexpr uint32_t kQualityLow = 0, kQualityMedium = 1, kQualityHigh = 2;
// Four wavelengths per path with Cauchy dispersion (no reference counterpart).
constexpr uint32_t kQualitySpectral = 3;

struct V3 {
    float x, y, z;
};

__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 v3s(float s) { return V3{s, s, s}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator-(V3 a) { return V3{-a.x, -a.y, -a.z}; }
__device__ __forceinline__ V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ V3 fma3(V3 a, V3 b, V3 c) {
    return V3{fmaf(a.x, b.x, c.x), fmaf(a.y, b.y, c.y), fmaf(a.z, b.z, c.z)};
}
__device__ __forceinline__ float dot3(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ V3 cross3(V3 a, V3 b) {
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float length3(V3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ V3 normalize3(V3 a) {
    float l = length3(a);
    return V3{a.x / l, a.y / l, a.z / l};
}
// WGSL reflect(e1, e2) = e1 - 2 * dot(e2, e1) * e2
__device__ __forceinline__ V3 reflect3(V3 e, V3 n) {
    float d2 = 2.0f * dot3(n, e);
    return e - d2 * n;
}
// WGSL refract(e1, e2, eta)
__device__ __forceinline__ V3 refract3(V3 e, V3 n, float eta) {
    float d = dot3(n, e);
    float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return v3s(0.0f);
    return eta * e - (eta * d + sqrtf(k)) * n;
}
// bindings.wesl:171-175
__device__ __forceinline__ bool near_zero(float f) { return f < 0.0001f && f > -0.0001f; }
__device__ __forceinline__ bool v3_near_zero(V3 v) {
    return near_zero(v.x) && near_zero(v.y) && near_zero(v.z);
}
__device__ __forceinline__ float saturate(float x) { return fminf(fmaxf(x, 0.0f), 1.0f); }

__device__ __forceinline__ float f16_to_f32(uint32_t bits16) {
    return (float)__builtin_bit_cast(_Float16, (uint16_t)bits16);
}
__device__ __forceinline__ float bf16_to_f32(uint32_t bits16) {
    return __uint_as_float(bits16 << 16);
}

// Texel fetch with the default sampler (nearest, clamp-to-edge),
// textures.rs:196; RGBA8 UNORM -> c / 255.
__device__ __forceinline__ float4 unorm4(uint32_t t) {
    return make_float4((float)(t & 0xFFu) / 255.0f, (float)((t >> 8) & 0xFFu) / 255.0f,
                       (float)((t >> 16) & 0xFFu) / 255.0f, (float)(t >> 24) / 255.0f);
}
__device__ __forceinline__ uint32_t tex_texel(const TexDesc* __restrict__ tex,
                                              const uint32_t* __restrict__ texels, uint32_t ti,
                                              float u, float v) {
    const TexDesc d = tex[ti];
    float fx = floorf(u * (float)d.width);
    float fy = floorf(v * (float)d.height);
    fx = fminf(fmaxf(fx, 0.0f), (float)(d.width - 1));
    fy = fminf(fmaxf(fy, 0.0f), (float)(d.height - 1));
    uint32_t ix = (uint32_t)fx, iy = (uint32_t)fy;
    return texels[d.offset + iy * d.width + ix];
}
__device__ __forceinline__ float4 tex_sample(const SceneDev& s, uint32_t ti, float u, float v) {
    return unorm4(tex_texel(s.tex, s.texels, ti, u, v));
}

// Texture coordinates of a hit (general.wesl:1060-1061, 1118-1119):
// barycentrics (u, v) weight vertices 1 and 2, so
// tex = tex_pos_1 (1 - u - v) + tex_pos_2 u + tex_pos_3 v.
__device__ __forceinline__ void hit_tex_coords(const Material& m, float bu, float bv, float* tu,
                                               float* tv) {
    const float bz = (1.0f - bu) - bv;
    const float t1u = f16_to_f32(m.tex_pos_1 & 0xFFFFu), t1v = f16_to_f32(m.tex_pos_1 >> 16);
    const float t2u = f16_to_f32(m.tex_pos_2 & 0xFFFFu), t2v = f16_to_f32(m.tex_pos_2 >> 16);
    const float t3u = f16_to_f32(m.tex_pos_3 & 0xFFFFu), t3v = f16_to_f32(m.tex_pos_3 >> 16);
    *tu = fmaf(t1u, bz, fmaf(t2u, bu, t3u * bv));
    *tv = fmaf(t1v, bz, fmaf(t2v, bu, t3v * bv));
}
__device__ __forceinline__ float4 tex_load(const SceneDev& s, uint32_t ti, uint32_t x,
                                           uint32_t y) {
    const TexDesc d = s.tex[ti];
    x = min(x, d.width - 1);
    y = min(y, d.height - 1);
    return unorm4(s.texels[d.offset + y * d.width + x]);
}
// Cube-map lookup (general.wesl:1180-1182) with the standard major-axis face
// selection; faces +x, -x, +y, -y, +z, -z.
__device__ __forceinline__ V3 background(const SceneDev& s, V3 d) {
    float ax = fabsf(d.x), ay = fabsf(d.y), az = fabsf(d.z);
    int face;
    float sc, tc, ma;
    if (ax >= ay && ax >= az) {
        face = d.x >= 0.0f ? 0 : 1;
        ma = ax;
        sc = d.x >= 0.0f ? -d.z : d.z;
        tc = -d.y;
    } else if (ay >= az) {
        face = d.y >= 0.0f ? 2 : 3;
        ma = ay;
        sc = d.x;
        tc = d.y >= 0.0f ? d.z : -d.z;
    } else {
        face = d.z >= 0.0f ? 4 : 5;
        ma = az;
        sc = d.z >= 0.0f ? d.x : -d.x;
        tc = -d.y;
    }
    float u = 0.5f * (sc / ma + 1.0f);
    float v = 0.5f * (tc / ma + 1.0f);
    float4 c = tex_sample(s, s.bg_base + face, u, v);
    return v3(c.x, c.y, c.z);
}

// general.wesl:58-72 Intersection (the fields the path tracer reads).
struct Isect {
    V3 color;
    V3 emission;
    V3 normal;
    V3 tangent;
    float t;
    uint32_t ty;
    bool front_face;
    float rough_x, rough_y;      // roughness
    float rough_tx, rough_ty;    // roughness_transparent
    float ior_low, ior_high;
    uint32_t prim;
};

// general.wesl:1139-1156 correct_tangent
__device__ __forceinline__ V3 correct_tangent(V3 t) {
    float ax = fabsf(t.x), ay = fabsf(t.y), az = fabsf(t.z);
    bool x_largest = ax > ay && ax > az;
    bool y_largest = ay > ax && ay > az;
    bool reverse;
    if (x_largest)
        reverse = t.x < 0.0f;
    else if (y_largest)
        reverse = t.y < 0.0f;
    else
        reverse = t.z < 0.0f;
    return reverse ? -t : t;
}

// general.wesl:1080-1136: committed-hit record of a triangle hit.  Front
// face: the ray direction opposes the winding normal (v1-v0) x (v2-v0).
__device__ __forceinline__ Isect make_isect(const SceneDev& s, uint32_t leaf, float t, float bu,
                                            float bv, V3 dir) {
    Isect it;
    const float4* tp = reinterpret_cast<const float4*>(s.tris + leaf);
    const float4 a = tp[0], b = tp[1], c = tp[2];
    const uint32_t prim = __float_as_uint(a.w);
    it.prim = prim;
    it.t = t;
    V3 v0 = v3(a.x, a.y, a.z), v1 = v3(b.x, b.y, b.z), v2 = v3(c.x, c.y, c.z);
    it.front_face = dot3(dir, cross3(v1 - v0, v2 - v0)) < 0.0f;
    if (s.obj_tris) {
        // Instanced geometry: the reference fetches object-space vertices
        // (bindings.wesl:102-110) for the tangent and normal, and a
        // triangle's facing is decided in object space, where a mirroring
        // transform reverses the winding seen in world space.
        const float4* o = s.obj_tris + 3 * (uint64_t)prim;
        const float4 oa = o[0], ob = o[1], oc = o[2];
        v0 = v3(oa.x, oa.y, oa.z);
        v1 = v3(ob.x, ob.y, ob.z);
        v2 = v3(oc.x, oc.y, oc.z);
        if (oa.w != 0.0f) it.front_face = !it.front_face;
    }
    const Material m = s.materials[__float_as_uint(b.w)];  // material index (leaf record v1.w)
    const uint32_t idx_diffuse = m.tex_idx_diffuse_emission & 0xFFFFu;
    const uint32_t idx_emission = m.tex_idx_diffuse_emission >> 16;
    const uint32_t idx_attr = m.tex_idx_attributes_ty & 0xFFFFu;
    it.ty = m.tex_idx_attributes_ty >> 16;
    it.ior_low = f16_to_f32(m.refractive_index & 0xFFFFu);
    it.ior_high = f16_to_f32(m.refractive_index >> 16);
    V3 tangent = correct_tangent(normalize3(v0 - v1));
    V3 normal = normalize3(cross3(tangent, v0 - v2));
    it.normal = dot3(dir, normal) > 0.0f ? -normal : normal;
    it.tangent = tangent;
    float tu, tv;
    hit_tex_coords(m, bu, bv, &tu, &tv);
    const float4 dc = tex_sample(s, s.diffuse_base + idx_diffuse, tu, tv);
    const float4 rc = tex_load(s, s.recolour, m.tex_pos_recolour & 0xFFFFu,
                               m.tex_pos_recolour >> 16);
    it.color = v3(dc.x * rc.x, dc.y * rc.y, dc.z * rc.z);
    const float es_x = bf16_to_f32(m.emission_scale & 0xFFFFu);
    const float es_y = bf16_to_f32(m.emission_scale >> 16);
    V3 em = v3s(0.0f);
    if (idx_emission != kNoTexture) {
        const float4 ec = tex_sample(s, s.emission_base + idx_emission, tu, tv);
        em = v3(ec.x, ec.y, ec.z) * es_x;
    }
    it.emission = em + it.color * es_y;
    if (idx_attr != kNoTexture) {
        const float4 at = tex_sample(s, s.attr_base + idx_attr, tu, tv);
        it.rough_x = at.x;
        it.rough_y = at.y;
        it.rough_tx = at.z;
        it.rough_ty = at.w;
    } else {
        it.rough_x = it.rough_y = it.rough_tx = it.rough_ty = 0.0f;
    }
    return it;
}

// ---------------------------------------------------------------- sampling
// general.wesl:1212-1217 from_onb
__device__ __forceinline__ V3 from_onb(V3 src, V3 up, V3 tangent) {
    V3 t = cross3(up, tangent);
    return fma3(v3s(src.x), tangent, fma3(v3s(src.y), up, src.z * t));
}
// importance_sampling/shared.wesl:38-46 to_onb (equals -from_onb)
__device__ __forceinline__ V3 to_onb(V3 src, V3 up, V3 tangent) {
    V3 t = cross3(up, tangent);
    V3 inv_s = cross3(up, t);
    V3 inv_up = cross3(t, tangent);
    V3 inv_t = cross3(tangent, up);
    return fma3(v3s(src.x), inv_s, fma3(v3s(src.y), inv_up, src.z * inv_t));
}
// general.wesl:1219-1228 from_onb_no_tangent
__device__ __forceinline__ V3 from_onb_no_tangent(V3 src, V3 up) {
    V3 other = near_zero(up.x) ? v3(1.0f, 0.0f, 0.0f) : v3(0.0f, 1.0f, 0.0f);
    V3 tangent = normalize3(cross3(up, other));
    return from_onb(src, up, tangent);
}
// general.wesl:1230-1244, from the two uniforms r1 (azimuth) and r2
__device__ __forceinline__ V3 onb_cosine_weighted_hemisphere_u(float r1, float r2) {
    float phi = kTau * r1;
    float sr2 = sqrtf(r2);
    return v3(cosf(phi) * sr2, sqrtf(1.0f - r2), sinf(phi) * sr2);
}
__device__ __forceinline__ V3 onb_cosine_weighted_hemisphere(uint32_t seed) {
    return onb_cosine_weighted_hemisphere_u(rand_f32(seed), rand_f32(seed + 1u));
}
__device__ __forceinline__ V3 cosine_weighted_hemisphere(uint32_t seed, V3 normal, V3 tangent) {
    return from_onb(onb_cosine_weighted_hemisphere(seed), normal, tangent);
}
// general.wesl:1272-1281
__device__ __forceinline__ V3 sample_ggx_internal(float u1, float u2, V3 wi) {
    float phi = kTau * u1;
    float y = fmaf(1.0f - u2, 1.0f + wi.y, -wi.y);
    float sin_theta = sqrtf(fminf(fmaxf(fmaf(y, -y, 1.0f), 0.0f), 1.0f));
    return v3(sin_theta * cosf(phi), y, sin_theta * sinf(phi)) + wi;
}
// general.wesl:1246-1270, from the two uniforms u1 (azimuth) and u2
__device__ __forceinline__ V3 sample_vndf_ggx_u(float u1, float u2, V3 view, float rx, float ry) {
    V3 in_dir = v3(view.x * rx, view.y, view.z * ry);
    if (v3_near_zero(in_dir)) in_dir = v3(0.0f, -1.0f, 0.0f);
    V3 wm_std = sample_ggx_internal(u1, u2, normalize3(in_dir));
    V3 out = v3(wm_std.x * rx, wm_std.y, wm_std.z * ry);
    if (v3_near_zero(out)) out = v3(0.0f, 1.0f, 0.0f);
    return normalize3(out);
}
__device__ __forceinline__ V3 sample_vndf_ggx(uint32_t seed, V3 view, float rx, float ry) {
    return sample_vndf_ggx_u(rand_f32(seed), rand_f32(seed + 1u), view, rx, ry);
}
// general.wesl:1206-1210 Schlick
__device__ __forceinline__ float reflectance(float cosine, float ref_idx) {
    float r0 = (1.0f - ref_idx) / (1.0f + ref_idx);
    float r0s = r0 * r0;
    return fmaf(1.0f - r0s, powf(1.0f - cosine, 5.0f), r0s);
}
// importance_sampling/shared.wesl:215-231
__device__ __forceinline__ V3 oren_nayar_weight(V3 incoming, V3 reflected, float roughness,
                                                V3 color) {
    float rr = roughness * kHalfPi;
    float rr2 = rr * rr;
    V3 a = fma3(0.17f * color, v3s(rr2 / (rr2 + 0.13f)),
                v3s(1.0f - 0.5f * (rr2 / (rr2 + 0.33f))));
    float b = 0.45f * (rr2 / (rr2 + 0.09f));
    float s = fmaf(-reflected.y, incoming.y, dot3(reflected, incoming));
    float t = 1.0f;
    if (s > 0.0f) {
        t = fmaxf(reflected.y, incoming.y);
        if (near_zero(t)) t = 0.001f;
    }
    return fma3(v3s(b), v3s(s / t), a);
}

// --------------------------------------------------------- quality hooks
// low.wesl / medium.wesl / high.wesl get_normal, from the uniforms u1, u2
// (the reference draws them at seed and seed + 1)
template <uint32_t Q>
__device__ __forceinline__ V3 get_normal_u(V3 ray_dir, V3 normal, V3 tangent, float u1, float u2,
                                           float rx, float ry) {
    if constexpr (Q == kQualityLow) {
        return normal;
    } else {
        V3 ray_onb = to_onb(ray_dir, normal, tangent);
        V3 n_onb = sample_vndf_ggx_u(u1, u2, ray_onb, rx, ry);
        return from_onb(n_onb, normal, tangent);
    }
}
template <uint32_t Q>
__device__ __forceinline__ V3 get_normal(V3 ray_dir, V3 normal, V3 tangent, uint32_t seed,
                                         float rx, float ry) {
    if constexpr (Q == kQualityLow) {
        return normal;
    } else {
        return get_normal_u<Q>(ray_dir, normal, tangent, rand_f32(seed), rand_f32(seed + 1u), rx, ry);
    }
}
// make_diffuse (note: the world-space new direction is passed as the
// ONB-local reflected vector, medium.wesl:10-13)
template <uint32_t Q>
__device__ __forceinline__ V3 make_diffuse(V3 hit_color, V3 ray_dir, V3 normal, V3 tangent,
                                           float roughness, V3 new_dir) {
    if constexpr (Q == kQualityLow) {
        return v3s(1.0f);
    } else {
        V3 ray_onb = to_onb(ray_dir, normal, tangent);
        return oren_nayar_weight(ray_onb, new_dir, roughness, hit_color);
    }
}
template <uint32_t Q>
__device__ __forceinline__ float conv_rough(float r) {
    return Q == kQualityLow ? 0.0f : r;
}
// high.wesl:79-90
__device__ __forceinline__ float cie_g(float x, float u, float t1, float t2) {
    float tau = x < u ? t1 : t2;
    float d = x - u;
    return expf(-((d * d) * (tau * tau)) * 0.5f);
}
// high.wesl:38-60: saturated sRGB of the multi-lobe CIE fit at lambda (nm).
__device__ __forceinline__ V3 cie_rgb(float lambda) {
    float x = fmaf(1.056f, cie_g(lambda, 599.8f, 0.0264f, 0.0323f),
                   fmaf(0.362f, cie_g(lambda, 442.0f, 0.0624f, 0.0374f),
                        -(0.065f * cie_g(lambda, 501.1f, 0.049f, 0.0382f))));
    float y = fmaf(0.821f, cie_g(lambda, 568.8f, 0.0213f, 0.0248f),
                   0.286f * cie_g(lambda, 530.9f, 0.0613f, 0.0322f));
    float z = fmaf(1.217f, cie_g(lambda, 437.0f, 0.0845f, 0.0278f),
                   0.681f * cie_g(lambda, 459.0f, 0.0385f, 0.0725f));
    float r = (3.2406f * x + -1.5372f * y) + -0.4986f * z;
    float g = (-0.9689f * x + 1.8758f * y) + 0.0415f * z;
    float b = (0.0557f * x + -0.204f * y) + 1.057f * z;
    return v3(saturate(r), saturate(g), saturate(b));
}

// Spectral quality: four wavelengths per path, lambda_j = 380 + 400
// fract(u + j/4) nm with u the hero draw (hero-wavelength sampling, Wilkie et
// al. 2014).  The weights are normalised so that their mean over 380-780 nm
// is 1 per channel (1 / mean of cie_rgb: 3.7238771, 4.2015504, 5.2524004, by
// numerical integration; the reference's High constants are the same
// normalisation over 340-760 nm).  Until the path meets a dispersive
// interface (IOR range low != high) the colour weight is the mean of the four
// wavelengths'; there only the hero can follow its refraction, so the path
// continues with the hero alone (weight = 4 x 1/4 of it) and its IOR from the
// Cauchy law n = A + B / lambda^2 (lambda in um) through n(760 nm) = low and
// n(340 nm) = high, the reference's IOR-range endpoints.  *picked >= 2 marks
// a collapsed path.
constexpr float kSpecK0 = 3.7238771f, kSpecK1 = 4.2015504f, kSpecK2 = 5.2524004f;
__device__ __forceinline__ V3 spec_weight(float lambda) {
    const V3 c = cie_rgb(lambda);
    return v3(c.x * kSpecK0, c.y * kSpecK1, c.z * kSpecK2);
}
__device__ __forceinline__ V3 spec_mean_weight(float u) {
    V3 acc = v3s(0.0f);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        float v = u + 0.25f * (float)j;
        v = v - floorf(v);
        acc = acc + spec_weight(fmaf(400.0f, v, 380.0f));
    }
    return acc * v3s(0.25f);
}
__device__ __forceinline__ float cauchy_ior(float low, float high, float lambda) {
    const float inv_hi = 1.0f / (0.34f * 0.34f), inv_lo = 1.0f / (0.76f * 0.76f);
    const float B = (high - low) / (inv_hi - inv_lo);
    const float A = low - B * inv_lo;
    const float um = lambda * 0.001f;
    return A + B / (um * um);
}

// pick_ref_idx: *picked < 0 means not yet picked (the index is in [0, 1)).
template <uint32_t Q>
__device__ __forceinline__ float pick_ref_idx(float* picked, float low, float high,
                                              uint32_t seed, V3* colour) {
    *colour = v3s(1.0f);
    if constexpr (Q == kQualityLow) {
        return low;
    } else {
        bool fresh = false;
        if (*picked < 0.0f) {
            *picked = rand_f32(seed);
            fresh = true;
        }
        if constexpr (Q == kQualitySpectral) {
            // u on the 2^-22 grid, so that the collapsed encoding u + 2 is exact
            if (fresh) *picked = (*picked + 2.0f) - 2.0f;
            const bool collapsed = *picked >= 2.0f;
            const float u = collapsed ? *picked - 2.0f : *picked;
            const float hero = fmaf(400.0f, u - floorf(u), 380.0f);
            if (low == high) {
                if (fresh) *colour = spec_mean_weight(u);
                return low;
            }
            if (!collapsed) {
                const V3 h = spec_weight(hero);
                if (fresh) {
                    *colour = h;
                } else {
                    const V3 m = spec_mean_weight(u);
                    *colour = v3(m.x > 0.0f ? h.x / m.x : 0.0f, m.y > 0.0f ? h.y / m.y : 0.0f,
                                 m.z > 0.0f ? h.z / m.z : 0.0f);
                }
                *picked = u + 2.0f;
            }
            return cauchy_ior(low, high, hero);
        } else {
            float idx = *picked;
            if constexpr (Q == kQualityHigh) {
                if (fresh) {
                    const V3 c = cie_rgb(760.0f * (1.0f - idx) + 340.0f * idx);
                    *colour = v3(c.x * 3.910f, c.y * 4.411f, c.z * 5.511f);
                }
            }
            return low * (1.0f - idx) + high * idx;
        }
    }
}

// Uniforms that stand in for the reference's draws in one scatter (a
// stratified sampler's point): u0, u1 the direction pair (the microfacet
// normal's, drawn at seed and seed + 1, and the cosine lobe's, drawn later;
// a scatter uses one of the two for its direction), u2 the Fresnel choice.
struct ScatterUniforms {
    float u0, u1, u2;
};

// general.wesl:714-901 with every guiding candidate invalid ("plain"): the
// Fresnel choice, the refractive-index pick, and diffuse / metal / dielectric
// scattering.  Advances `seed` exactly as the reference does, multiplies
// `color` by the sampled weight and returns the continuation direction.  The
// caller accumulates the vertex emission first (general.wesl:714).  With
// kGiven the direction pair and the Fresnel choice come from `su` instead of
// the seed's draws (the seed still advances as in the reference).
template <uint32_t Q, bool kGiven = false>
__device__ __forceinline__ V3 surface_scatter(const Isect& it, V3 dir, uint32_t& seed, V3& color,
                                              float& picked, ScatterUniforms su = {}) {
    const float g1 = kGiven ? su.u0 : rand_f32(seed);
    const float g2 = kGiven ? su.u1 : rand_f32(seed + 1u);
    const V3 assign_normal = get_normal_u<Q>(dir, it.normal, it.tangent, g1, g2, it.rough_x, it.rough_y);
    const V3 transparent_normal =
        get_normal_u<Q>(dir, it.normal, it.tangent, g1, g2, it.rough_tx, it.rough_ty);
    V3 normal = transparent_normal;
    bool should_reflect = false;
    seed += 1u;
    V3 ccol;
    float ri = pick_ref_idx<Q>(&picked, it.ior_low, it.ior_high, seed, &ccol);
    seed += 1u;
    if (!(!it.front_face && it.ty == kMatTransparent)) ri = 1.0f / ri;
    color = color * ccol;
    if (ri != 1.0f) {
        const float cos_theta = fminf(dot3(-dir, normal), 1.0f);
        const float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
        const bool cannot_refract = ri * sin_theta > 1.0f;
        const float pdf = reflectance(cos_theta, ri);
        const float rnd = kGiven ? su.u2 : rand_f32(seed);
        if (cannot_refract || pdf > rnd) should_reflect = true;
        seed += 1u;
    }
    if (!should_reflect || it.ty == kMatTransparent) color = color * it.color;
    if (!should_reflect) {
        normal = assign_normal;
        seed += 1u;
    }
    const bool reflecting = should_reflect || it.ty == kMatMetallic;
    V3 ndir = dir;
    if (!reflecting) {
        if (it.ty == kMatDiffuse) {
            const V3 dir_cos =
                kGiven ? from_onb(onb_cosine_weighted_hemisphere_u(su.u0, su.u1), it.normal, it.tangent)
                       : cosine_weighted_hemisphere(seed, it.normal, it.tangent);
            const float cos_pdf = fmaxf(dot3(dir_cos, it.normal) / kPi, 0.0f);
            seed += 3u;  // theta, xi and the lobe choice
            const float avg_pdf = 1.0f * cos_pdf;
            const V3 diffuse = make_diffuse<Q>(it.color, dir, it.normal, it.tangent, it.rough_x, dir_cos);
            const float cached_brdf = fmaxf(dot3(dir_cos, it.normal) / kPi, 0.0f);
            const float nf = fabsf(cached_brdf) / fmaxf(avg_pdf, 0.001f);
            color = (color * nf) * diffuse;
            ndir = dir_cos;
        } else if (it.ty == kMatTransparent) {
            ndir = refract3(dir, normal, ri);
        }
    } else {
        ndir = reflect3(dir, normal);
        if (dot3(ndir, it.normal) < 0.0f) ndir = reflect3(ndir, it.normal);
    }
    return ndir;
}

}  // namespace lmb
