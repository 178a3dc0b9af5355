// oracle_pt.cpp - CPU ORACLE (test infrastructure only) for the path tracer.
//
// A per-pixel restatement of the reference megakernel
//   rt_main    path_tracing/general.wesl:506-610 (seeds, sample loop, outputs)
//   rt_sample  general.wesl:643-975 (bounce loop, emission, Fresnel, scatter)
//   ray_hit    general.wesl:1080-1136 (hit record: tangent, normal, textures)
//   helpers    general.wesl:1139-1281, importance_sampling/shared.wesl:38-46,
//              215-231, quality modules low.wesl / medium.wesl / high.wesl
// in "plain" mode: Markov-chain guiding off, i.e. every guiding candidate
// invalid, so the diffuse lobe is the cosine lobe with probability 1
// (general.wesl:812-815) and the guiding random draws are left out.
// Optional Russian roulette from bounce `rr_start` (not in the reference).
//
// Closest hits come from the oracle BVH in oracle_geom.cpp, which returns
// exactly the brute-force watertight result.
//
// Pinned by (tests/test_oracle_pt.py): closed-box furnace identities
// (diffuse / metal / Low quality: radiance E * sum_{k<D} rho^k, exact per
// path), emission-only and background-only images, Russian-roulette
// unbiasedness, the Oren-Nayar weight at roughness 0, VNDF / cosine sample
// distributions (chi-square), the Schlick closed form, and the CIE colour
// weights of the High quality against a float64 evaluation.
#include <omp.h>
#include <stdint.h>

#include <cmath>
#include <cstring>
#include <vector>

extern "C" {
uint32_t orc_rand_u32(uint32_t s);
float orc_rand_f32(uint32_t s);
void orc_create_ray(const float* cam, uint32_t px, uint32_t py, uint32_t width, uint32_t height,
                    float* o, float* d);
void* orc_bvh_create(const float* verts, uint64_t nverts, const uint32_t* idx, uint64_t ntri);
void orc_bvh_destroy(void* h);
struct OrcHit {
    uint32_t prim;
    float t, u, v;
};
void orc_bvh_intersect1(void* hb, const float* ray, OrcHit* hit);
typedef bool (*OrcAccept)(const void* ctx, uint32_t prim, float u, float v);
void orc_bvh_intersect1_filtered(void* hb, const float* ray, OrcHit* hit, OrcAccept accept,
                                 const void* ctx);
}

namespace {

const float PI = 3.14159265358979323846f;
const float HALF_PI = PI / 2.0f;
const float TAU = 2.0f * PI;

struct F3 {
    float x, y, z;
};
inline F3 mk(float x, float y, float z) { return F3{x, y, z}; }
inline F3 splat(float s) { return F3{s, s, s}; }
inline F3 add(F3 a, F3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
inline F3 sub(F3 a, F3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
inline F3 neg(F3 a) { return mk(-a.x, -a.y, -a.z); }
inline F3 mul(F3 a, F3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
inline F3 scale(F3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
inline F3 divs(F3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
inline F3 fma3(F3 a, F3 b, F3 c) {
    return mk(std::fma(a.x, b.x, c.x), std::fma(a.y, b.y, c.y), std::fma(a.z, b.z, c.z));
}
inline float dot(F3 a, F3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
inline F3 cross(F3 a, F3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
inline float length(F3 a) { return std::sqrt(dot(a, a)); }
inline F3 normalize(F3 a) {
    float l = length(a);
    return mk(a.x / l, a.y / l, a.z / l);
}
inline F3 reflect(F3 e1, F3 e2) { return sub(e1, scale(e2, 2.0f * dot(e2, e1))); }
inline F3 refract(F3 e1, F3 e2, float eta) {
    float d = dot(e2, e1);
    float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return splat(0.0f);
    return sub(scale(e1, eta), scale(e2, eta * d + std::sqrt(k)));
}
inline bool near_zero(float f) { return f < 0.0001f && f > -0.0001f; }
inline bool vec3_near_zero(F3 v) { return near_zero(v.x) && near_zero(v.y) && near_zero(v.z); }
inline float saturate(float x) { return std::fmin(std::fmax(x, 0.0f), 1.0f); }

float half_to_float(uint32_t h) {
    uint32_t s = (h >> 15) & 1u, e = (h >> 10) & 0x1Fu, m = h & 0x3FFu;
    float v;
    if (e == 0)
        v = std::ldexp((float)m, -24);
    else if (e == 31)
        v = m ? NAN : INFINITY;
    else
        v = std::ldexp((float)(m | 0x400u), (int)e - 25);
    return s ? -v : v;
}
float bfloat_to_float(uint32_t b) {
    uint32_t x = b << 16;
    float f;
    std::memcpy(&f, &x, 4);
    return f;
}

struct Scene {
    void* bvh;
    std::vector<float> verts;
    std::vector<uint32_t> idx;
    uint64_t ntri;
    std::vector<uint32_t> matidx;
    std::vector<uint32_t> mats;  // 8 words per material
    std::vector<uint32_t> texels;
    std::vector<uint32_t> descs;  // offset, width, height per texture
    uint32_t diffuse_base, emission_base, attr_base, recolour, bg_base;
    bool alpha_test = false;  // non-opaque geometry: any-hit alpha test
    // Instanced geometry: per primitive the object-space triangle (9 floats)
    // and whether its instance transform mirrors (determinant < 0).  The
    // reference's vertex fetch returns object-space positions
    // (bindings.wesl:102-110), so the tangent and normal come from these,
    // and the facing of a triangle is decided in object space.
    std::vector<float> obj;
    std::vector<uint8_t> mirrored;
};

struct Tex4 {
    float r, g, b, a;
};

Tex4 texel_value(const Scene& s, uint32_t ti, uint32_t x, uint32_t y) {
    uint32_t off = s.descs[3 * ti], w = s.descs[3 * ti + 1], h = s.descs[3 * ti + 2];
    if (x >= w) x = w - 1;
    if (y >= h) y = h - 1;
    uint32_t t = s.texels[off + y * w + x];
    return Tex4{(float)(t & 255u) / 255.0f, (float)((t >> 8) & 255u) / 255.0f,
                (float)((t >> 16) & 255u) / 255.0f, (float)(t >> 24) / 255.0f};
}

// Nearest filtering, clamp to edge (the default wgpu sampler).
Tex4 sample_tex(const Scene& s, uint32_t ti, float u, float v) {
    uint32_t w = s.descs[3 * ti + 1], h = s.descs[3 * ti + 2];
    float fx = std::floor(u * (float)w), fy = std::floor(v * (float)h);
    if (!(fx >= 0.0f)) fx = 0.0f;
    if (!(fy >= 0.0f)) fy = 0.0f;
    if (fx > (float)(w - 1)) fx = (float)(w - 1);
    if (fy > (float)(h - 1)) fy = (float)(h - 1);
    return texel_value(s, ti, (uint32_t)fx, (uint32_t)fy);
}

// Cube map: major axis picks the face (+x,-x,+y,-y,+z,-z), then (s, t) as in
// the Vulkan cube-map face table.
F3 background(const Scene& s, F3 d) {
    float ax = std::fabs(d.x), ay = std::fabs(d.y), az = std::fabs(d.z);
    uint32_t face;
    float sc, tc, ma;
    if (ax >= ay && ax >= az) {
        ma = ax;
        if (d.x >= 0.0f) {
            face = 0;
            sc = -d.z;
        } else {
            face = 1;
            sc = d.z;
        }
        tc = -d.y;
    } else if (ay >= az) {
        ma = ay;
        sc = d.x;
        if (d.y >= 0.0f) {
            face = 2;
            tc = d.z;
        } else {
            face = 3;
            tc = -d.z;
        }
    } else {
        ma = az;
        if (d.z >= 0.0f) {
            face = 4;
            sc = d.x;
        } else {
            face = 5;
            sc = -d.x;
        }
        tc = -d.y;
    }
    Tex4 c = sample_tex(s, s.bg_base + face, 0.5f * (sc / ma + 1.0f), 0.5f * (tc / ma + 1.0f));
    return mk(c.r, c.g, c.b);
}

struct Hit {
    bool hit;
    float t;
    F3 color, emission, normal, tangent;
    uint32_t ty;
    bool front_face;
    float rx, ry, rtx, rty, n_lo, n_hi;
};

F3 correct_tangent(F3 dir) {
    F3 a = mk(std::fabs(dir.x), std::fabs(dir.y), std::fabs(dir.z));
    bool x_largest = a.x > a.y && a.x > a.z;
    bool y_largest = a.y > a.x && a.y > a.z;
    bool rev;
    if (x_largest)
        rev = dir.x < 0.0f;
    else if (y_largest)
        rev = dir.y < 0.0f;
    else
        rev = dir.z < 0.0f;
    return rev ? neg(dir) : dir;
}

// Texture coordinates of a hit: tex_pos_1 (1 - u - v) + tex_pos_2 u +
// tex_pos_3 v (general.wesl:1060-1061, 1118-1119).
void hit_tex_coords(const uint32_t* m, float bx, float by, float* tu, float* tv) {
    float bz = (1.0f - bx) - by;
    *tu = std::fma(half_to_float(m[0] & 0xFFFFu), bz,
                   std::fma(half_to_float(m[1] & 0xFFFFu), bx, half_to_float(m[2] & 0xFFFFu) * by));
    *tv = std::fma(half_to_float(m[0] >> 16), bz,
                   std::fma(half_to_float(m[1] >> 16), bx, half_to_float(m[2] >> 16) * by));
}

// Any-hit rule of non-opaque geometry (general.wesl:1057-1066): the diffuse
// texel's alpha is not near zero, or (path-tracing shaders) the material is
// TRANSPARENT; the debugging shaders test the alpha alone.
struct AcceptCtx {
    const Scene* s;
    bool transparent_passes;
};
bool alpha_accept(const void* c, uint32_t prim, float u, float v) {
    const AcceptCtx& a = *(const AcceptCtx*)c;
    const uint32_t* m = &a.s->mats[8 * a.s->matidx[prim]];
    if (a.transparent_passes && (m[5] >> 16) == 2u) return true;
    float tu, tv;
    hit_tex_coords(m, u, v, &tu, &tv);
    float alpha = sample_tex(*a.s, a.s->diffuse_base + (m[4] & 0xFFFFu), tu, tv).a;
    return !(alpha < 0.0001f && alpha > -0.0001f);
}

Hit ray_hit(const Scene& s, F3 org, F3 dir, float tmin, float tmax,
            bool transparent_passes = true) {
    float ray[8] = {org.x, org.y, org.z, tmin, dir.x, dir.y, dir.z, tmax};
    OrcHit oh;
    if (s.alpha_test) {
        AcceptCtx ctx{&s, transparent_passes};
        orc_bvh_intersect1_filtered(s.bvh, ray, &oh, alpha_accept, &ctx);
    } else {
        orc_bvh_intersect1(s.bvh, ray, &oh);
    }
    Hit h;
    std::memset(&h, 0, sizeof h);
    if (oh.prim == 0xFFFFFFFFu) return h;
    h.hit = true;
    h.t = oh.t;
    uint64_t p = oh.prim;
    uint64_t i0 = s.idx.empty() ? 3 * p : s.idx[3 * p], i1 = s.idx.empty() ? 3 * p + 1 : s.idx[3 * p + 1],
             i2 = s.idx.empty() ? 3 * p + 2 : s.idx[3 * p + 2];
    F3 v0 = mk(s.verts[4 * i0], s.verts[4 * i0 + 1], s.verts[4 * i0 + 2]);
    F3 v1 = mk(s.verts[4 * i1], s.verts[4 * i1 + 1], s.verts[4 * i1 + 2]);
    F3 v2 = mk(s.verts[4 * i2], s.verts[4 * i2 + 1], s.verts[4 * i2 + 2]);
    // front face: counter-clockwise seen from the ray origin
    h.front_face = dot(dir, cross(sub(v1, v0), sub(v2, v0))) < 0.0f;
    if (!s.obj.empty()) {
        const float* o = &s.obj[9 * p];
        v0 = mk(o[0], o[1], o[2]);
        v1 = mk(o[3], o[4], o[5]);
        v2 = mk(o[6], o[7], o[8]);
        if (s.mirrored[p]) h.front_face = !h.front_face;
    }
    const uint32_t* m = &s.mats[8 * s.matidx[p]];
    uint32_t idx_diffuse = m[4] & 0xFFFFu, idx_emission = m[4] >> 16;
    uint32_t idx_attr = m[5] & 0xFFFFu;
    h.ty = m[5] >> 16;
    h.n_lo = half_to_float(m[7] & 0xFFFFu);
    h.n_hi = half_to_float(m[7] >> 16);
    F3 tangent = correct_tangent(normalize(sub(v0, v1)));
    F3 normal = normalize(cross(tangent, sub(v0, v2)));
    h.normal = dot(dir, normal) > 0.0f ? neg(normal) : normal;
    h.tangent = tangent;
    float tu, tv;
    hit_tex_coords(m, oh.u, oh.v, &tu, &tv);
    Tex4 dt = sample_tex(s, s.diffuse_base + idx_diffuse, tu, tv);
    Tex4 rc = texel_value(s, s.recolour, m[3] & 0xFFFFu, m[3] >> 16);
    h.color = mk(dt.r * rc.r, dt.g * rc.g, dt.b * rc.b);
    float es0 = bfloat_to_float(m[6] & 0xFFFFu), es1 = bfloat_to_float(m[6] >> 16);
    F3 em = splat(0.0f);
    if (idx_emission != 0xFFFFu) {
        Tex4 et = sample_tex(s, s.emission_base + idx_emission, tu, tv);
        em = scale(mk(et.r, et.g, et.b), es0);
    }
    h.emission = add(em, scale(h.color, es1));
    if (idx_attr != 0xFFFFu) {
        Tex4 at = sample_tex(s, s.attr_base + idx_attr, tu, tv);
        h.rx = at.r;
        h.ry = at.g;
        h.rtx = at.b;
        h.rty = at.a;
    }
    return h;
}

F3 from_onb(F3 src, F3 up, F3 tangent) {
    F3 t = cross(up, tangent);
    return fma3(splat(src.x), tangent, fma3(splat(src.y), up, scale(t, src.z)));
}
F3 to_onb(F3 src, F3 up, F3 tangent) {
    F3 t = cross(up, tangent);
    F3 inv_s = cross(up, t), inv_up = cross(t, tangent), inv_t = cross(tangent, up);
    return fma3(splat(src.x), inv_s, fma3(splat(src.y), inv_up, scale(inv_t, src.z)));
}
F3 cosine_weighted_hemisphere(uint32_t seed, F3 normal, F3 tangent) {
    float r1 = orc_rand_f32(seed), r2 = orc_rand_f32(seed + 1u);
    float phi = TAU * r1;
    F3 hemi = mk(std::cos(phi) * std::sqrt(r2), std::sqrt(1.0f - r2), std::sin(phi) * std::sqrt(r2));
    return from_onb(hemi, normal, tangent);
}
F3 sample_vndf_ggx(uint32_t seed, F3 view, float rx, float ry) {
    float u1 = orc_rand_f32(seed), u2 = orc_rand_f32(seed + 1u);
    F3 in_dir = mk(view.x * rx, view.y, view.z * ry);
    if (vec3_near_zero(in_dir)) in_dir = mk(0.0f, -1.0f, 0.0f);
    F3 wi = normalize(in_dir);
    float phi = TAU * u1;
    float y = std::fma(1.0f - u2, 1.0f + wi.y, -wi.y);
    float sin_theta = std::sqrt(std::fmin(std::fmax(std::fma(y, -y, 1.0f), 0.0f), 1.0f));
    F3 wm = add(mk(sin_theta * std::cos(phi), y, sin_theta * std::sin(phi)), wi);
    F3 out = mk(wm.x * rx, wm.y, wm.z * ry);
    if (vec3_near_zero(out)) out = mk(0.0f, 1.0f, 0.0f);
    return normalize(out);
}
float reflectance(float cosine, float ref_idx) {
    float r0 = (1.0f - ref_idx) / (1.0f + ref_idx);
    float r0s = r0 * r0;
    return std::fma(1.0f - r0s, std::pow(1.0f - cosine, 5.0f), r0s);
}
F3 oren_nayar(F3 incoming, F3 reflected, float roughness, F3 color) {
    float rr = roughness * HALF_PI;
    float rr2 = rr * rr;
    F3 a = fma3(scale(color, 0.17f), splat(rr2 / (rr2 + 0.13f)),
                splat(1.0f - 0.5f * (rr2 / (rr2 + 0.33f))));
    float b = 0.45f * (rr2 / (rr2 + 0.09f));
    float s = std::fma(-reflected.y, incoming.y, dot(reflected, incoming));
    float t = 1.0f;
    if (s > 0.0f) {
        t = std::fmax(reflected.y, incoming.y);
        if (near_zero(t)) t = 0.001f;
    }
    return fma3(splat(b), splat(s / t), a);
}

F3 get_normal(uint32_t q, F3 ray_dir, F3 normal, F3 tangent, uint32_t seed, float rx, float ry) {
    if (q == 0) return normal;
    F3 n_onb = sample_vndf_ggx(seed, to_onb(ray_dir, normal, tangent), rx, ry);
    return from_onb(n_onb, normal, tangent);
}
F3 make_diffuse(uint32_t q, F3 hit_color, F3 ray_dir, F3 normal, F3 tangent, float rough,
                F3 new_dir) {
    if (q == 0) return splat(1.0f);
    return oren_nayar(to_onb(ray_dir, normal, tangent), new_dir, rough, hit_color);
}
float cie_g(float x, float u, float t1, float t2) {
    float tau = x < u ? t1 : t2;
    float d = x - u;
    return std::exp(-((d * d) * (tau * tau)) * 0.5f);
}
// high.wesl:38-60: saturated sRGB of the multi-lobe CIE fit at lambda (nm)
F3 cie_rgb(float lambda) {
    float x = std::fma(1.056f, cie_g(lambda, 599.8f, 0.0264f, 0.0323f),
                       std::fma(0.362f, cie_g(lambda, 442.0f, 0.0624f, 0.0374f),
                                -(0.065f * cie_g(lambda, 501.1f, 0.049f, 0.0382f))));
    float y = std::fma(0.821f, cie_g(lambda, 568.8f, 0.0213f, 0.0248f),
                       0.286f * cie_g(lambda, 530.9f, 0.0613f, 0.0322f));
    float z = std::fma(1.217f, cie_g(lambda, 437.0f, 0.0845f, 0.0278f),
                       0.681f * cie_g(lambda, 459.0f, 0.0385f, 0.0725f));
    float r = (3.2406f * x + -1.5372f * y) + -0.4986f * z;
    float g = (-0.9689f * x + 1.8758f * y) + 0.0415f * z;
    float b = (0.0557f * x + -0.204f * y) + 1.057f * z;
    return mk(saturate(r), saturate(g), saturate(b));
}

// Spectral quality (no reference counterpart): four wavelengths per path,
// lambda_j = 380 + 400 fract(u + j/4) nm, u the hero draw.  The colour
// weight is the mean of the four normalised CIE weights until the path meets
// a dispersive interface (IOR range lo != hi); there the hero wavelength
// alone continues (the others cannot follow its refraction) and the weight
// becomes the hero's.  IOR: Cauchy n = A + B / lambda^2 (lambda in um)
// through n(760 nm) = lo and n(340 nm) = hi.  picked >= 2 marks "collapsed".
constexpr float kSpecK0 = 3.7238771f, kSpecK1 = 4.2015504f, kSpecK2 = 5.2524004f;
F3 spec_weight(float lambda) {
    F3 c = cie_rgb(lambda);
    return mk(c.x * kSpecK0, c.y * kSpecK1, c.z * kSpecK2);
}
F3 spec_mean_weight(float u) {
    F3 acc = splat(0.0f);
    for (int j = 0; j < 4; ++j) {
        float v = u + 0.25f * (float)j;
        v = v - std::floor(v);
        acc = add(acc, spec_weight(std::fma(400.0f, v, 380.0f)));
    }
    return scale(acc, 0.25f);
}
float cauchy_ior(float lo, float hi, float lambda) {
    const float inv_hi = 1.0f / (0.34f * 0.34f), inv_lo = 1.0f / (0.76f * 0.76f);
    const float B = (hi - lo) / (inv_hi - inv_lo);
    const float A = lo - B * inv_lo;
    const float um = lambda * 0.001f;
    return A + B / (um * um);
}

// returns IOR; *colour = wavelength weight; *picked < 0 = not picked yet
float pick_ref_idx(uint32_t q, float* picked, float lo, float hi, uint32_t seed, F3* colour) {
    *colour = splat(1.0f);
    if (q == 0) return lo;
    bool fresh = *picked < 0.0f;
    if (fresh) *picked = orc_rand_f32(seed);
    if (q == 3) {
        // u on the 2^-22 grid, so that the collapsed encoding u + 2 is exact
        if (fresh) *picked = (*picked + 2.0f) - 2.0f;
        const bool collapsed = *picked >= 2.0f;
        const float u = collapsed ? *picked - 2.0f : *picked;
        const float hero = std::fma(400.0f, u - std::floor(u), 380.0f);
        if (lo == hi) {
            if (fresh) *colour = spec_mean_weight(u);
            return lo;
        }
        if (!collapsed) {
            const F3 h = spec_weight(hero);
            if (fresh) {
                *colour = h;
            } else {
                const F3 m = spec_mean_weight(u);
                *colour = mk(m.x > 0.0f ? h.x / m.x : 0.0f, m.y > 0.0f ? h.y / m.y : 0.0f,
                             m.z > 0.0f ? h.z / m.z : 0.0f);
            }
            *picked = u + 2.0f;
        }
        return cauchy_ior(lo, hi, hero);
    }
    float u = *picked;
    if (q == 2 && fresh) {
        float lambda = 760.0f * (1.0f - u) + 340.0f * u;
        F3 c = cie_rgb(lambda);
        *colour = mk(c.x * 3.910f, c.y * 4.411f, c.z * 5.511f);
    }
    return lo * (1.0f - u) + hi * u;
}

struct PlainOpts {
    uint32_t quality, bounces, rr_start;
    float tmin, tmax;
};

struct SampleOut {
    F3 color, normal, albedo;
    uint32_t vertices;
};

SampleOut rt_sample_plain(const Scene& s, const PlainOpts& o, F3 org, F3 dir, uint32_t own_seed) {
    SampleOut out;
    out.color = splat(0.0f);
    out.normal = splat(0.0f);
    out.albedo = splat(0.0f);
    out.vertices = 0;
    uint32_t self_seed = own_seed;
    F3 color = splat(1.0f);
    float picked = -1.0f;
    for (uint32_t i = 0; i < o.bounces; ++i) {
        out.vertices++;
        Hit it = ray_hit(s, org, dir, o.tmin, o.tmax);
        if (!it.hit) {
            out.color = fma3(color, background(s, dir), out.color);
            break;
        }
        F3 new_pos = fma3(dir, splat(it.t), org);
        org = new_pos;
        if (i == 0) {
            out.normal = it.normal;
            out.albedo = it.color;
        }
        out.color = fma3(color, it.emission, out.color);
        F3 assign_normal = get_normal(o.quality, dir, it.normal, it.tangent, self_seed, it.rx, it.ry);
        F3 transparent_normal =
            get_normal(o.quality, dir, it.normal, it.tangent, self_seed, it.rtx, it.rty);
        F3 normal = transparent_normal;
        bool should_reflect = false;
        self_seed += 1u;
        F3 ccol;
        float ri = pick_ref_idx(o.quality, &picked, it.n_lo, it.n_hi, self_seed, &ccol);
        self_seed += 1u;
        if (!(!it.front_face && it.ty == 2u)) ri = 1.0f / ri;
        color = mul(color, ccol);
        if (ri != 1.0f) {
            float cos_theta = std::fmin(dot(neg(dir), normal), 1.0f);
            float sin_theta = std::sqrt(1.0f - cos_theta * cos_theta);
            bool cannot = ri * sin_theta > 1.0f;
            float pdf = reflectance(cos_theta, ri);
            if (cannot || pdf > orc_rand_f32(self_seed)) should_reflect = true;
            self_seed += 1u;
        }
        if (!should_reflect || it.ty == 2u) color = mul(color, it.color);
        if (!should_reflect) {
            normal = assign_normal;
            self_seed += 1u;
        }
        bool reflecting = should_reflect || it.ty == 1u;
        if (!reflecting) {
            if (it.ty == 0u) {
                F3 dir_cos = cosine_weighted_hemisphere(self_seed, it.normal, it.tangent);
                float cosine_pdf = std::fmax(dot(dir_cos, it.normal) / PI, 0.0f);
                self_seed += 3u;
                float avg_pdf = 1.0f * cosine_pdf;
                F3 diffuse = make_diffuse(o.quality, it.color, dir, it.normal, it.tangent, it.rx, dir_cos);
                float cached_brdf = std::fmax(dot(dir_cos, it.normal) / PI, 0.0f);
                float nf = std::fabs(cached_brdf) / std::fmax(avg_pdf, 0.001f);
                color = mul(scale(color, nf), diffuse);
                dir = dir_cos;
            } else if (it.ty == 2u) {
                dir = refract(dir, normal, ri);
            }
        } else {
            dir = reflect(dir, normal);
            if (dot(dir, it.normal) < 0.0f) dir = reflect(dir, it.normal);
        }
        if (i + 1 < o.bounces && i >= o.rr_start) {
            float q = std::fmin(std::fmax(std::fmax(color.x, color.y), color.z), 1.0f);
            float r = orc_rand_f32(self_seed);
            self_seed += 1u;
            if (r >= q) break;
            color = divs(color, q);
        }
    }
    return out;
}

// ===================================================================== guided
// The reference's full rt_main / rt_sample: Markov-chain (vMF) guiding over
// 21 candidates, the world hash grid + radiance cache, the ReSTIR reservoir.
// Interleaving: sample-major; within a sample pixels run one after another in
// row order, each updating the world grid immediately; subgroup shuffles of
// the tentative chains act on the values at sample start (lockstep lanes).

struct Chain {
    F3 ls;
    float mean_cosine, weight_sum;
    uint32_t num_samples;
    float score;
};
Chain czero() { return Chain{splat(0.0f), 0.0f, 0.0f, 0u, 0.0f}; }
Chain chain_from_words(const uint32_t* w) {
    Chain c;
    float f[7];
    std::memcpy(f, w, 28);
    c.ls = mk(f[0], f[1], f[2]);
    c.mean_cosine = f[3];
    c.weight_sum = f[4];
    std::memcpy(&c.num_samples, w + 5, 4);
    std::memcpy(&c.score, w + 6, 4);
    return c;
}
void chain_to_words(const Chain& c, uint32_t* w) {
    float f[5] = {c.ls.x, c.ls.y, c.ls.z, c.mean_cosine, c.weight_sum};
    std::memcpy(w, f, 20);
    std::memcpy(w + 5, &c.num_samples, 4);
    std::memcpy(w + 6, &c.score, 4);
    w[7] = 0u;
}

struct WorldC {  // bindings.wesl:33-40, 56 bytes
    uint32_t secondary_hash, num_samples;
    float lum[3];
    uint32_t lock;
    float cls[3];
    float cmc, cws;
    uint32_t cn;
    float cscore;
    uint32_t timeout;
};
static_assert(sizeof(WorldC) == 56, "world chain layout");

float strip_nan(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    if ((b & 0x7F800000u) == 0x7F800000u) return (b & 0x007FFFFFu) == 0u ? 10000.0f : 0.0f;
    return f;
}
F3 strip_nan3(F3 v) { return mk(strip_nan(v.x), strip_nan(v.y), strip_nan(v.z)); }

}  // namespace
extern "C" uint32_t orc_normalised_to_u32(const float* n);
namespace {

uint32_t n2u(F3 n) {
    float a[3] = {n.x, n.y, n.z};
    return orc_normalised_to_u32(a);
}

int32_t f2i_sat(float f) {
    if (!(f == f)) return 0;
    if (f >= 2147483520.0f) return 2147483647;
    if (f <= -2147483648.0f) return (int32_t)0x80000000;
    return (int32_t)f;
}

struct GKey {
    uint32_t idx, secondary;
    uint32_t primary;   // before the modulo
    uint32_t loc[3];    // masked level-of-detail cell (i32 bits)
};

// general.wesl:148-188: level-of-detail cell, normal and incoming octant
// hashed with the prime multiplies; floor(log2(x)) for x >= 1 from frexp.
GKey grid_key(F3 pos, F3 normal, F3 incoming, F3 cam, uint32_t table) {
    F3 d = sub(pos, cam);
    float len = std::fmax(std::fabs(d.x), std::fmax(std::fabs(d.y), std::fabs(d.z)));
    float x = std::fmax(len / 1.0f, 1.0f);
    uint32_t shifts;
    if (std::isnan(x)) {
        shifts = 0;
    } else if (std::isinf(x)) {
        shifts = 127;
    } else {
        int e;
        std::frexp(x, &e);
        shifts = (uint32_t)(e - 1);
    }
    uint32_t mask = 0xFFFFFFFFu << (shifts < 30u ? shifts : 30u);
    uint32_t lx = (uint32_t)(f2i_sat(pos.x / 0.1f)) & mask;
    uint32_t ly = (uint32_t)(f2i_sat(pos.y / 0.1f)) & mask;
    uint32_t lz = (uint32_t)(f2i_sat(pos.z / 0.1f)) & mask;
    uint32_t n = n2u(normal) & 0xFF00FF00u;
    uint32_t inc_hash = (incoming.x > 0.0f ? 1u : 0u) | (incoming.y > 0.0f ? 2u : 0u) |
                        (incoming.z > 0.0f ? 4u : 0u);
    uint32_t inc = orc_rand_u32(inc_hash);
    uint32_t primary = ((((lx * 1u) ^ (ly * 687603047u)) ^ (lz * 571899763u)) ^ (n * 894271171u)) ^
                       (inc * 840057787u);
    uint32_t s3 = ((lx * 1699644091u) ^ (ly * 550105817u)) ^ (lz * 956114303u);
    uint32_t secondary = ((orc_rand_u32(s3) ^ (n * 155096983u)) ^ 0u) | 1u;
    return GKey{primary % table, secondary, primary, {lx, ly, lz}};
}

struct Grid {
    WorldC* cur;
    const WorldC* old;
    uint32_t table;
    F3 cam;
};

Chain world_chain(const WorldC& w) {
    return Chain{mk(w.cls[0], w.cls[1], w.cls[2]), w.cmc, w.cws, w.cn, w.cscore};
}

// general.wesl:192-226
Chain grid_load(Grid& g, const GKey& k, bool force_old) {
    Chain s = world_chain(g.cur[k.idx]);
    if (g.cur[k.idx].secondary_hash != k.secondary || force_old) {
        if (g.old[k.idx].secondary_hash == k.secondary)
            s = world_chain(g.old[k.idx]);
        else
            s = czero();
    } else {
        g.cur[k.idx].timeout = 30u;
    }
    s.ls = strip_nan3(s.ls);
    s.mean_cosine = strip_nan(s.mean_cosine);
    s.weight_sum = strip_nan(s.weight_sum);
    s.score = strip_nan(s.score);
    return s;
}

// general.wesl:228-246
void grid_store(Grid& g, const GKey& k, const Chain& s) {
    WorldC& w = g.cur[k.idx];
    if (w.secondary_hash != k.secondary) {
        if (w.secondary_hash != 0u) return;
        w.secondary_hash = k.secondary;
    }
    w.cls[0] = s.ls.x;
    w.cls[1] = s.ls.y;
    w.cls[2] = s.ls.z;
    w.cmc = s.mean_cosine;
    w.cws = s.weight_sum;
    w.cn = s.num_samples;
    w.cscore = s.score;
    w.timeout = 30u;
}

// general.wesl:248-271
F3 grid_radiance(Grid& g, const GKey& k) {
    uint32_t n = g.cur[k.idx].num_samples;
    F3 rad = mk(g.cur[k.idx].lum[0], g.cur[k.idx].lum[1], g.cur[k.idx].lum[2]);
    if (g.cur[k.idx].secondary_hash != k.secondary) {
        if (g.old[k.idx].secondary_hash == k.secondary) {
            n = g.old[k.idx].num_samples;
            rad = mk(g.old[k.idx].lum[0], g.old[k.idx].lum[1], g.old[k.idx].lum[2]);
        } else {
            return splat(0.0f);
        }
    } else {
        g.cur[k.idx].timeout = 30u;
    }
    if (n == 0u) return splat(0.0f);
    F3 r = divs(rad, (float)n);
    return mk(std::fmin(std::fmax(r.x, 0.0f), 1000.0f), std::fmin(std::fmax(r.y, 0.0f), 1000.0f),
              std::fmin(std::fmax(r.z, 0.0f), 1000.0f));
}

// general.wesl:273-311 (no lock contention in a sequential run)
void grid_add(Grid& g, const GKey& k, F3 radiance) {
    WorldC& w = g.cur[k.idx];
    F3 c = mk(std::fmin(std::fmax(radiance.x, 0.0f), 100.0f), std::fmin(std::fmax(radiance.y, 0.0f), 100.0f),
              std::fmin(std::fmax(radiance.z, 0.0f), 100.0f));
    if (w.secondary_hash != k.secondary) {
        if (w.secondary_hash != 0u) return;
        w.secondary_hash = k.secondary;
    }
    if (w.lock == 1u) return;
    if (w.num_samples >= 0x0FFFFFFFu) {
        w.timeout = 30u;
        return;
    }
    w.num_samples += 1u;
    w.lum[0] = std::fmax(strip_nan(c.x + w.lum[0]), 0.0f);
    w.lum[1] = std::fmax(strip_nan(c.y + w.lum[1]), 0.0f);
    w.lum[2] = std::fmax(strip_nan(c.z + w.lum[2]), 0.0f);
    w.timeout = 30u;
}

F3 safe_div3(F3 v, float d) { return near_zero(d) ? splat(0.0f) : divs(v, d); }
F3 light_dir(const Chain& s, F3 cur) { return sub(safe_div3(s.ls, s.weight_sum), cur); }
F3 mix3(F3 a, F3 b, float t) { return add(scale(a, 1.0f - t), scale(b, t)); }
float mix1(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// general.wesl:464-492
Chain update_state(F3 x1, F3 x2, float p_hat, uint32_t seed, bool not_vmf, Chain state, float conf,
                   float sum_mc) {
    if (not_vmf) state = czero();
    float rnd = orc_rand_f32(seed);
    if (rnd * sum_mc < p_hat * 21.0f) {
        if (near_zero(length(light_dir(state, x1)))) return state;
        state.num_samples = std::min(state.num_samples + 1u, 1024u);
        float alpha = state.num_samples == 1u ? 1.0f
                                              : std::fmax(1.0f / (float)state.num_samples, 0.01f) * conf;
        state.weight_sum = mix1(state.weight_sum, p_hat, alpha);
        state.ls = mix3(state.ls, scale(x2, p_hat), alpha);
        F3 mean = normalize(light_dir(state, x1));
        state.mean_cosine = mix1(state.mean_cosine, p_hat * dot(normalize(sub(x2, x1)), mean), alpha);
    }
    return state;
}

// general.wesl:1000-1047
float get_kappa(const Chain& s) {
    float n2 = (float)(s.num_samples * s.num_samples);
    float r = ((n2 * (s.mean_cosine / s.weight_sum)) + (0.2f * 0.99f)) / (n2 + 0.2f);
    float k = ((3.0f * r) - ((r * r) * r)) / (1.0f - (r * r));
    return std::fmin(std::fmax(k, 0.001f), 8.50705867e37f);
}
float expm1_ref(float x) {
    float u = std::exp(x);
    if (u == 1.0f) return x;
    float y = u - 1.0f;
    if (std::fabs(x) < 1.0f) return (x * y) / std::log(u);
    return y;
}
float x_div_expm1(float x) {
    float u = std::exp(x);
    if (u == 1.0f) return 1.0f;
    float y = u - 1.0f;
    if (std::fabs(x) < 1.0f) return std::log(u) / y;
    return x / y;
}
float vmf_pdf(F3 mean, float kappa, F3 d) {
    float fraction = x_div_expm1(-2.0f * kappa) / (4.0f * PI);
    return std::fmax(fraction * std::exp(kappa * (dot(d, mean) - 1.0f)), 0.0f);
}
F3 from_onb_no_tangent(F3 src, F3 up) {
    F3 other = near_zero(up.x) ? mk(1.0f, 0.0f, 0.0f) : mk(0.0f, 1.0f, 0.0f);
    return from_onb(src, up, normalize(cross(up, other)));
}

// general.wesl:797-808: Jakob (2012) vMF sample around `mean` from the two
// draws at seed and seed + 1; returns the unnormalised Xs (the caller
// normalises it and tests its length).
F3 vmf_sample(F3 mean, float kappa, uint32_t seed) {
    float theta = orc_rand_f32(seed) * TAU;
    float vx = std::cos(theta), vy = std::sin(theta);
    float xi = orc_rand_f32(seed + 1u);
    float wm1 = std::log(1.0f + xi * expm1_ref(-2.0f * kappa)) / kappa;
    float W = 1.0f + wm1;
    float norm_v = std::sqrt(-std::fma(wm1, wm1, 2.0f * wm1));
    return from_onb_no_tangent(mk(strip_nan(vx * norm_v), W, strip_nan(vy * norm_v)), mean);
}

constexpr int kNumResamples = 21;  // general.wesl:341-345

// general.wesl:404-425 resample_mcmm: one pass of weighted reservoir
// selection over the 21 candidates; returns the selected index (-1 = none,
// the zero chain) and leaves the running weight sum in *sum_mc.
int resample_mcmm(const Chain* all, uint32_t* seed, float* sum_mc) {
    int sel = -1;
    float sum = 0.0f;
    uint32_t s = *seed;
    for (int k = 0; k < kNumResamples; ++k) {
        s = orc_rand_u32(s);
        sum += all[k].weight_sum;
        if (orc_rand_f32(s) * sum < all[k].weight_sum) sel = k;
        s = orc_rand_u32(s);
    }
    *seed = s;
    *sum_mc = sum;
    return sel;
}

// general.wesl:984-998 all_vmf_pdfs: the mixture pdf of direction d over the
// valid candidates, each weighted by weight_sum / sum_mc.
float all_vmf_pdfs(const Chain* all, float sum_mc, F3 hit, F3 d) {
    float pdfs = 0.0f;
    if (sum_mc > 0.0f) {
        for (int k = 0; k < kNumResamples; ++k) {
            if (!(all[k].weight_sum > 0.0f)) continue;
            pdfs += vmf_pdf(normalize(light_dir(all[k], hit)), get_kappa(all[k]), d) *
                    (all[k].weight_sum / sum_mc);
        }
    }
    return pdfs;
}

// fill_mc_screen_space_states (general.wesl:367-385): the lane drawn once
// and used 4x (the first loop does not advance the seed), then 12
// previous-frame pixels in a 91 x 91 window, indices wrapping as u32.
struct ScreenDraw {
    uint32_t lane_pix;
    bool lane_ok;
    uint32_t idx[12];
    uint32_t seed_after;
};
ScreenDraw screen_draw(uint32_t seed, uint32_t x, uint32_t y, uint32_t width, uint32_t subgroup) {
    ScreenDraw d;
    uint32_t base = (x / subgroup) * subgroup;
    uint32_t lane = orc_rand_u32(seed) % subgroup;
    d.lane_pix = y * width + base + lane;
    d.lane_ok = base + lane < width;
    for (int k = 0; k < 12; ++k) {
        uint32_t xr = orc_rand_u32(seed);
        seed = xr;
        uint32_t yr = orc_rand_u32(seed);
        seed = yr;
        int32_t dx = (int32_t)(xr % 91u) - 45, dy = (int32_t)(yr % 91u) - 45;
        uint32_t nx = (uint32_t)((int32_t)x + dx), ny = (uint32_t)((int32_t)y + dy);
        d.idx[k] = nx + ny * width;
    }
    d.seed_after = seed;
    return d;
}

struct GuidedOpts {
    uint32_t quality, bounces, rr_start, samples, restir, subgroup;
    float tmin, tmax;
};

struct SampleRet {
    F3 color, normal, albedo, point, emission, visible_point, visible_normal, out_radiance;
    uint32_t ty;
    float pdf;
    uint32_t vertices;
};

// general.wesl:643-975 rt_sample (cands[0..15] = screen-space candidates)
SampleRet rt_sample_guided(const Scene& s, Grid& grid, const GuidedOpts& o, F3 org, F3 dir,
                           uint32_t own_seed, const Chain* cands, Chain& tent, float& sum_mc) {
    SampleRet out;
    std::memset(&out, 0, sizeof out);
    out.ty = 0xFFu;
    out.pdf = 1.0f;
    uint32_t self_seed = own_seed;
    F3 color = splat(1.0f), orc = splat(1.0f);
    float picked = -1.0f;
    F3 last_normal = splat(0.0f), last_dir = splat(0.0f);
    float cached_pdf = 0.0f, cached_brdf = 0.0f;
    bool last_not_vmf = false;
    F3 accum_emission = splat(0.0f), accum_albedo = splat(1.0f);
    const F3 start_dir = dir;
    for (uint32_t i = 0; i < o.bounces; ++i) {
        out.vertices++;
        Hit it = ray_hit(s, org, dir, o.tmin, o.tmax);
        if (it.hit) {
            F3 new_pos = fma3(dir, splat(it.t), org);
            if (i != 0) {
                GKey kn = grid_key(new_pos, it.normal, dir, grid.cam, grid.table);
                F3 reflected = strip_nan3(divs(scale(mul(it.color, grid_radiance(grid, kn)), cached_brdf), cached_pdf));
                F3 rout = divs(scale(it.emission, cached_brdf), cached_pdf);
                GKey kp = grid_key(org, last_normal, last_dir, grid.cam, grid.table);
                grid_add(grid, kp, add(rout, reflected));
                float weight = length(add(rout, reflected));
                if (weight > 0.0f) {
                    Chain st = update_state(org, new_pos, weight, self_seed, false,
                                            grid_load(grid, kp, false), 1.0f, sum_mc);
                    grid_store(grid, kp, st);
                    tent = update_state(org, new_pos, length(rout), self_seed, last_not_vmf, tent,
                                        i == 1 ? 1.0f : 0.5f, sum_mc);
                    self_seed++;
                }
            }
            last_normal = it.normal;
            last_dir = dir;
            last_not_vmf = false;
            org = new_pos;
            accum_emission = fma3(it.emission, accum_albedo, accum_emission);
            accum_albedo = mul(accum_albedo, it.color);
            if (i == 0) {
                out.point = org;
                out.emission = accum_emission;
                out.ty = it.ty;
                out.normal = it.normal;
                out.albedo = accum_albedo;
            } else if (i == 1) {
                out.visible_point = org;
                out.visible_normal = it.normal;
            }
            out.color = fma3(color, it.emission, out.color);
            if (i != 0) out.out_radiance = fma3(orc, it.emission, out.out_radiance);
            F3 assign_normal = get_normal(o.quality, dir, it.normal, it.tangent, self_seed, it.rx, it.ry);
            F3 transparent_normal = get_normal(o.quality, dir, it.normal, it.tangent, self_seed, it.rtx, it.rty);
            F3 normal = transparent_normal;
            bool should_reflect = false;
            self_seed += 1u;
            F3 ccol;
            float ri = pick_ref_idx(o.quality, &picked, it.n_lo, it.n_hi, self_seed, &ccol);
            self_seed += 1u;
            if (!(!it.front_face && it.ty == 2u)) ri = 1.0f / ri;
            color = mul(color, ccol);
            orc = mul(orc, ccol);
            if (ri != 1.0f) {
                float cos_theta = std::fmin(dot(neg(dir), normal), 1.0f);
                float sin_theta = std::sqrt(1.0f - cos_theta * cos_theta);
                bool cannot = ri * sin_theta > 1.0f;
                float pdf = reflectance(cos_theta, ri);
                if (cannot || pdf > orc_rand_f32(self_seed)) should_reflect = true;
                self_seed += 1u;
            }
            if (!should_reflect || it.ty == 2u) {
                color = mul(color, it.color);
                if (i != 0) orc = mul(orc, it.color);
            }
            if (!should_reflect) {
                normal = assign_normal;
                self_seed += 1u;
            }
            // fill_mc_world_space_states (general.wesl:389-400)
            Chain all[21];
            for (int k = 0; k < 16; ++k) all[k] = cands[k];
            for (int k = 16; k < 20; ++k) {
                bool rand_candidate = orc_rand_f32(self_seed) < 0.7f;
                float jitter_mul = rand_candidate ? 0.75f : 1.0f;
                self_seed = orc_rand_u32(self_seed);
                float jx = (float)self_seed * (1.0f / 4294967296.0f);
                self_seed = orc_rand_u32(self_seed);
                float jz = (float)self_seed * (1.0f / 4294967296.0f);
                F3 jitter = scale(scale(mk(jx - std::floor(jx), 0.0f, jz - std::floor(jz)), 0.1f * 1.5f), jitter_mul);
                GKey k2 = grid_key(add(org, from_onb_no_tangent(jitter, it.normal)), it.normal, dir,
                                   grid.cam, grid.table);
                all[k] = grid_load(grid, k2, rand_candidate);
                self_seed++;
            }
            all[20] = tent;
            const int sel_k = resample_mcmm(all, &self_seed, &sum_mc);
            const Chain sel = sel_k < 0 ? czero() : all[sel_k];
            bool reflecting = should_reflect || it.ty == 1u;
            if (!reflecting) {
                if (it.ty == 0u) {
                    F3 ld = light_dir(sel, org);
                    F3 mean = normalize(ld);
                    float kappa = get_kappa(sel);
                    F3 dir_cos = cosine_weighted_hemisphere(self_seed, it.normal, it.tangent);
                    float cos_pdf = std::fmax(dot(dir_cos, it.normal) / PI, 0.0f);
                    F3 Xs = vmf_sample(mean, kappa, self_seed);
                    self_seed += 2u;
                    F3 dir_markov = normalize(Xs);
                    float c0 = 0.2f, c1 = 0.8f;
                    if (!(sel.weight_sum > 0.0f) || near_zero(length(ld)) || near_zero(sel.weight_sum) ||
                        near_zero(length(Xs))) {
                        c0 = 1.0f;
                        c1 = 0.0f;
                    }
                    float csum = c0 + c1;
                    float conf0 = c0 / csum, conf1 = c1 / csum;
                    float rnd = orc_rand_f32(self_seed);
                    self_seed += 1u;
                    F3 d;
                    if (rnd < conf0) {
                        d = dir_cos;
                        last_not_vmf = true;
                    } else {
                        d = dir_markov;
                    }
                    float mix_pdf = conf1 != 0.0f ? all_vmf_pdfs(all, sum_mc, org, d) : 0.0f;
                    float avg_pdf = conf0 * cos_pdf + (conf1 == 0.0f ? 0.0f : conf1 * std::fmax(mix_pdf, 0.0f));
                    F3 diffuse = make_diffuse(o.quality, it.color, dir, it.normal, it.tangent, it.rx, d);
                    cached_pdf = avg_pdf;
                    cached_brdf = std::fmax(dot(d, it.normal) / PI, 0.0f);
                    float nf = std::fabs(cached_brdf) / std::fmax(avg_pdf, 0.001f);
                    color = mul(scale(color, nf), diffuse);
                    orc = mul(scale(orc, nf), diffuse);
                    dir = d;
                    if (i == 0) out.pdf = avg_pdf;
                } else if (it.ty == 2u) {
                    cached_pdf = 1.0f;
                    cached_brdf = 1.0f;
                    dir = refract(dir, normal, ri);
                }
            } else {
                cached_pdf = 1.0f;
                cached_brdf = 1.0f;
                dir = reflect(dir, normal);
                if (dot(dir, it.normal) < 0.0f) dir = reflect(dir, it.normal);
            }
            if (i + 1 < o.bounces && i >= o.rr_start) {
                float q = std::fmin(std::fmax(std::fmax(color.x, color.y), color.z), 1.0f);
                float r = orc_rand_f32(self_seed);
                self_seed += 1u;
                if (r >= q) break;
                color = divs(color, q);
                orc = divs(orc, q);
            }
        } else {
            F3 bg = background(s, dir);
            F3 new_pos = fma3(dir, splat(o.tmax), org);
            if (i == 1) {
                out.visible_point = new_pos;
                out.visible_normal = neg(dir);
            }
            if (i != 0) {
                out.out_radiance = fma3(orc, bg, out.out_radiance);
                F3 rout = divs(scale(bg, cached_brdf), cached_pdf);
                GKey kp = grid_key(org, last_normal, last_dir, grid.cam, grid.table);
                grid_add(grid, kp, rout);
                float weight = length(rout);
                if (weight > 0.0f) {
                    Chain st = update_state(org, new_pos, weight, self_seed, false,
                                            grid_load(grid, kp, false), 1.0f, sum_mc);
                    grid_store(grid, kp, st);
                    tent = update_state(org, new_pos, length(rout), self_seed, false, tent,
                                        i == 1 ? 1.0f : 0.5f, sum_mc);
                    self_seed++;
                }
            }
            out.color = fma3(color, bg, out.color);
            break;
        }
    }
    // general.wesl:944-948
    GKey k = grid_key(out.point, out.normal, start_dir, grid.cam, grid.table);
    Chain wc = grid_load(grid, k, false);
    Chain st = update_state(out.point, out.visible_point, length(out.out_radiance), self_seed, false,
                            wc, 1.0f, sum_mc);
    grid_store(grid, k, st);
    grid_add(grid, k, out.out_radiance);
    return out;
}

struct Res {  // bindings.wesl:158-169
    float vp[3];
    uint32_t vn;
    float sp[3];
    uint32_t sn;
    float L[3];
    uint32_t ty;
    float rough, w;
    uint32_t mv;
    float W;
};
static_assert(sizeof(Res) == 64, "reservoir layout");

void res_update(Res& R, const Res& smp, float w, uint32_t seed) {
    R.w = R.w + w;
    R.mv = ((R.mv + 1u) & 0x7FFFFFFFu) | (R.mv & 0x80000000u);
    if (orc_rand_f32(seed) * R.w < w) {
        std::memcpy(R.vp, smp.vp, 12);
        R.vn = smp.vn;
        std::memcpy(R.sp, smp.sp, 12);
        R.sn = smp.sn;
        std::memcpy(R.L, smp.L, 12);
        R.ty = smp.ty;
        R.rough = smp.rough;
        R.mv |= 0x80000000u;
    }
}
float res_p_hat(const Res& r) {
    return std::fmax(std::sqrt((r.L[0] * r.L[0] + r.L[1] * r.L[1]) + r.L[2] * r.L[2]), 0.00001f);
}
float safe_div1(float n, float d) { return near_zero(d) ? 0.0f : n / d; }

}  // namespace

extern "C" {

// general.wesl:506-610 over a whole image.  world/world_old: 56-byte chains
// (world is updated in place); mc_old / mc_out: 32-byte chains (mc_count
// entries each); gi_out: 64-byte reservoirs (gi_count entries); info_out:
// 9 floats per pixel.  Returns the number of path vertices.
uint64_t orc_trace_guided(void* h, const float* cam, uint32_t width, uint32_t height,
                          uint32_t quality, uint32_t samples, float tmin, float tmax,
                          uint32_t bounces, uint32_t rr_start, uint32_t seed_offset, uint32_t restir,
                          uint32_t subgroup, void* world, const void* world_old, uint32_t world_size,
                          const uint32_t* mc_old, uint32_t* mc_out, uint64_t mc_count, void* gi_out,
                          uint64_t gi_count, float* info_out, float* out_color, float* out_normal,
                          float* out_albedo) {
    const Scene& s = *(const Scene*)h;
    const uint64_t n = (uint64_t)width * height;
    GuidedOpts o{quality, bounces, rr_start, samples, restir, subgroup, tmin, tmax};
    Grid grid{(WorldC*)world, (const WorldC*)world_old, world_size, mk(cam[16 + 12], cam[16 + 13], cam[16 + 14])};
    std::vector<Chain> tent(n), snap(n);
    std::vector<uint32_t> own(n);
    std::vector<float> sum_mc(n, 0.0f), pdf(n, 0.0f);
    std::vector<F3> color_sum(n, splat(0.0f)), out_rad(n, splat(0.0f)), last_n(n), last_a(n), last_e(n);
    std::vector<Res> R(n), smp(n);
    std::vector<uint8_t> smp_valid(n, 0);
    std::memset(R.data(), 0, n * sizeof(Res));
    std::memset(smp.data(), 0, n * sizeof(Res));
    auto mc_at = [&](uint64_t idx) {
        return idx < mc_count ? chain_from_words(mc_old + 8 * idx) : czero();
    };
    for (uint64_t p = 0; p < n; ++p) {
        uint32_t x = (uint32_t)(p % width), y = (uint32_t)(p / width);
        own[p] = orc_rand_u32(x * y) + seed_offset + y;
        tent[p] = mc_at(p);
    }
    uint64_t verts = 0;
    for (uint32_t si = 0; si < samples; ++si) {
        snap = tent;
        for (uint64_t p = 0; p < n; ++p) {
            uint32_t x = (uint32_t)(p % width), y = (uint32_t)(p / width);
            Chain cands[16];
            ScreenDraw sd = screen_draw(own[p], x, y, width, subgroup);
            own[p] = sd.seed_after;
            Chain lc = sd.lane_ok ? snap[sd.lane_pix] : czero();
            for (int k = 0; k < 4; ++k) cands[k] = lc;
            for (int k = 4; k < 16; ++k) cands[k] = mc_at(sd.idx[k - 4]);
            float o3[3], d3[3];
            orc_create_ray(cam, x, y, width, height, o3, d3);
            SampleRet rs = rt_sample_guided(s, grid, o, mk(o3[0], o3[1], o3[2]), mk(d3[0], d3[1], d3[2]),
                                            own[p], cands, tent[p], sum_mc[p]);
            verts += rs.vertices;
            last_n[p] = rs.normal;
            last_a[p] = rs.albedo;
            last_e[p] = rs.emission;
            color_sum[p] = add(color_sum[p], rs.color);
            out_rad[p] = add(out_rad[p], rs.out_radiance);
            own[p] = orc_rand_u32(own[p]);
            Res cur;
            std::memset(&cur, 0, sizeof cur);
            cur.vp[0] = rs.visible_point.x;
            cur.vp[1] = rs.visible_point.y;
            cur.vp[2] = rs.visible_point.z;
            cur.vn = n2u(rs.visible_normal);
            cur.sp[0] = rs.point.x;
            cur.sp[1] = rs.point.y;
            cur.sp[2] = rs.point.z;
            cur.sn = n2u(rs.normal);
            cur.ty = rs.ty;
            cur.rough = 0.0f;
            if (restir == 2u) {
                F3 L = divs(out_rad[p], (float)samples);
                cur.L[0] = L.x;
                cur.L[1] = L.y;
                cur.L[2] = L.z;
                smp[p] = cur;
                smp_valid[p] = 1;
                pdf[p] += rs.pdf;
            } else if (restir == 1u) {
                cur.L[0] = rs.out_radiance.x;
                cur.L[1] = rs.out_radiance.y;
                cur.L[2] = rs.out_radiance.z;
                if (cur.ty != 0xFFu) {
                    res_update(R[p], cur, res_p_hat(cur), own[p]);
                    own[p] = orc_rand_u32(own[p]);
                    R[p].W = safe_div1(R[p].w, (float)(R[p].mv & 0x7FFFFFFFu) * res_p_hat(R[p]));
                } else {
                    R[p].ty = 0xFFu;
                }
            }
        }
        // subgroupShuffle(tentative, (lane + 1) % sub_size)
        std::vector<Chain> next(n);
        for (uint64_t p = 0; p < n; ++p) {
            uint32_t x = (uint32_t)(p % width), y = (uint32_t)(p / width);
            uint32_t base = (x / subgroup) * subgroup;
            uint32_t active = std::min(subgroup, width - base);
            next[p] = tent[(uint64_t)y * width + base + ((x - base + 1u) % active)];
        }
        tent.swap(next);
    }
    for (uint64_t p = 0; p < n; ++p) {
        if (restir != 0u) {
            float* inf = info_out + 9 * p;
            inf[0] = grid.cam.x;
            inf[1] = grid.cam.y;
            inf[2] = grid.cam.z;
            inf[3] = last_e[p].x;
            inf[4] = last_e[p].y;
            inf[5] = last_e[p].z;
            inf[6] = last_a[p].x;
            inf[7] = last_a[p].y;
            inf[8] = last_a[p].z;
        }
        if (p < mc_count) chain_to_words(tent[p], mc_out + 8 * p);
        float* oc = out_color + 4 * p;
        oc[0] = std::fmax(color_sum[p].x / (float)samples, 0.0f);
        oc[1] = std::fmax(color_sum[p].y / (float)samples, 0.0f);
        oc[2] = std::fmax(color_sum[p].z / (float)samples, 0.0f);
        oc[3] = 1.0f;
        float* on = out_normal + 4 * p;
        on[0] = last_n[p].x;
        on[1] = last_n[p].y;
        on[2] = last_n[p].z;
        on[3] = 1.0f;
        float* oa = out_albedo + 4 * p;
        oa[0] = last_a[p].x;
        oa[1] = last_a[p].y;
        oa[2] = last_a[p].z;
        oa[3] = 1.0f;
        if (restir == 2u) {
            if (smp[p].ty != 0xFFu && smp_valid[p]) {
                res_update(R[p], smp[p], res_p_hat(smp[p]), own[p]);
                own[p] = orc_rand_u32(own[p]);
                R[p].W = safe_div1(R[p].w, (float)(R[p].mv & 0x7FFFFFFFu) * res_p_hat(R[p]));
            } else if (smp[p].ty == 0xFFu) {
                R[p].ty = 0xFFu;
            }
        }
        if (p < gi_count) std::memcpy((char*)gi_out + 64 * p, &R[p], 64);
    }
    return verts;
}

// descs: offset, width, height per texture.
void* orc_scene_create(const float* verts, uint64_t nverts, const uint32_t* idx, uint64_t ntri,
                       const uint32_t* matidx, const uint32_t* mats, uint64_t nmat,
                       const uint32_t* texels, uint64_t ntexels, const uint32_t* descs,
                       uint32_t ntex, uint32_t diffuse_base, uint32_t emission_base,
                       uint32_t attr_base, uint32_t recolour, uint32_t bg_base) {
    Scene* s = new Scene();
    s->verts.assign(verts, verts + 4 * nverts);
    if (idx) s->idx.assign(idx, idx + 3 * ntri);
    s->ntri = ntri;
    s->matidx.assign(matidx, matidx + ntri);
    s->mats.assign(mats, mats + 8 * nmat);
    s->texels.assign(texels, texels + ntexels);
    s->descs.assign(descs, descs + 3 * ntex);
    s->diffuse_base = diffuse_base;
    s->emission_base = emission_base;
    s->attr_base = attr_base;
    s->recolour = recolour;
    s->bg_base = bg_base;
    s->bvh = orc_bvh_create(s->verts.data(), nverts, idx ? s->idx.data() : nullptr, ntri);
    return s;
}

void orc_scene_set_alpha_test(void* h, int on) { ((Scene*)h)->alpha_test = on != 0; }

// obj: 9 floats per primitive (object-space v0, v1, v2); mirrored: 1 byte
// per primitive.
void orc_scene_set_object_space(void* h, const float* obj, const uint8_t* mirrored) {
    Scene* s = (Scene*)h;
    s->obj.assign(obj, obj + 9 * s->ntri);
    s->mirrored.assign(mirrored, mirrored + s->ntri);
}

void orc_scene_destroy(void* h) {
    Scene* s = (Scene*)h;
    orc_bvh_destroy(s->bvh);
    delete s;
}

// Plain-mode trace of a width x height image (or of the pixel rows
// [row0, row0 + rows)).  Outputs rgba per pixel of the traced rows; colour is
// max(mean, 0) (general.wesl:592), accumulated in float64 when fp64 != 0.
// Returns the number of path vertices (closest-hit queries).
uint64_t orc_trace_plain(void* h, const float* cam, uint32_t width, uint32_t height,
                         uint32_t row0, uint32_t rows, uint32_t quality, uint32_t samples,
                         float tmin, float tmax, uint32_t bounces, uint32_t rr_start,
                         uint32_t seed_offset, int fp64, float* out_color, float* out_normal,
                         float* out_albedo) {
    const Scene& s = *(const Scene*)h;
    PlainOpts o{quality, bounces, rr_start, tmin, tmax};
    uint64_t total = 0;
    const int64_t n = (int64_t)width * rows;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total)
    for (int64_t k = 0; k < n; ++k) {
        uint32_t x = (uint32_t)(k % width), y = row0 + (uint32_t)(k / width);
        float o3[3], d3[3];
        orc_create_ray(cam, x, y, width, height, o3, d3);
        uint32_t own_seed = orc_rand_u32(x * y) + seed_offset + y;
        double acc[3] = {0, 0, 0};
        F3 accf = splat(0.0f);
        SampleOut last{};
        for (uint32_t i = 0; i < samples; ++i) {
            SampleOut so = rt_sample_plain(s, o, mk(o3[0], o3[1], o3[2]), mk(d3[0], d3[1], d3[2]),
                                           own_seed);
            total += so.vertices;
            if (fp64) {
                acc[0] += so.color.x;
                acc[1] += so.color.y;
                acc[2] += so.color.z;
            } else {
                accf = add(accf, so.color);
            }
            last = so;
            own_seed = orc_rand_u32(own_seed);
        }
        float c[3];
        if (fp64) {
            for (int j = 0; j < 3; ++j) c[j] = (float)std::fmax(acc[j] / (double)samples, 0.0);
        } else {
            c[0] = std::fmax(accf.x / (float)samples, 0.0f);
            c[1] = std::fmax(accf.y / (float)samples, 0.0f);
            c[2] = std::fmax(accf.z / (float)samples, 0.0f);
        }
        float* oc = out_color + 4 * k;
        oc[0] = c[0];
        oc[1] = c[1];
        oc[2] = c[2];
        oc[3] = 1.0f;
        float* on = out_normal + 4 * k;
        on[0] = last.normal.x;
        on[1] = last.normal.y;
        on[2] = last.normal.z;
        on[3] = 1.0f;
        float* oa = out_albedo + 4 * k;
        oa[0] = last.albedo.x;
        oa[1] = last.albedo.y;
        oa[2] = last.albedo.z;
        oa[3] = 1.0f;
    }
    return total;
}

// The debugging shaders (src/debug/front_face.wesl:52-63,
// reflectance.wesl:49-57 + 92-96, tangent.wesl:52-72 + 111-122): the camera
// ray's closest hit only.  kind 0 front face: front (0, 1/(t+1), 0), back
// (1/(t+1), 0, 0); kind 1 reflectance: (ior_lo / 5, ior_hi / 5, 0); kind 2
// tangent: (correct_tangent(normalize(v0 - v1)) + 1) / 2, the zero vector on
// a miss.  Misses are 0 for kinds 0 and 1.  rgba per pixel, w = 1.
void orc_trace_debug(void* h, const float* cam, uint32_t width, uint32_t height, uint32_t kind,
                     float tmin, float tmax, float* out_color) {
    const Scene& s = *(const Scene*)h;
    const int64_t n = (int64_t)width * height;
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t k = 0; k < n; ++k) {
        uint32_t x = (uint32_t)(k % width), y = (uint32_t)(k / width);
        float o3[3], d3[3];
        orc_create_ray(cam, x, y, width, height, o3, d3);
        Hit hit = ray_hit(s, mk(o3[0], o3[1], o3[2]), mk(d3[0], d3[1], d3[2]), tmin, tmax, false);
        F3 c = splat(0.0f);
        if (kind == 0 && hit.hit) {
            float shade = 1.0f / (hit.t + 1.0f);
            c = hit.front_face ? mk(0.0f, shade, 0.0f) : mk(shade, 0.0f, 0.0f);
        } else if (kind == 1 && hit.hit) {
            c = mk(hit.n_lo / 5.0f, hit.n_hi / 5.0f, 0.0f);
        } else if (kind == 2) {
            F3 t = hit.hit ? hit.tangent : splat(0.0f);
            c = mk((t.x + 1.0f) / 2.0f, (t.y + 1.0f) / 2.0f, (t.z + 1.0f) / 2.0f);
        }
        float* oc = out_color + 4 * k;
        oc[0] = c.x;
        oc[1] = c.y;
        oc[2] = c.z;
        oc[3] = 1.0f;
    }
}

// Test hooks for single functions.
void orc_oren_nayar(const float* incoming, const float* reflected, float roughness,
                    const float* color, float* out) {
    F3 r = oren_nayar(mk(incoming[0], incoming[1], incoming[2]),
                      mk(reflected[0], reflected[1], reflected[2]), roughness,
                      mk(color[0], color[1], color[2]));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}
float orc_reflectance(float cosine, float ref_idx) { return reflectance(cosine, ref_idx); }
void orc_sample_vndf_ggx(uint32_t seed, const float* view, float rx, float ry, float* out) {
    F3 r = sample_vndf_ggx(seed, mk(view[0], view[1], view[2]), rx, ry);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}
void orc_cosine_hemisphere(uint32_t seed, const float* normal, const float* tangent, float* out) {
    F3 r = cosine_weighted_hemisphere(seed, mk(normal[0], normal[1], normal[2]),
                                      mk(tangent[0], tangent[1], tangent[2]));
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}
void orc_high_colour(uint32_t seed, float* out) {
    float picked = -1.0f;
    F3 c;
    pick_ref_idx(2, &picked, 1.0f, 1.0f, seed, &c);
    out[0] = c.x;
    out[1] = c.y;
    out[2] = c.z;
    out[3] = picked;
}
// Spectral-quality hook: pick at a vertex with IOR range (lo, hi) starting
// from *picked (-1 = fresh); returns the IOR, writes the colour weight.
float orc_spectral_pick(float* picked, float lo, float hi, uint32_t seed, float* colour) {
    F3 c;
    float ri = pick_ref_idx(3, picked, lo, hi, seed, &c);
    colour[0] = c.x;
    colour[1] = c.y;
    colour[2] = c.z;
    return ri;
}
void orc_background(void* h, const float* d, float* out) {
    F3 c = background(*(const Scene*)h, mk(d[0], d[1], d[2]));
    out[0] = c.x;
    out[1] = c.y;
    out[2] = c.z;
}

// Guided-mode hooks.  Chains travel as 8 u32 words (bindings.wesl:3-9 plus
// a zero pad word), world-grid cells as 56-byte WorldMarcovChainState.
float orc_strip_nan(float f) { return strip_nan(f); }

// out: idx, secondary, primary (before the modulo), loc.x, loc.y, loc.z.
void orc_grid_key(const float* pos, const float* normal, const float* incoming, const float* cam,
                  uint32_t table, uint32_t* out) {
    GKey k = grid_key(mk(pos[0], pos[1], pos[2]), mk(normal[0], normal[1], normal[2]),
                      mk(incoming[0], incoming[1], incoming[2]), mk(cam[0], cam[1], cam[2]), table);
    out[0] = k.idx;
    out[1] = k.secondary;
    out[2] = k.primary;
    out[3] = k.loc[0];
    out[4] = k.loc[1];
    out[5] = k.loc[2];
}

float orc_get_kappa(const uint32_t* chain) { return get_kappa(chain_from_words(chain)); }
float orc_expm1(float x) { return expm1_ref(x); }
float orc_x_div_expm1(float x) { return x_div_expm1(x); }

void orc_vmf_pdf_many(const float* mean, float kappa, const float* dirs, uint64_t n, float* out) {
    const F3 m = mk(mean[0], mean[1], mean[2]);
    for (uint64_t i = 0; i < n; ++i)
        out[i] = vmf_pdf(m, kappa, mk(dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]));
}

// Unnormalised Xs per seed (general.wesl:797-808).
void orc_vmf_sample_many(const float* mean, float kappa, const uint32_t* seeds, uint64_t n,
                         float* out) {
    const F3 m = mk(mean[0], mean[1], mean[2]);
    for (uint64_t i = 0; i < n; ++i) {
        F3 x = vmf_sample(m, kappa, seeds[i]);
        out[3 * i] = x.x;
        out[3 * i + 1] = x.y;
        out[3 * i + 2] = x.z;
    }
}

void orc_update_state(const float* x1, const float* x2, float p_hat, uint32_t seed, int not_vmf,
                      const uint32_t* chain_in, float confidence, float sum_mc, uint32_t* chain_out) {
    Chain c = update_state(mk(x1[0], x1[1], x1[2]), mk(x2[0], x2[1], x2[2]), p_hat, seed,
                           not_vmf != 0, chain_from_words(chain_in), confidence, sum_mc);
    chain_to_words(c, chain_out);
}

// out: lane pixel, lane valid, 12 neighbour indices, seed after (15 words).
void orc_screen_draw(uint32_t seed, uint32_t x, uint32_t y, uint32_t width, uint32_t subgroup,
                     uint32_t* out) {
    ScreenDraw d = screen_draw(seed, x, y, width, subgroup);
    out[0] = d.lane_pix;
    out[1] = d.lane_ok ? 1u : 0u;
    for (int k = 0; k < 12; ++k) out[2 + k] = d.idx[k];
    out[14] = d.seed_after;
}

// 21 chains in; returns the selected index (-1: none); *seed is advanced.
int orc_resample_mcmm(const uint32_t* chains, uint32_t* seed, float* sum_mc) {
    Chain all[kNumResamples];
    for (int k = 0; k < kNumResamples; ++k) all[k] = chain_from_words(chains + 8 * k);
    return resample_mcmm(all, seed, sum_mc);
}

float orc_all_vmf_pdfs(const uint32_t* chains, float sum_mc, const float* hit, const float* d) {
    Chain all[kNumResamples];
    for (int k = 0; k < kNumResamples; ++k) all[k] = chain_from_words(chains + 8 * k);
    return all_vmf_pdfs(all, sum_mc, mk(hit[0], hit[1], hit[2]), mk(d[0], d[1], d[2]));
}

// One sequential world-grid operation on cur / old (table cells each):
// op 0 grid_load (-> chain_out), 1 grid_store_markov_chain (chain_in),
// 2 grid_load_radiance_estimate (-> rad_out), 3 grid_add_radiance (rad_in).
void orc_grid_op(int op, void* cur, const void* old, uint32_t table, const float* cam,
                 const float* pos, const float* normal, const float* incoming, int force_old,
                 const uint32_t* chain_in, const float* rad_in, uint32_t* chain_out,
                 float* rad_out) {
    Grid g{(WorldC*)cur, (const WorldC*)old, table, mk(cam[0], cam[1], cam[2])};
    GKey k = grid_key(mk(pos[0], pos[1], pos[2]), mk(normal[0], normal[1], normal[2]),
                      mk(incoming[0], incoming[1], incoming[2]), g.cam, table);
    if (op == 0) {
        chain_to_words(grid_load(g, k, force_old != 0), chain_out);
    } else if (op == 1) {
        grid_store(g, k, chain_from_words(chain_in));
    } else if (op == 2) {
        F3 r = grid_radiance(g, k);
        rad_out[0] = r.x;
        rad_out[1] = r.y;
        rad_out[2] = r.z;
    } else {
        grid_add(g, k, mk(rad_in[0], rad_in[1], rad_in[2]));
    }
}

}  // extern "C"
