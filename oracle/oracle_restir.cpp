// oracle_restir.cpp - CPU ORACLE (test infrastructure only) of the ReSTIR GI
// passes and the end-of-frame kernels:
//
//   temporal   importance_sampling/temporal_resampling.wesl:37-64
//   spatial    importance_sampling/spatial_resampling.wesl:35-194 (kept with
//              its candidate array Q[129] as written)
//   reservoir  importance_sampling/shared.wesl:63-168 (update, merge,
//              assign, p_hat, confidence packing)
//   octahedral general.wesl:1293-1304 normalised_to_u32,
//              shared.wesl:233-262 u32_to_normalised / copy_sign
//   EOF        data_buffer/end_of_frame.wesl:14-99
//
// Reservoir layout: 16 x 32-bit words (bindings.wesl:158-169).
// Pinned by (tests/test_cpu_restir.py): the octahedral round trip error
// bound and exactness on the axes, merge selection frequencies proportional
// to the weights (binomial test), W = w / (M p_hat) on 2-sample cases in
// closed form, confidence-cap and decay identities in closed form.
#include <stdint.h>

#include <cmath>
#include <cstring>
#include <vector>

extern "C" {
uint32_t orc_rand_u32(uint32_t s);
float orc_rand_f32(uint32_t s);
}

namespace {

struct R {
    float vp[3];
    uint32_t vn;
    float sp[3];
    uint32_t sn;
    float L[3];
    uint32_t ty;
    float rough;
    float w;
    uint32_t mv;
    float W;
};
static_assert(sizeof(R) == 64, "reservoir layout");

const uint32_t MAX_CONF = 0x7FFFFFFFu;
const uint32_t SKY = 0xFFu;

bool near_zero(float f) { return f < 0.0001f && f > -0.0001f; }
float safe_div(float n, float d) { return near_zero(d) ? 0.0f : n / d; }
float copy_sign(float src, float dst) { return src < 0.0f ? -std::fabs(dst) : std::fabs(dst); }

void unpack_snorm(uint32_t u, float* a, float* b) {
    int16_t lo = (int16_t)(uint16_t)(u & 0xFFFFu), hi = (int16_t)(uint16_t)(u >> 16);
    *a = std::fmax((float)lo / 32767.0f, -1.0f);
    *b = std::fmax((float)hi / 32767.0f, -1.0f);
}
uint32_t pack_snorm(float a, float b) {
    auto q = [](float v) {
        float c = std::fmin(std::fmax(v, -1.0f), 1.0f);
        return (uint32_t)(int32_t)std::floor(0.5f + 32767.0f * c) & 0xFFFFu;
    };
    return q(a) | (q(b) << 16);
}

void u32_to_normalised(uint32_t u, float out[3]) {
    float x, z;
    unpack_snorm(u, &x, &z);
    float y = 1.0f - (std::fabs(x) + std::fabs(z));
    if (y < 0.0f) {
        float x0 = x;
        x = copy_sign(x0, 1.0f - std::fabs(z));
        z = copy_sign(z, 1.0f - std::fabs(x0));
    }
    float l = std::sqrt((x * x + y * y) + z * z);
    out[0] = x / l;
    out[1] = y / l;
    out[2] = z / l;
}

uint32_t normalised_to_u32(const float n[3]) {
    float s = std::fmax((std::fabs(n[0]) + std::fabs(n[1])) + std::fabs(n[2]), 0.01f);
    float v[3] = {n[0] / s, n[1] / s, n[2] / s};
    if (v[1] >= 0.0f) return pack_snorm(v[0], v[2]);
    return pack_snorm(copy_sign(v[0], 1.0f - std::fabs(v[2])), copy_sign(v[2], 1.0f - std::fabs(v[0])));
}

float dot3(const float a[3], const float b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

float p_hat(const R& r) { return std::fmax(std::sqrt(dot3(r.L, r.L)), 0.00001f); }
uint32_t conf(uint32_t mv) { return mv & MAX_CONF; }

void assign(R& dst, const R& s) {
    std::memcpy(dst.vp, s.vp, 12);
    dst.vn = s.vn;
    std::memcpy(dst.sp, s.sp, 12);
    dst.sn = s.sn;
    std::memcpy(dst.L, s.L, 12);
    dst.ty = s.ty;
    dst.rough = s.rough;
    dst.mv = (dst.mv & MAX_CONF) | (1u << 31);
}

// shared.wesl:99-115
bool merge(R& res, const R& r, float p, uint32_t seed) {
    uint32_t c = conf(res.mv), v = res.mv >> 31, oc = conf(r.mv);
    float w_new = p * r.W * (float)oc;
    c += oc;
    res.mv = (c & MAX_CONF) | (v << 31);
    res.w = res.w + w_new;
    if (orc_rand_f32(seed) * res.w < w_new) {
        assign(res, r);
        return true;
    }
    return false;
}

// shared.wesl:82-92
void update(R& res, const R& s_new, float w_new, uint32_t seed) {
    res.w = res.w + w_new;
    uint32_t c = conf(res.mv) + 1u, v = res.mv >> 31;
    res.mv = (c & MAX_CONF) | (v << 31);
    if (orc_rand_f32(seed) * res.w < w_new) assign(res, s_new);
}

float len_sqrd(const float v[3]) { return std::fma(v[0], v[0], std::fma(v[1], v[1], v[2] * v[2])); }

void normalize(float v[3]) {
    float l = std::sqrt(dot3(v, v));
    v[0] /= l;
    v[1] /= l;
    v[2] /= l;
}

float jacobian(const R& q, const R& r) {
    float q1[3], r1[3], nv[3];
    for (int k = 0; k < 3; ++k) {
        q1[k] = q.sp[k] - q.vp[k];
        r1[k] = r.sp[k] - q.vp[k];
    }
    u32_to_normalised(q.vn, nv);
    float nr[3] = {r1[0], r1[1], r1[2]}, nq[3] = {q1[0], q1[1], q1[2]};
    normalize(nr);
    normalize(nq);
    float cr = dot3(nr, nv), cq = dot3(nq, nv);
    return safe_div(cr * len_sqrd(q1), cq * len_sqrd(r1));
}

}  // namespace

extern "C" {

uint32_t orc_normalised_to_u32(const float* n) { return normalised_to_u32(n); }
void orc_u32_to_normalised(uint32_t u, float* out) { u32_to_normalised(u, out); }

// Merges reservoir r into res with target weight p (64-byte records).
int orc_merge(void* res, const void* r, float p, uint32_t seed) {
    return merge(*(R*)res, *(const R*)r, p, seed) ? 1 : 0;
}
void orc_update(void* res, const void* s, float w, uint32_t seed) {
    update(*(R*)res, *(const R*)s, w, seed);
}

// temporal_resampling.wesl:37-64 over a width x height image.
void orc_temporal(void* gi_v, const void* old_v, uint32_t width, uint32_t height,
                  uint32_t seed_offset) {
    R* gi = (R*)gi_v;
    const R* old = (const R*)old_v;
    for (uint32_t y = 0; y < height; ++y)
        for (uint32_t x = 0; x < width; ++x) {
            uint64_t idx = x + (uint64_t)y * width;
            uint32_t own_seed = orc_rand_u32(x * y) + orc_rand_u32(seed_offset + y);
            R Rs = gi[idx], Rr = old[idx];
            float a[3], b[3];
            u32_to_normalised(Rr.sn, a);
            u32_to_normalised(Rs.sn, b);
            if (dot3(a, b) < 0.95f) std::memset(&Rr, 0, sizeof Rr);
            if (Rs.ty == SKY) continue;
            float w = p_hat(Rr);
            merge(Rs, Rr, w, own_seed);
            Rs.W = safe_div(Rs.w, (float)conf(Rs.mv) * p_hat(Rs));
            gi[idx] = Rs;
        }
}

// spatial_resampling.wesl:52-194; info: 9 floats per pixel; out: rgba.
void orc_spatial(const void* gi_v, const float* info, uint32_t width, uint32_t height,
                 uint32_t is_samples, uint32_t is_space, uint32_t seed_offset, float* out,
                 int32_t* written) {
    const R* gi = (const R*)gi_v;
    std::vector<R> Q(is_samples + 1);
    for (uint32_t y = 0; y < height; ++y)
        for (uint32_t x = 0; x < width; ++x) {
            uint64_t idx = x + (uint64_t)y * width;
            written[idx] = 0;
            uint32_t own_seed = orc_rand_u32(x * y) + orc_rand_u32(seed_offset) + y;
            R Rs = gi[idx];
            float Rs_normal[3];
            u32_to_normalised(Rs.sn, Rs_normal);
            Q[0] = Rs;
            if (Rs.ty == SKY) continue;
            uint32_t k = 0;
            for (uint32_t i = 0; i < is_samples; ++i) {
                uint32_t max_val = is_space * is_space - 1u;
                uint32_t min_space = is_space / 2u;
                uint32_t self_idx = min_space + min_space * is_space;
                uint32_t rnd = orc_rand_u32(own_seed) % max_val;
                if (rnd >= self_idx) rnd = rnd + 1u;
                int32_t nx = (int32_t)(rnd % is_space) - (int32_t)min_space;
                int32_t ny = (int32_t)(rnd / is_space) - (int32_t)min_space;
                own_seed = orc_rand_u32(own_seed);
                int32_t ox = (int32_t)x + nx, oy = (int32_t)y + ny;
                if (ox < 0 || oy < 0 || (nx == 0 && ny == 0)) continue;
                if ((uint32_t)ox >= width || (uint32_t)oy >= height) continue;
                R Rn = gi[(uint64_t)ox + (uint64_t)oy * width];
                float Rn_normal[3];
                u32_to_normalised(Rn.sn, Rn_normal);
                float sim = dot3(Rn_normal, Rs_normal);
                if (sim < 0.9f || Rn.ty != Rs.ty || std::fabs(Rs.rough - Rn.rough) > 0.1f) continue;
                float jac = jacobian(Rn, Rs);
                if (Rs.ty != 0u) continue;
                float phd = p_hat(Rn) * jac;
                bool picked = merge(Rs, Rn, phd, own_seed);
                own_seed = orc_rand_u32(own_seed);
                if (picked) u32_to_normalised(Rs.sn, Rs_normal);
                Q[k + 1] = Rn;
                own_seed = orc_rand_u32(own_seed);
                ++k;
            }
            uint32_t Z = 0;
            for (uint32_t j = 0; j < k + 1; ++j)
                if (p_hat(Q[j]) > 0.0f) Z += conf(Q[j].mv);
            float zp = (float)Z * p_hat(Rs);
            Rs.W = safe_div(Rs.w, zp);
            const float* inf = info + 9 * idx;
            float* o = out + 4 * idx;
            for (int c = 0; c < 3; ++c) o[c] = std::fma(Rs.L[c] * Rs.W, inf[6 + c], inf[3 + c]);
            o[3] = 1.0f;
            written[idx] = 1;
        }
}

// end_of_frame.wesl:14-31
void orc_eof_gi(void* res_v, uint64_t n, uint32_t cap) {
    R* res = (R*)res_v;
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t c = conf(res[i].mv);
        if (c > cap) {
            float mult = (float)cap / (float)c;
            res[i].w *= mult;
            res[i].mv = (cap & MAX_CONF) | ((res[i].mv >> 31) << 31);
        }
    }
}

// end_of_frame.wesl:73-99 on 14-word world chains
void orc_eof_world(uint32_t* w, uint64_t n) {
    const uint32_t mul = 8u, div = 9u;
    for (uint64_t i = 0; i < n; ++i) {
        uint32_t* c = w + 14 * i;
        if (c[13] == 0u) {
            c[0] = c[1] = c[2] = c[3] = c[4] = c[5] = 0u;
            continue;
        }
        uint32_t old_n = c[1];
        if (old_n != 0u) {
            uint32_t prod = old_n * (mul < div ? mul : div);
            uint32_t new_n = prod / div + (prod % div != 0u ? 1u : 0u);
            float m = (float)new_n / (float)old_n;
            for (int k = 0; k < 3; ++k) {
                float f;
                std::memcpy(&f, &c[2 + k], 4);
                f *= m;
                std::memcpy(&c[2 + k], &f, 4);
            }
            c[1] = new_n;
        }
        c[13] -= 1u;
    }
}

}  // extern "C"
