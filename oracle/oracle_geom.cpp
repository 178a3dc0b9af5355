// oracle_geom.cpp - CPU ORACLE (test infrastructure only; never linked into
// or called by the product library).  Restates, in plain C++/OpenMP:
//
//   * rand_u32 / rand_f32        importance_sampling/shared.wesl:266-284
//   * create_ray                  path_tracing/general.wesl:1158-1173
//   * closest-hit ray query       general.wesl:1052-1078 (semantics of the
//     Vulkan ray query the reference calls: closest hit with parametric t in
//     [T_MIN, T_MAX], barycentrics weighting vertices 1 and 2), with the
//     watertight triangle test of Woop, Benthin & Wald (JCGT 2013), ties in t
//     resolved to the lower primitive index.
//   * a binary median-split BVH used only to make the same closest-hit query
//     fast on large scenes; it returns exactly what the brute force returns.
//
// Compiled with -ffp-contract=off: every fused multiply-add is explicit.
//
// How each function is pinned (tests/test_oracle_*.py):
//   rand_u32      numpy restatement on 2^20 inputs + bijectivity on a 2^24 range
//                 + uniformity (chi-square) of rand_f32
//   create_ray    float64 closed form within a few ulp; direction length
//                 1/sqrt(1 + (w/|xyz|)^2) (the 4-D normalize quirk)
//   brute force   float64 Moller-Trumbore nearest hit on random soups (IDs
//                 agree except near ties), shared-edge watertightness, t
//                 range bounds
//   oracle BVH    identical to the brute force on every ray
#include <omp.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

extern "C" {

uint32_t orc_rand_u32(uint32_t s) {
    uint32_t state = s * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

float orc_rand_f32(uint32_t s) {
    // f32(u32) rounds to nearest even; RECIP_U32_MAX = 1 / f32(0xFFFFFFFF) = 2^-32.
    float f = (float)orc_rand_u32(s) * (1.0f / 4294967296.0f);
    return f - std::floor(f);
}

void orc_rand_u32_many(const uint32_t* in, uint64_t n, uint32_t* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) out[i] = orc_rand_u32(in[i]);
}

void orc_rand_f32_many(const uint32_t* in, uint64_t n, float* out) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) out[i] = orc_rand_f32(in[i]);
}

// cam: projection_inverse[16], view_inverse[16], column-major (camera.rs:8-11).
static void mat_vec(const float* m, float x, float y, float z, float w, float* out) {
    for (int r = 0; r < 4; ++r) {
        float a = m[r] * x;
        float b = m[4 + r] * y;
        float c = m[8 + r] * z;
        float d = m[12 + r] * w;
        out[r] = ((a + b) + c) + d;
    }
}

void orc_create_ray(const float* cam, uint32_t px, uint32_t py, uint32_t width,
                    uint32_t height, float* o, float* d) {
    const float* pinv = cam;
    const float* vinv = cam + 16;
    float hx = (float)width / 2.0f;
    float hy = (float)height / 2.0f;
    float nx = (float)px / hx - 1.0f;
    float ny = (float)py / hy - 1.0f;
    float P[4];
    mat_vec(pinv, nx, -ny, 1.0f, 1.0f, P);
    float dot = ((P[0] * P[0] + P[1] * P[1]) + P[2] * P[2]) + P[3] * P[3];
    float len = std::sqrt(dot);
    float q[3] = {P[0] / len, P[1] / len, P[2] / len};
    float D[4], O[4];
    mat_vec(vinv, q[0], q[1], q[2], 0.0f, D);
    mat_vec(vinv, 0.0f, 0.0f, 0.0f, 1.0f, O);
    for (int k = 0; k < 3; ++k) {
        d[k] = D[k];
        o[k] = O[k];
    }
}

void orc_create_rays(const float* cam, uint32_t width, uint32_t height, float tmin, float tmax,
                     float* rays) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)width * height; ++i) {
        float* r = rays + 8 * i;
        orc_create_ray(cam, (uint32_t)(i % width), (uint32_t)(i / width), width, height, r,
                       r + 4);
        r[3] = tmin;
        r[7] = tmax;
    }
}

struct OrcHit {
    uint32_t prim;
    float t, u, v;
};

// Watertight test, Woop et al. 2013, two-sided; the edge functions are
// recomputed in double precision when one of them is exactly zero.
static bool orc_watertight(const float* org, const float* dir, const float* v0,
                           const float* v1, const float* v2, float tmin, float tmax, float* t,
                           float* u, float* v) {
    float ad[3] = {std::fabs(dir[0]), std::fabs(dir[1]), std::fabs(dir[2])};
    int kz = 0;
    if (ad[0] > ad[1]) {
        kz = ad[0] > ad[2] ? 0 : 2;
    } else {
        kz = ad[1] > ad[2] ? 1 : 2;
    }
    int kx = (kz + 1) % 3;
    int ky = (kx + 1) % 3;
    if (dir[kz] < 0.0f) std::swap(kx, ky);
    float Sx = dir[kx] / dir[kz];
    float Sy = dir[ky] / dir[kz];
    float Sz = 1.0f / dir[kz];
    float A[3], B[3], C[3];
    for (int k = 0; k < 3; ++k) {
        A[k] = v0[k] - org[k];
        B[k] = v1[k] - org[k];
        C[k] = v2[k] - org[k];
    }
    float Ax = A[kx] - Sx * A[kz];
    float Ay = A[ky] - Sy * A[kz];
    float Bx = B[kx] - Sx * B[kz];
    float By = B[ky] - Sy * B[kz];
    float Cx = C[kx] - Sx * C[kz];
    float Cy = C[ky] - Sy * C[kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        U = (float)((double)Cx * (double)By - (double)Cy * (double)Bx);
        V = (float)((double)Ax * (double)Cy - (double)Ay * (double)Cx);
        W = (float)((double)Bx * (double)Ay - (double)By * (double)Ax);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = Sz * A[kz];
    float Bz = Sz * B[kz];
    float Cz = Sz * C[kz];
    float T = U * Az + V * Bz + W * Cz;
    float rcp = 1.0f / det;
    float tt = T * rcp;
    if (!(tt >= tmin && tt <= tmax)) return false;
    *t = tt;
    *u = V * rcp;
    *v = W * rcp;
    return true;
}

static inline void tri_verts(const float* verts, const uint32_t* idx, uint64_t t,
                             const float** a, const float** b, const float** c) {
    uint64_t i0 = idx ? idx[3 * t] : 3 * t, i1 = idx ? idx[3 * t + 1] : 3 * t + 1,
             i2 = idx ? idx[3 * t + 2] : 3 * t + 2;
    *a = verts + 4 * i0;
    *b = verts + 4 * i1;
    *c = verts + 4 * i2;
}

// Any-hit filter of non-opaque geometry: accept(ctx, prim, u, v) decides
// whether a candidate counts (null: every candidate counts).
typedef bool (*OrcAccept)(const void* ctx, uint32_t prim, float u, float v);

static inline void consider(const float* r, const float* a, const float* b, const float* c,
                            uint32_t prim, OrcHit* h, OrcAccept accept = nullptr,
                            const void* ctx = nullptr) {
    float t, u, v;
    if (orc_watertight(r, r + 4, a, b, c, r[3], h->t, &t, &u, &v)) {
        if ((t < h->t || prim < h->prim) && (!accept || accept(ctx, prim, u, v))) {
            h->t = t;
            h->u = u;
            h->v = v;
            h->prim = prim;
        }
    }
}

// rays: n x 8 floats (origin, tmin, dir, tmax); vertices: 4 floats each.
void orc_intersect_brute(const float* verts, const uint32_t* idx, uint64_t ntri,
                         const float* rays, uint64_t n, OrcHit* hits) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float* r = rays + 8 * i;
        OrcHit h{0xFFFFFFFFu, r[7], 0.f, 0.f};
        for (uint64_t t = 0; t < ntri; ++t) {
            const float *a, *b, *c;
            tri_verts(verts, idx, t, &a, &b, &c);
            consider(r, a, b, c, (uint32_t)t, &h);
        }
        hits[i] = h;
    }
}

// ---------------------------------------------------------------------------
// Oracle BVH: binary, median split along the longest centroid axis, leaves of
// at most 4 triangles, boxes tested in double precision with a relative
// widening so that no box containing a hit is culled.
struct OrcNode {
    double lo[3], hi[3];
    uint32_t left, first, count;  // count > 0: leaf
};

struct OrcBvh {
    std::vector<OrcNode> nodes;
    std::vector<uint32_t> prims;
    const float* verts;
    const uint32_t* idx;
    uint64_t ntri;
    std::vector<float> vcopy;
    std::vector<uint32_t> icopy;
};

static void orc_build_rec(OrcBvh* b, std::vector<double>& cent, uint32_t ni, uint32_t first,
                          uint32_t count) {
    OrcNode nd;
    for (int k = 0; k < 3; ++k) {
        nd.lo[k] = INFINITY;
        nd.hi[k] = -INFINITY;
    }
    double clo[3] = {INFINITY, INFINITY, INFINITY}, chi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t i = first; i < first + count; ++i) {
        uint32_t p = b->prims[i];
        const float *a, *bb, *c;
        tri_verts(b->verts, b->idx, p, &a, &bb, &c);
        for (int k = 0; k < 3; ++k) {
            nd.lo[k] = std::min(nd.lo[k], (double)std::min(a[k], std::min(bb[k], c[k])));
            nd.hi[k] = std::max(nd.hi[k], (double)std::max(a[k], std::max(bb[k], c[k])));
            clo[k] = std::min(clo[k], cent[3 * p + k]);
            chi[k] = std::max(chi[k], cent[3 * p + k]);
        }
    }
    if (count <= 4) {
        nd.left = 0;
        nd.first = first;
        nd.count = count;
        b->nodes[ni] = nd;
        return;
    }
    int axis = 0;
    double ext = chi[0] - clo[0];
    for (int k = 1; k < 3; ++k)
        if (chi[k] - clo[k] > ext) {
            ext = chi[k] - clo[k];
            axis = k;
        }
    uint32_t mid = first + count / 2;
    std::nth_element(b->prims.begin() + first, b->prims.begin() + mid,
                     b->prims.begin() + first + count, [&](uint32_t x, uint32_t y) {
                         double cx = cent[3 * x + axis], cy = cent[3 * y + axis];
                         return cx < cy || (cx == cy && x < y);
                     });
    uint32_t left = (uint32_t)b->nodes.size();
    b->nodes.resize(b->nodes.size() + 2);
    nd.left = left;
    nd.first = 0;
    nd.count = 0;
    b->nodes[ni] = nd;
    orc_build_rec(b, cent, left, first, mid - first);
    orc_build_rec(b, cent, left + 1, mid, first + count - mid);
}

void* orc_bvh_create(const float* verts, uint64_t nverts, const uint32_t* idx, uint64_t ntri) {
    OrcBvh* b = new OrcBvh();
    b->vcopy.assign(verts, verts + 4 * nverts);
    if (idx) b->icopy.assign(idx, idx + 3 * ntri);
    b->verts = b->vcopy.data();
    b->idx = idx ? b->icopy.data() : nullptr;
    b->ntri = ntri;
    std::vector<double> cent(3 * ntri);
    for (uint64_t t = 0; t < ntri; ++t) {
        const float *a, *bb, *c;
        tri_verts(b->verts, b->idx, t, &a, &bb, &c);
        for (int k = 0; k < 3; ++k) cent[3 * t + k] = ((double)a[k] + bb[k] + c[k]) / 3.0;
        b->prims.push_back((uint32_t)t);
    }
    b->nodes.resize(1);
    if (ntri == 0) {
        b->nodes.clear();
        return b;
    }
    orc_build_rec(b, cent, 0, 0, (uint32_t)ntri);
    return b;
}

void orc_bvh_destroy(void* h) { delete (OrcBvh*)h; }

static bool orc_box(const OrcNode& n, const double* o, const double* inv, double tmin,
                    double tmax) {
    double tn = tmin, tf = tmax;
    for (int k = 0; k < 3; ++k) {
        double lo = n.lo[k], hi = n.hi[k];
        double pad = 1e-7 * (std::fabs(lo) + std::fabs(hi) + 1e-30);
        lo -= pad;
        hi += pad;
        if (std::isinf(inv[k])) {
            if (o[k] < lo || o[k] > hi) return false;
            continue;
        }
        double t0 = (lo - o[k]) * inv[k], t1 = (hi - o[k]) * inv[k];
        if (t0 > t1) std::swap(t0, t1);
        tn = std::max(tn, t0);
        tf = std::min(tf, t1);
    }
    return tn <= tf * (1.0 + 1e-9) + 1e-12;
}

// Closest hit among the candidates that `accept` takes (all when null).
void orc_bvh_intersect1_filtered(void* hb, const float* r, OrcHit* out, OrcAccept accept,
                                 const void* ctx) {
    OrcBvh* b = (OrcBvh*)hb;
    {
        OrcHit h{0xFFFFFFFFu, r[7], 0.f, 0.f};
        if (!b->nodes.empty()) {
            double o[3] = {r[0], r[1], r[2]};
            double inv[3];
            for (int k = 0; k < 3; ++k) inv[k] = r[4 + k] == 0.0f ? INFINITY : 1.0 / r[4 + k];
            uint32_t stack[128];
            int sp = 0;
            stack[sp++] = 0;
            while (sp) {
                const OrcNode& nd = b->nodes[stack[--sp]];
                if (!orc_box(nd, o, inv, r[3], (double)h.t)) continue;
                if (nd.count) {
                    for (uint32_t k = nd.first; k < nd.first + nd.count; ++k) {
                        uint32_t p = b->prims[k];
                        const float *a, *bb, *c;
                        tri_verts(b->verts, b->idx, p, &a, &bb, &c);
                        consider(r, a, bb, c, p, &h, accept, ctx);
                    }
                } else {
                    stack[sp++] = nd.left;
                    stack[sp++] = nd.left + 1;
                }
            }
        }
        *out = h;
    }
}

void orc_bvh_intersect1(void* hb, const float* r, OrcHit* out) {
    orc_bvh_intersect1_filtered(hb, r, out, nullptr, nullptr);
}

void orc_bvh_intersect(void* hb, const float* rays, uint64_t n, OrcHit* hits) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)n; ++i) orc_bvh_intersect1(hb, rays + 8 * i, hits + i);
}

int orc_num_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
