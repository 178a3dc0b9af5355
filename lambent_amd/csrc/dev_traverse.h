// dev_traverse.h - gfx950 device code for closest-hit ray queries over the
// compressed 8-wide BVH (replaces the reference's hardware ray query,
// general.wesl:1052-1078, whose semantics are: closest hit with parametric
// t in [T_MIN, T_MAX], barycentrics (u, v) weighting vertices 1 and 2,
// opaque geometry auto-committed).
//
// All code here is compiled with -ffp-contract=off; every fused multiply-add
// is an explicit fmaf, so the arithmetic is identical on any IEEE target.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lmb_layout.h"
#include "lmb_scene_dev.h"

namespace lmb {

struct Ray {
    float ox, oy, oz, tmin;
    float dx, dy, dz, tmax;
};

struct Hit {
    uint32_t prim;  // original primitive index (kMissPrim = no hit)
    uint32_t leaf;  // position of the triangle in leaf order
    float t, u, v;
};

// Conservative widening of the far slab distance (1 + 2*gamma(3), Ize 2013,
// "Robust BVH Ray Traversal").
constexpr float kRobustFar = 1.0000004f;
constexpr float kZeroDirInv = 1e30f;

__device__ __forceinline__ float comp3(float x, float y, float z, int k) {
    return k == 0 ? x : (k == 1 ? y : z);
}

// Per-ray constants of the watertight test (Woop, Benthin, Wald 2013,
// "Watertight Ray/Triangle Intersection", JCGT 2(1)).
struct Shear {
    int kx, ky, kz;
    float sx, sy, sz;
};

__device__ __forceinline__ Shear make_shear(float dx, float dy, float dz) {
    float ax = fabsf(dx), ay = fabsf(dy), az = fabsf(dz);
    Shear s;
    s.kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
    s.kx = s.kz + 1;
    if (s.kx == 3) s.kx = 0;
    s.ky = s.kx + 1;
    if (s.ky == 3) s.ky = 0;
    float dkz = comp3(dx, dy, dz, s.kz);
    if (dkz < 0.0f) {
        int t = s.kx;
        s.kx = s.ky;
        s.ky = t;
    }
    float dkx = comp3(dx, dy, dz, s.kx);
    float dky = comp3(dx, dy, dz, s.ky);
    s.sx = dkx / dkz;
    s.sy = dky / dkz;
    s.sz = 1.0f / dkz;
    return s;
}

// Two-sided watertight intersection.  Returns true and (t, u, v) when the
// hit lies in [tmin, tmax]; u, v weight vertices 1 and 2.
__device__ __forceinline__ bool intersect_triangle(const Shear& sh, float ox, float oy,
                                                   float oz, const float4 a4,
                                                   const float4 b4, const float4 c4,
                                                   float tmin, float tmax, float* t_out,
                                                   float* u_out, float* v_out) {
    float Ax0 = a4.x - ox, Ay0 = a4.y - oy, Az0 = a4.z - oz;
    float Bx0 = b4.x - ox, By0 = b4.y - oy, Bz0 = b4.z - oz;
    float Cx0 = c4.x - ox, Cy0 = c4.y - oy, Cz0 = c4.z - oz;
    float Akx = comp3(Ax0, Ay0, Az0, sh.kx), Aky = comp3(Ax0, Ay0, Az0, sh.ky),
          Akz = comp3(Ax0, Ay0, Az0, sh.kz);
    float Bkx = comp3(Bx0, By0, Bz0, sh.kx), Bky = comp3(Bx0, By0, Bz0, sh.ky),
          Bkz = comp3(Bx0, By0, Bz0, sh.kz);
    float Ckx = comp3(Cx0, Cy0, Cz0, sh.kx), Cky = comp3(Cx0, Cy0, Cz0, sh.ky),
          Ckz = comp3(Cx0, Cy0, Cz0, sh.kz);
    float Ax = Akx - sh.sx * Akz;
    float Ay = Aky - sh.sy * Akz;
    float Bx = Bkx - sh.sx * Bkz;
    float By = Bky - sh.sy * Bkz;
    float Cx = Ckx - sh.sx * Ckz;
    float Cy = Cky - sh.sy * Ckz;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        double CxBy = (double)Cx * (double)By;
        double CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        double AxCy = (double)Ax * (double)Cy;
        double AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        double BxAy = (double)Bx * (double)Ay;
        double ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f))
        return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = sh.sz * Akz;
    float Bz = sh.sz * Bkz;
    float Cz = sh.sz * Ckz;
    float T = U * Az + V * Bz + W * Cz;
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= tmin && t <= tmax)) return false;
    *t_out = t;
    *u_out = V * rcp;
    *v_out = W * rcp;
    return true;
}

__device__ __forceinline__ uint32_t byte_of(uint32_t lo4, uint32_t hi4, int s) {
    uint32_t w = s < 4 ? lo4 : hi4;
    return (w >> ((s & 3) * 8)) & 0xFFu;
}

// Traversal stack: kStackLds entries per lane in LDS (`lds` points at this
// lane's column, stride BLOCK), deeper entries in `spill`.
template <int BLOCK>
struct ShortStack {
    uint2* lds;
    uint2* spill;
    int sp;
    __device__ __forceinline__ void push(uint2 e) {
        uint2* slot = lds + (sp & (kStackLds - 1)) * BLOCK;
        if (sp >= kStackLds) spill[sp - kStackLds] = *slot;
        *slot = e;
        ++sp;
    }
    __device__ __forceinline__ uint2 pop() {
        --sp;
        uint2* slot = lds + (sp & (kStackLds - 1)) * BLOCK;
        uint2 e = *slot;
        if (sp >= kStackLds) *slot = spill[sp - kStackLds];
        return e;
    }
};

// Closest-hit traversal as a per-lane state machine: init() sets up a ray,
// each step() visits one node (box tests of its 8 children, then the
// triangles of its hit leaves) and chooses the next node, or pops the stack.
// Ties in t resolve to the lower primitive index, so the result does not
// depend on the traversal order.
struct Traversal {
    Ray r;
    float ix, iy, iz;
    uint32_t oct;
    Shear sh;
    Hit h;
    uint32_t node;
    bool active;

    __device__ __forceinline__ void init(const Ray& ray, uint32_t num_nodes) {
        r = ray;
        h.prim = kMissPrim;
        h.leaf = kMissPrim;
        h.t = r.tmax;
        h.u = 0.0f;
        h.v = 0.0f;
        // Near-zero direction components get a large positive inverse: the
        // slab test then accepts exactly the boxes whose slab contains the
        // origin.
        ix = fabsf(r.dx) < 1e-30f ? kZeroDirInv : 1.0f / r.dx;
        iy = fabsf(r.dy) < 1e-30f ? kZeroDirInv : 1.0f / r.dy;
        iz = fabsf(r.dz) < 1e-30f ? kZeroDirInv : 1.0f / r.dz;
        oct = (ix < 0.0f ? 1u : 0u) | (iy < 0.0f ? 2u : 0u) | (iz < 0.0f ? 4u : 0u);
        sh = make_shear(r.dx, r.dy, r.dz);
        node = 0;
        active = num_nodes != 0;
    }

    // The hit leaves' triangles are handed out as (*tri_base_out,
    // *tri_mask_out) for the wave's shared triangle phase (shared_triangles).
    template <int BLOCK>
    __device__ __forceinline__ void step(const Bvh8Node* __restrict__ nodes, ShortStack<BLOCK>& st,
                                         uint32_t* tri_base_out, uint32_t* tri_mask_out) {
        const uint4* np = reinterpret_cast<const uint4*>(nodes + node);
        const uint4 w0 = np[0];
        const uint4 w1 = np[1];
        const uint4 w2 = np[2];
        const uint4 w3 = np[3];
        const uint4 w4 = np[4];
        const float px = __uint_as_float(w0.x);
        const float py = __uint_as_float(w0.y);
        const float pz = __uint_as_float(w0.z);
        const float sx = __uint_as_float((w0.w & 0xFFu) << 23);
        const float sy = __uint_as_float(((w0.w >> 8) & 0xFFu) << 23);
        const float sz = __uint_as_float(((w0.w >> 16) & 0xFFu) << 23);
        const uint32_t imask = w0.w >> 24;
        const uint32_t child_base = w1.x;
        const uint32_t tri_base = w1.y;
        // near / far plane bytes per axis, chosen by the ray's sign
        const uint32_t nx0 = (oct & 1u) ? w3.z : w2.x, nx1 = (oct & 1u) ? w3.w : w2.y;
        const uint32_t fx0 = (oct & 1u) ? w2.x : w3.z, fx1 = (oct & 1u) ? w2.y : w3.w;
        const uint32_t ny0 = (oct & 2u) ? w4.x : w2.z, ny1 = (oct & 2u) ? w4.y : w2.w;
        const uint32_t fy0 = (oct & 2u) ? w2.z : w4.x, fy1 = (oct & 2u) ? w2.w : w4.y;
        const uint32_t nz0 = (oct & 4u) ? w4.z : w3.x, nz1 = (oct & 4u) ? w4.w : w3.y;
        const uint32_t fz0 = (oct & 4u) ? w3.x : w4.z, fz1 = (oct & 4u) ? w3.y : w4.w;
        // Plane distance t = q * (2^e / d) + (p - o) / d: one fma per plane.
        // With u = 2^-24, S = 2^e * fl(1/d) carries the one rounding of 1/d and
        // O = fl(fl(p - o) * fl(1/d)) three, so against the distance to the
        // exact dequantized plane the fma result errs by at most
        // u |q S| + 3.0001 u |O| + u |t| <= 4.0002 u |O| + 2 u |t|.
        // The |O| part is folded into the per-axis offsets (near planes use
        // O - 6u|O|, far planes O + 6u|O|, covering the rounding of that fma
        // too), so it never couples axes: a near-parallel axis with a huge
        // |O| widens only its own planes.  The 2u |t| parts of tn and tf are
        // covered on the far side only: if the exact slab test hits, then
        // tn - 2u|tn| <= tf + 2u|tf|, which (by the signs of tn and tf) gives
        // tn <= tf + 4.0001u|tf|, and tf_hi = fl(tf + 6u|tf|) >= tf + 4.99u|tf|;
        // likewise tn <= best_hi = fl(best + 6u|best|), once per node, and
        // tmin <= tf + 2u|tf| <= tf_hi.  Every box the exact slab test hits
        // is therefore hit here.
        const float Sx = fminf(fmaxf(sx * ix, -1e38f), 1e38f);
        const float Sy = fminf(fmaxf(sy * iy, -1e38f), 1e38f);
        const float Sz = fminf(fmaxf(sz * iz, -1e38f), 1e38f);
        const float Ox = (px - r.ox) * ix;
        const float Oy = (py - r.oy) * iy;
        const float Oz = (pz - r.oz) * iz;
        constexpr float kOffsetSlack = 6.0f * 5.9604645e-8f;
        const float Onx = fmaf(-fabsf(Ox), kOffsetSlack, Ox), Ofx = fmaf(fabsf(Ox), kOffsetSlack, Ox);
        const float Ony = fmaf(-fabsf(Oy), kOffsetSlack, Oy), Ofy = fmaf(fabsf(Oy), kOffsetSlack, Oy);
        const float Onz = fmaf(-fabsf(Oz), kOffsetSlack, Oz), Ofz = fmaf(fabsf(Oz), kOffsetSlack, Oz);
        constexpr float kChildSlack = 6.0f * 5.9604645e-8f;
        const float best_hi = fmaf(fabsf(h.t), kChildSlack, h.t);
        uint32_t node_hits = 0, tri_hits = 0;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            const uint32_t meta = byte_of(w1.z, w1.w, s);
            const float tnx = fmaf((float)byte_of(nx0, nx1, s), Sx, Onx);
            const float tfx = fmaf((float)byte_of(fx0, fx1, s), Sx, Ofx);
            const float tny = fmaf((float)byte_of(ny0, ny1, s), Sy, Ony);
            const float tfy = fmaf((float)byte_of(fy0, fy1, s), Sy, Ofy);
            const float tnz = fmaf((float)byte_of(nz0, nz1, s), Sz, Onz);
            const float tfz = fmaf((float)byte_of(fz0, fz1, s), Sz, Ofz);
            const float tn = fmaxf(fmaxf(tnx, tny), tnz);
            const float tf = fminf(fminf(tfx, tfy), tfz);
            const float tf_hi = fmaf(fabsf(tf), kChildSlack, tf);
            // Branch-free: an empty slot's meta byte is 0, so a hit on it adds
            // no internal bit and an empty triangle run (a branch per slot
            // cost about ten scalar instructions of exec-mask bookkeeping).
            static_assert(kMetaEmpty == 0, "empty slots must carry no triangles");
            const bool hit = tn <= fminf(tf_hi, best_hi) && r.tmin <= tf_hi;
            const bool internal = meta == kMetaInternal;
            node_hits |= (hit && internal) ? (1u << ((uint32_t)s ^ oct)) : 0u;
            tri_hits |= (hit && !internal) ? ((meta >> 5) << (meta & 31u)) : 0u;
        }
        *tri_base_out = tri_base;
        *tri_mask_out = tri_hits;
        uint32_t gbase, ghits, gimask;
        if (node_hits) {
            gbase = child_base;
            ghits = node_hits;
            gimask = imask;
        } else {
            if (st.sp == 0) {
                active = false;
                return;
            }
            const uint2 e = st.pop();
            gbase = e.x;
            ghits = e.y & 0xFFu;
            gimask = e.y >> 8;
        }
        descend(gbase, ghits, gimask, st);
    }

    // Goes to the nearest child of a hit set (ghits: hit children in the
    // ray's octant order, never empty; gimask: the node's internal-child
    // mask, whose children lie contiguously from gbase) and pushes the rest.
    template <int BLOCK>
    __device__ __forceinline__ void descend(uint32_t gbase, uint32_t ghits, uint32_t gimask,
                                            ShortStack<BLOCK>& st) {
        const uint32_t key = __builtin_ctz(ghits);
        ghits &= ghits - 1u;
        if (ghits) st.push(make_uint2(gbase, ghits | (gimask << 8)));
        const uint32_t slot = key ^ oct;
        node = gbase + __builtin_popcount(gimask & ((1u << slot) - 1u));
    }
};

// Scene binding of the queue traversal: every ray of a OneScene queue uses
// the same BVH (uniform pointers, kept in scalar registers); a SceneTable
// queue tags each ray with a scene index and each lane binds that scene's
// BVH when it starts the ray.
struct OneScene {
    static constexpr bool kAlphaTest = false;  // opaque: every candidate hit counts
    const Bvh8Node* nodes_;
    const LeafTri* tris_;
    uint32_t num_nodes_;
    __device__ __forceinline__ void bind(uint32_t) {}
    __device__ __forceinline__ const Bvh8Node* nodes() const { return nodes_; }
    __device__ __forceinline__ const LeafTri* tris() const { return tris_; }
    __device__ __forceinline__ uint32_t num_nodes() const { return num_nodes_; }
    // triangle array of the ray of lane `owner` (all lanes call this)
    __device__ __forceinline__ const LeafTri* tris_of(uint32_t) const { return tris_; }
};

struct SceneTable {
    static constexpr bool kAlphaTest = false;
    const SceneDev* table;
    const Bvh8Node* nodes_ = nullptr;
    const LeafTri* tris_ = nullptr;
    uint32_t num_nodes_ = 0;
    __device__ __forceinline__ void bind(uint32_t item) {
        nodes_ = table[item].nodes;
        tris_ = table[item].tris;
        num_nodes_ = table[item].num_nodes;
    }
    __device__ __forceinline__ const Bvh8Node* nodes() const { return nodes_; }
    __device__ __forceinline__ const LeafTri* tris() const { return tris_; }
    __device__ __forceinline__ uint32_t num_nodes() const { return num_nodes_; }
    // Every lane binds its own scene: the owner's triangle array comes from
    // the owner lane (all lanes call this).
    __device__ __forceinline__ const LeafTri* tris_of(uint32_t owner) const {
        const uint64_t p = reinterpret_cast<uint64_t>(tris_);
        const uint32_t lo = __shfl((uint32_t)p, (int)owner);
        const uint32_t hi = __shfl((uint32_t)(p >> 32), (int)owner);
        return reinterpret_cast<const LeafTri*>(((uint64_t)hi << 32) | lo);
    }
};

// Per-wave LDS of the shared triangle phase: 64 entries each (one per lane).
struct TriShare {
    unsigned long long* key;  // (orderable t << 32) | primitive id of the best hit
    float2* uv;
    uint32_t* leaf;
    uint32_t* busy;  // stamp of the last pass in which a helper of this lane's ray ran
};

// Order-preserving map of floats to unsigned integers (NaN excluded).
__device__ __forceinline__ uint32_t orderable(float t) {
    const uint32_t u = __float_as_uint(t);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float from_orderable(uint32_t o) {
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

// Inclusive prefix sum over the 64 lanes of a wave: four DPP row shifts
// (16-lane rows; lanes shifted in from outside the row read 0), then lane 15
// of each row added to the next row and lane 31 to the upper half (the GFX9
// row broadcasts).  Each step is one VALU instruction with no LDS round trip.
__device__ __forceinline__ uint32_t wave_inclusive_sum(uint32_t x) {
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xF, 0xF, true);   // row_shr:1
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xF, 0xF, true);   // row_shr:2
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xF, 0xF, true);   // row_shr:4
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xF, 0xF, true);   // row_shr:8
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xA, 0xF, false);  // row_bcast:15
    x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return x;
}

// Position of the j-th set bit (j from 0) of a 24-bit mask.
__device__ __forceinline__ uint32_t nth_set_bit(uint32_t m, uint32_t j) {
    uint32_t b = 0, c;
    c = __popc(m & 0xFFu);
    if (j >= c) { j -= c; b = 8; }
    c = __popc((m >> b) & 0xFFu);
    if (j >= c) { j -= c; b += 8; }
    c = __popc((m >> b) & 0xFu);
    if (j >= c) { j -= c; b += 4; }
    c = __popc((m >> b) & 0x3u);
    if (j >= c) { j -= c; b += 2; }
    c = (m >> b) & 1u;
    if (j >= c) b += 1;
    return b;
}

// Position of the j-th set bit (j from 0) of a 64-bit mask with more than j
// bits set.
__device__ __forceinline__ uint32_t nth_set_bit64(uint64_t m, uint32_t j) {
    uint32_t w = (uint32_t)m, b = 0, c = __popc(w);
    if (j >= c) { j -= c; w = (uint32_t)(m >> 32); b = 32; }
    c = __popc(w & 0xFFFFu);
    if (j >= c) { j -= c; w >>= 16; b += 16; }
    c = __popc(w & 0xFFu);
    if (j >= c) { j -= c; w >>= 8; b += 8; }
    c = __popc(w & 0xFu);
    if (j >= c) { j -= c; w >>= 4; b += 4; }
    c = __popc(w & 0x3u);
    if (j >= c) { j -= c; w >>= 2; b += 2; }
    if (j >= (w & 1u)) b += 1;
    return b;
}

// Shared triangle phase, called by all 64 lanes of a wave after each node
// step.  Lane l holds two runs of untested triangles: (pb_l, pm_l), carried
// from the pass before, and (nb_l, nm_l) from this pass's step (base + bit
// of mask).  The wave's (lane, triangle) pairs - every carried run first,
// then every new run - are spread over all lanes, 64 per round, each tested
// against its lane's ray (fetched by lane shuffles).  Rounds run only when
// full, except that carried pairs are always tested: a pass with carried
// pairs tests everything if fewer than 64 are pending, and otherwise the full
// rounds; the untested rest (fewer than 64, from this pass's runs) is carried
// to the next pass.  A lane therefore waits at most one pass for its
// triangles, and a lane whose traversal has ended keeps its ray until its
// carried pairs are tested.  Per ray the closest hit is the minimum of
// (t, primitive id) - the order the one-lane loop applies (nearer t first,
// then the lower primitive id) - taken with a 64-bit LDS atomicMin on the hit
// record of lane key_lane_l (the lane itself, or for a helper the lane whose
// ray it traverses); the lane holding the winning triangle then writes its
// (u, v) and leaf index there.  Testing later only delays the pruning
// distance of a ray, never the result.
template <class Scenes>
__device__ __forceinline__ void shared_triangles(Traversal& tv, uint32_t& pb, uint32_t& pm,
                                                 uint32_t nb, uint32_t nm, const Scenes& sc,
                                                 const TriShare& ts, uint32_t lane,
                                                 uint32_t key_lane, bool helper) {
    // counts and offsets of both runs in one word (each total is below 2^16:
    // at most 64 x 24 pairs per pass)
    const uint32_t c = (uint32_t)__popc(pm) | ((uint32_t)__popc(nm) << 16);
    const uint32_t incl = wave_inclusive_sum(c);
    const uint32_t totals = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    const uint32_t t_old = totals & 0xFFFFu, total = t_old + (totals >> 16);
    const uint32_t n_test = (t_old != 0u && total < 64u) ? total : (total & ~63u);
    const uint32_t excl = incl - c;
    const uint32_t off_new = t_old + (excl >> 16);
    if (n_test == 0u) {  // nothing carried and fewer than 64 new pairs: carry them
        pb = nb;
        pm = nm;
        return;
    }
    const uint32_t offs = (excl & 0xFFFFu) | (off_new << 16);
    const unsigned long long init = ((unsigned long long)orderable(tv.h.t) << 32) | tv.h.prim;
    ts.key[lane] = init;
    const uint32_t kxyz = (uint32_t)tv.sh.kx | ((uint32_t)tv.sh.ky << 2) | ((uint32_t)tv.sh.kz << 4) |
                          (key_lane << 6);
    for (uint32_t r0 = 0; r0 < n_test; r0 += 64) {
        const uint32_t i = r0 + lane;
        const bool carried = i < t_old;
        const uint32_t half = carried ? 0u : 16u;
        uint32_t owner = 0;
#pragma unroll
        for (uint32_t stp = 32; stp; stp >>= 1) {
            const uint32_t m = owner + stp;
            if (((__shfl(offs, (int)m) >> half) & 0xFFFFu) <= i) owner = m;
        }
        const uint32_t o_pm = __shfl(pm, (int)owner), o_nm = __shfl(nm, (int)owner);
        const uint32_t o_pb = __shfl(pb, (int)owner), o_nb = __shfl(nb, (int)owner);
        const uint32_t o_mask = carried ? o_pm : o_nm;
        const uint32_t o_base = carried ? o_pb : o_nb;
        const uint32_t o_off = (__shfl(offs, (int)owner) >> half) & 0xFFFFu;
        const float ox = __shfl(tv.r.ox, (int)owner);
        const float oy = __shfl(tv.r.oy, (int)owner);
        const float oz = __shfl(tv.r.oz, (int)owner);
        const float otmin = __shfl(tv.r.tmin, (int)owner);
        Shear sh;
        sh.sx = __shfl(tv.sh.sx, (int)owner);
        sh.sy = __shfl(tv.sh.sy, (int)owner);
        sh.sz = __shfl(tv.sh.sz, (int)owner);
        const uint32_t ok = __shfl(kxyz, (int)owner);
        const uint32_t o_key = ok >> 6;
        const LeafTri* __restrict__ tris = sc.tris_of(owner);
        sh.kx = (int)(ok & 3u);
        sh.ky = (int)((ok >> 2) & 3u);
        sh.kz = (int)((ok >> 4) & 3u);
        bool hit = false;
        unsigned long long mykey = 0;
        float t = 0.f, u = 0.f, v = 0.f;
        uint32_t leaf = 0;
        if (i < n_test) {
            leaf = o_base + nth_set_bit(o_mask, i - o_off);
            const float4* tp = reinterpret_cast<const float4*>(tris + leaf);
            const float4 a = tp[0], b = tp[1], cc = tp[2];
            const float tmax = from_orderable((uint32_t)(ts.key[o_key] >> 32));
            bool found = intersect_triangle(sh, ox, oy, oz, a, b, cc, otmin, tmax, &t, &u, &v);
            if constexpr (Scenes::kAlphaTest) found = found && sc.accept(b, u, v);
            if (found) {
                hit = true;
                mykey = ((unsigned long long)orderable(t) << 32) | __float_as_uint(a.w);
                atomicMin(ts.key + o_key, mykey);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (hit && ts.key[o_key] == mykey) {
            ts.uv[o_key] = make_float2(u, v);
            ts.leaf[o_key] = leaf;
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
        __builtin_amdgcn_wave_barrier();
    }
    // Every carried pair has been tested; of this pass's run, the pairs from
    // index n_test on are carried (the run's highest set bits).
    const uint32_t cn = (uint32_t)__popc(nm);
    pb = nb;
    if (n_test >= off_new + cn) {
        pm = 0u;
    } else if (n_test <= off_new) {
        pm = nm;
    } else {
        const uint32_t first_left = nth_set_bit(nm, n_test - off_new);
        pm = nm & ~((1u << first_left) - 1u);
    }
    const unsigned long long k = ts.key[lane];
    if (k != init) {
        tv.h.t = from_orderable((uint32_t)(k >> 32));
        tv.h.prim = (uint32_t)k;
        const float2 uv = ts.uv[lane];
        tv.h.u = uv.x;
        tv.h.v = uv.y;
        tv.h.leaf = ts.leaf[lane];
    }
    // a helper prunes with the best distance found so far for its ray
    if (helper) tv.h.t = from_orderable((uint32_t)(ts.key[key_lane] >> 32));
}

// Persistent traversal over a ray queue with wave-level refill: whenever at
// least `refill` lanes of the wave are idle, the wave takes that many rays
// from the queue with one atomic and each idle lane starts one, so lanes do
// not wait for the longest ray of the wave.  `fetch`: kFetchSegments zeroed
// counters kFetchStride words apart, one per queue segment, counting the rays
// handed out there.
//
// Once the queue is drained the wave's idle lanes help the lanes still
// traversing: each pass, the k-th idle lane takes the top stack entry (a set
// of hit children of one node) of the k-th lane with a non-empty stack and
// traverses those subtrees for the same ray, as a helper.  A helper's
// triangles update its ray's hit record (the LDS key of the owning lane), it
// prunes with that record's distance, and it can itself donate.  A ray's
// result is stored once its own lane has finished, its carried triangles are
// tested and no helper of it is still at work.  The closest hit (and the tie
// rule) does not depend on which lane visits which subtree, so the result is
// unchanged; the wave's tail - lanes waiting for the longest ray - gets
// shorter.
// Source must provide: Ray load(uint64_t slot, uint32_t* scene),
// void store(uint64_t slot, Hit).
template <int BLOCK, class Scenes, class Source>
__device__ __forceinline__ void traverse_queue_t(Scenes& sc, uint64_t count, uint32_t* fetch,
                                                 uint32_t refill, ShortStack<BLOCK>& st,
                                                 const TriShare& ts, Source& src) {
    const uint32_t lane = __lane_id();
    if (count == 0) return;
    // A short queue is spread over every wave of the grid (a few rays each)
    // instead of filling the first waves: the traversal of a short queue is
    // latency-bound, and a wave lasts as long as its longest ray.
    const uint64_t waves = (uint64_t)gridDim.x * BLOCK / 64u;
    const uint32_t quota = (uint32_t)min<uint64_t>(64u, max<uint64_t>(1u, (count + waves - 1) / waves));
    const uint64_t seg_len = (count + kFetchSegments - 1) / kFetchSegments;
    uint32_t seg = blockIdx.x % kFetchSegments, drained = 0;  // wave-uniform
    Traversal tv;
    tv.active = false;
    uint64_t slot = 0;
    uint32_t item = 0;
    uint32_t key_lane = lane;  // lane holding the hit record of this lane's ray
    bool own = false;          // holds a ray whose hit is not yet stored
    bool helper = false;       // works on subtrees of key_lane's ray
    uint32_t pb = 0, pm = 0;   // triangles carried to the next pass (base, mask)
    bool exhausted = false;    // wave-uniform
    uint32_t stamp = 0;        // wave-uniform
    ts.busy[lane] = ~0u;
    while (true) {
        if (!exhausted) {
            const uint64_t idle = __ballot(!own);
            const uint32_t n_idle = (uint32_t)__popcll(idle);
            if (n_idle >= min(refill, quota)) {
                const uint32_t take = min(n_idle, quota);
                const uint32_t leader = (uint32_t)__ffsll((unsigned long long)idle) - 1u;
                const uint64_t lo = (uint64_t)seg * seg_len;
                const uint64_t hi = min(lo + seg_len, count);
                uint32_t base = 0;
                if (lane == leader) base = atomicAdd(fetch + seg * kFetchStride, take);
                base = __shfl(base, (int)leader);
                if (lo + base + take >= hi) {  // this segment is drained: go on to the next
                    seg = (seg + 1u) % kFetchSegments;
                    if (++drained == kFetchSegments) exhausted = true;
                }
                const uint32_t rank = (uint32_t)__popcll(idle & ((1ull << lane) - 1ull));
                if (!own && rank < take) {
                    const uint64_t mine = lo + base + rank;
                    if (mine < hi) {
                        slot = mine;
                        const Ray r = src.load(slot, &item);
                        sc.bind(item);
                        tv.init(r, sc.num_nodes());
                        st.sp = 0;
                        own = tv.active;
                        if (!tv.active) src.store(slot, tv.h);
                    }
                }
            }
        } else {
            const uint64_t idle = __ballot(!own && !helper);
            const uint64_t donors = __ballot(tv.active && st.sp > 0);
            if (idle && donors) {
                const uint32_t n = (uint32_t)min(__popcll(idle), __popcll(donors));
                const uint64_t below = (1ull << lane) - 1ull;
                uint2 e = make_uint2(0u, 0u);
                if (((donors >> lane) & 1ull) && (uint32_t)__popcll(donors & below) < n) e = st.pop();
                const uint32_t idle_rank = (uint32_t)__popcll(idle & below);
                const bool takes = ((idle >> lane) & 1ull) && idle_rank < n;
                const int from = takes ? (int)nth_set_bit64(donors, idle_rank) : (int)lane;
                const uint32_t kxyz = (uint32_t)tv.sh.kx | ((uint32_t)tv.sh.ky << 2) |
                                      ((uint32_t)tv.sh.kz << 4) | (tv.oct << 6) | (key_lane << 9);
                const uint32_t ex = __shfl(e.x, from), ey = __shfl(e.y, from);
                const float ox = __shfl(tv.r.ox, from), oy = __shfl(tv.r.oy, from);
                const float oz = __shfl(tv.r.oz, from), tmin = __shfl(tv.r.tmin, from);
                const float ix = __shfl(tv.ix, from), iy = __shfl(tv.iy, from);
                const float iz = __shfl(tv.iz, from), best = __shfl(tv.h.t, from);
                const float sx = __shfl(tv.sh.sx, from), sy = __shfl(tv.sh.sy, from);
                const float sz = __shfl(tv.sh.sz, from);
                const uint32_t k = __shfl(kxyz, from);
                const uint32_t it = __shfl(item, from);
                if (takes) {
                    tv.r.ox = ox;
                    tv.r.oy = oy;
                    tv.r.oz = oz;
                    tv.r.tmin = tmin;
                    tv.ix = ix;
                    tv.iy = iy;
                    tv.iz = iz;
                    tv.h.t = best;
                    tv.sh.sx = sx;
                    tv.sh.sy = sy;
                    tv.sh.sz = sz;
                    tv.sh.kx = (int)(k & 3u);
                    tv.sh.ky = (int)((k >> 2) & 3u);
                    tv.sh.kz = (int)((k >> 4) & 3u);
                    tv.oct = (k >> 6) & 7u;
                    key_lane = k >> 9;
                    item = it;
                    sc.bind(item);
                    helper = true;
                    tv.active = true;
                    st.sp = 0;
                    tv.descend(ex, ey & 0xFFu, ey >> 8, st);
                }
            }
        }
        uint32_t tbase = 0, tmask = 0;
        if (tv.active) tv.step<BLOCK>(sc.nodes(), st, &tbase, &tmask);
        shared_triangles(tv, pb, pm, tbase, tmask, sc, ts, lane, key_lane, helper);
        const bool done = !tv.active && pm == 0u;  // traversal ended, triangles tested
        if (helper && done) helper = false;
        const bool ready = own && done;
        if (__ballot(ready)) {
            bool waits = false;
            if (__ballot(helper)) {
                // helpers still at work mark their rays' lanes
                ++stamp;
                if (helper) ts.busy[key_lane] = stamp;
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
                waits = ts.busy[lane] == stamp;
                __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
                __builtin_amdgcn_wave_barrier();
            }
            if (ready && !waits) {
                src.store(slot, tv.h);
                own = false;
            }
        }
        if (exhausted && !__any(own || helper)) break;
    }
}

// Single-scene form; Source provides Ray load(uint64_t) and store().
template <class Source>
struct UntaggedSource {
    Source& s;
    __device__ __forceinline__ Ray load(uint64_t slot, uint32_t* item) const {
        *item = 0;
        return s.load(slot);
    }
    __device__ __forceinline__ void store(uint64_t slot, const Hit& h) const { s.store(slot, h); }
};

template <int BLOCK, int REFILL, class Source>
__device__ __forceinline__ void traverse_queue(const Bvh8Node* __restrict__ nodes,
                                               const LeafTri* __restrict__ tris,
                                               uint32_t num_nodes, uint64_t count,
                                               uint32_t* fetch, ShortStack<BLOCK>& st,
                                               const TriShare& ts, Source& src) {
    OneScene sc{nodes, tris, num_nodes};
    UntaggedSource<Source> us{src};
    traverse_queue_t<BLOCK>(sc, count, fetch, (uint32_t)REFILL, st, ts, us);
}

// LDS of the shared triangle phase for a BLOCK-thread block, and this
// thread's wave slice of it.
#define LMB_TRI_SHARE(BLOCK, name)                                                   \
    __shared__ unsigned long long name##_key[BLOCK];                                 \
    __shared__ float2 name##_uv[BLOCK];                                              \
    __shared__ uint32_t name##_leaf[BLOCK];                                          \
    __shared__ uint32_t name##_busy[BLOCK];                                          \
    const TriShare name{name##_key + (threadIdx.x & ~63u), name##_uv + (threadIdx.x & ~63u), \
                        name##_leaf + (threadIdx.x & ~63u), name##_busy + (threadIdx.x & ~63u)}

}  // namespace lmb
