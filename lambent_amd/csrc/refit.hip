// refit.hip - world-space placement of instance triangles and the bottom-up
// refit of the 8-wide BVH after instances move (lmb_scene_update_instances).
//
// The node format and the quantization rule are those of the host builder
// (bvh_build.cpp): per axis a power-of-two scale 2^e >= extent / 253 and
// >= max|coordinate| / 2^22, an origin p that is a multiple of it, and 8-bit
// planes stepped outward until fmaf(q, 2^e, p) encloses the child's exact
// box, so every refit box is conservative and closest hits stay exact.
#include <hip/hip_runtime.h>

#include "lmb_instances.h"

namespace lmb {

namespace {

constexpr int kRefitBlock = 256;

// One row of a 3x4 row-major transform, summed left to right with separate
// roundings: ((m0 x + m1 y) + m2 z) + m3, as the host computes it at creation.
__device__ __forceinline__ float xf_row(const float* m, float x, float y, float z) {
    return __fadd_rn(__fadd_rn(__fadd_rn(__fmul_rn(m[0], x), __fmul_rn(m[1], y)),
                               __fmul_rn(m[2], z)),
                     m[3]);
}

__global__ __launch_bounds__(kRefitBlock) void k_instance_place(float4* obj, LeafTri* tris,
                                                                uint32_t n, const uint32_t* first,
                                                                uint32_t num_parts,
                                                                const PartTransform* xf) {
    const uint32_t stride = gridDim.x * kRefitBlock;
    for (uint32_t t = blockIdx.x * kRefitBlock + threadIdx.x; t < n; t += stride) {
        // the part p with first[p] <= t < first[p + 1]
        uint32_t lo = 0, hi = num_parts;
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (first[mid] <= t)
                lo = mid;
            else
                hi = mid;
        }
        float m[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) m[k] = xf[lo].m[k];
        const uint32_t mirrored = xf[lo].mirrored;
        LeafTri tri = tris[t];
        const uint32_t prim = __float_as_uint(tri.v0[3]);
        float4 a = obj[3 * (uint64_t)prim];
        const float4 b = obj[3 * (uint64_t)prim + 1], c = obj[3 * (uint64_t)prim + 2];
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            tri.v0[r] = xf_row(m + 4 * r, a.x, a.y, a.z);
            tri.v1[r] = xf_row(m + 4 * r, b.x, b.y, b.z);
            tri.v2[r] = xf_row(m + 4 * r, c.x, c.y, c.z);
        }
        tris[t] = tri;
        a.w = mirrored ? 1.0f : 0.0f;
        obj[3 * (uint64_t)prim] = a;
    }
}

// Quantization exponent of one axis of a node covering [lo, hi] (the host
// builder's rule).
__device__ __forceinline__ int pick_exponent(float lo, float hi) {
    const double extent = (double)hi - (double)lo;
    double need = fmax(extent / 253.0, fmax(fabs((double)lo), fabs((double)hi)) / 4194304.0);
    need = fmax(need, 1e-30);
    int e = (int)ceil(log2(need));
    while (ldexp(1.0, e) < need) ++e;
    return min(max(e, -126), 127);
}

__global__ __launch_bounds__(kRefitBlock) void k_refit(const RefitEntry* a, uint32_t na,
                                                       const RefitEntry* b, uint32_t nb,
                                                       const uint32_t* root_pos,
                                                       const LeafTri* tris, Bvh8Node* nodes,
                                                       float4* box, uint32_t* error) {
    const uint32_t i = blockIdx.x * kRefitBlock + threadIdx.x;
    if (i >= na + nb) return;
    const RefitEntry e = i < na ? a[i] : b[i - na];
    const uint32_t dst = (e.dst & kRootOfPart) ? root_pos[e.dst & ~kRootOfPart] : e.dst;
    float lo[8][3], hi[8][3];
    float nlo[3] = {INFINITY, INFINITY, INFINITY}, nhi[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
    for (int s = 0; s < 8; ++s) {
        const uint32_t meta = e.meta[s];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            lo[s][ax] = INFINITY;
            hi[s][ax] = -INFINITY;
        }
        if (meta == kMetaEmpty) continue;
        if (meta == kMetaInternal) {
            // internal children sit at child_base in slot order
            const uint32_t c = e.child_base + __popc((uint32_t)e.imask & ((1u << s) - 1u));
            const float4 l = box[2 * (uint64_t)c], h = box[2 * (uint64_t)c + 1];
            lo[s][0] = l.x;
            lo[s][1] = l.y;
            lo[s][2] = l.z;
            hi[s][0] = h.x;
            hi[s][1] = h.y;
            hi[s][2] = h.z;
        } else {
            const uint32_t off = meta & 31u, cnt = __popc(meta >> 5);
            for (uint32_t k = 0; k < cnt; ++k) {
                const LeafTri& t = tris[e.tri_base + off + k];
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) {
                    lo[s][ax] = fminf(lo[s][ax], fminf(t.v0[ax], fminf(t.v1[ax], t.v2[ax])));
                    hi[s][ax] = fmaxf(hi[s][ax], fmaxf(t.v0[ax], fmaxf(t.v1[ax], t.v2[ax])));
                }
            }
        }
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            nlo[ax] = fminf(nlo[ax], lo[s][ax]);
            nhi[ax] = fmaxf(nhi[ax], hi[s][ax]);
        }
    }
    Bvh8Node node;
    node.imask = e.imask;
    node.child_base = e.child_base;
    node.tri_base = e.tri_base;
    bool ok = true;
    const bool empty = !(nlo[0] <= nhi[0]);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        float p = 0.0f, scale = 1.0f;
        int ex = 0;
        if (!empty) {
            ok = ok && isfinite(nlo[ax]) && isfinite(nhi[ax]);
            ex = pick_exponent(nlo[ax], nhi[ax]);
            const double sd = ldexp(1.0, ex);
            p = (float)(floor((double)nlo[ax] / sd) * sd);
            scale = (float)sd;
        }
        node.p[ax] = p;
        node.e[ax] = (uint8_t)(ex + 127);
#pragma unroll
        for (int s = 0; s < 8; ++s) {
            node.qlo[ax][s] = 0;
            node.qhi[ax][s] = 0;
            if (e.meta[s] == kMetaEmpty) continue;
            double ql = floor(((double)lo[s][ax] - (double)p) / (double)scale);
            double qh = ceil(((double)hi[s][ax] - (double)p) / (double)scale);
            ql = fmin(fmax(ql, 0.0), 255.0);
            qh = fmin(fmax(qh, 0.0), 255.0);
            while (ql > 0.0 && !(fmaf((float)ql, scale, p) <= lo[s][ax])) ql -= 1.0;
            while (qh < 255.0 && !(fmaf((float)qh, scale, p) >= hi[s][ax])) qh += 1.0;
            ok = ok && fmaf((float)ql, scale, p) <= lo[s][ax] &&
                 fmaf((float)qh, scale, p) >= hi[s][ax];
            node.qlo[ax][s] = (uint8_t)ql;
            node.qhi[ax][s] = (uint8_t)qh;
        }
    }
#pragma unroll
    for (int s = 0; s < 8; ++s) node.meta[s] = e.meta[s];
    nodes[dst] = node;
    box[2 * (uint64_t)dst] = make_float4(nlo[0], nlo[1], nlo[2], 0.0f);
    box[2 * (uint64_t)dst + 1] = make_float4(nhi[0], nhi[1], nhi[2], 0.0f);
    if (!ok) atomicOr(error, 1u);
}

}  // namespace

hipError_t launch_instance_place(float4* obj, LeafTri* tris, uint32_t num_tris,
                                 const uint32_t* first, uint32_t num_parts,
                                 const PartTransform* xf, hipStream_t stream) {
    if (num_tris == 0 || num_parts == 0) return hipSuccess;
    const uint32_t blocks = (uint32_t)((num_tris + kRefitBlock - 1) / kRefitBlock);
    const uint32_t grid = blocks < 65536u ? blocks : 65536u;
    hipLaunchKernelGGL(k_instance_place, dim3(grid), dim3(kRefitBlock), 0, stream, obj, tris,
                       num_tris, first, num_parts, xf);
    return hipGetLastError();
}

hipError_t launch_refit(const RefitEntry* a, uint32_t na, const RefitEntry* b, uint32_t nb,
                        const uint32_t* root_pos, const LeafTri* tris, Bvh8Node* nodes,
                        float4* box, uint32_t* error, hipStream_t stream) {
    const uint64_t n = (uint64_t)na + nb;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_refit, dim3((uint32_t)((n + kRefitBlock - 1) / kRefitBlock)),
                       dim3(kRefitBlock), 0, stream, a, na, b, nb, root_pos, tris, nodes, box,
                       error);
    return hipGetLastError();
}

}  // namespace lmb
