// pathtrace.hip - wavefront path tracer for gfx950.
//
// The reference runs one megakernel thread per pixel (rt_main,
// general.wesl:506-610) that loops over SAMPLES paths of <= BOUNCES segments
// (rt_sample, general.wesl:643-975).  Here each sample is a wavefront:
//
//   generate  per pixel: seed, primary ray (create_ray) into queue 0
//   extend    per queued ray: closest hit over the 8-wide BVH
//   shade     per queued ray: hit record (ray_hit), emission, quality hooks,
//             Fresnel / diffuse / metal / dielectric scattering, optional
//             Russian roulette; survivors are compacted into the other queue
//             with one atomic per wavefront (ballot + prefix popcount)
//   resolve   colour = max(sum / SAMPLES, 0)
//
// This file implements guiding mode LMB_GUIDING_OFF ("plain"): the BSDF
// sampling of rt_sample with every Markov chain invalid, i.e. the cosine
// lobe with probability 1 (general.wesl:812-815).  Its random-number stream
// is the reference's with the guiding draws left out.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dev_anyhit.h"
#include "dev_common.h"
#include "dev_compact.h"
#include "dev_shade.h"
#include "dev_traverse.h"
#include "lmb_batch.h"
#include "lmb_frame_dev.h"
#include "lmb_guided.h"

namespace lmb {

constexpr int kPtBlock = 256;

// Grids of the wavefront stages, in 256-thread blocks per CU: the
// grid-stride generation / shading kernels, and the persistent traversal
// (k_extend, whose 94 VGPRs allow 5 blocks of 4 waves per CU).  Measured on
// the 1M-triangle 1080p frame (5 bounces): shading 3 / 4 / 5 blocks 0.478 /
// 0.471 / 0.531 ms per frame; traversal 3 / 4 / 5 blocks 3.53 / 3.18 /
// 3.04 ms per frame (6 blocks at 80 VGPRs spill: 3.41 ms).
#ifndef LMB_FRAME_BLOCKS_PER_CU
#define LMB_FRAME_BLOCKS_PER_CU 4
#endif
#ifndef LMB_EXTEND_BLOCKS_PER_CU
#define LMB_EXTEND_BLOCKS_PER_CU 5
#endif

static uint64_t grid_threads(uint64_t npix, uint64_t blocks_per_cu) {
    uint64_t blocks = (npix + kPtBlock - 1) / kPtBlock;
    uint64_t cap = (uint64_t)device_cu_count() * blocks_per_cu;
    blocks = std::max<uint64_t>(1, std::min(blocks, cap));
    return blocks * kPtBlock;
}

uint64_t frame_threads(uint64_t npix) { return grid_threads(npix, LMB_FRAME_BLOCKS_PER_CU); }

uint64_t extend_threads(uint64_t npix) { return grid_threads(npix, LMB_EXTEND_BLOCKS_PER_CU); }

__device__ __forceinline__ uint64_t* vertex_counter(const FrameDev& f) {
    return reinterpret_cast<uint64_t*>(f.counters + kCounterVertices);
}

// general.wesl:516: own_seed = rand_u32(x * y) + seed_offset + y; later
// samples advance it with rand_u32 (general.wesl:544).  The camera rays go
// into the bounce-0 queue in camera_slot_pixel order (8 x 8 pixel blocks);
// every later sample of a trace uses the same slots.
__global__ __launch_bounds__(kPtBlock) void k_generate(CameraDev cam, FrameDev f, TraceParams p,
                                                       uint32_t sample) {
    const uint64_t gtid = (uint64_t)blockIdx.x * kPtBlock + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * kPtBlock;
    if (gtid == 0) {
        f.counters[0] = (uint32_t)f.npix;
        for (uint32_t i = 1; i <= kMaxBounces; ++i) f.counters[i] = 0;
        for (uint32_t k = 0; k < 2 * kFetchSegments; ++k) f.counters[kCounterFetch + k * kFetchStride] = 0;
        if (sample == 0) *vertex_counter(f) = 0;
    }
    for (uint64_t slot = gtid; slot < f.npix; slot += stride) {
        uint32_t x, y;
        camera_slot_pixel(slot, f.width, f.height, x, y);
        const uint64_t pix = (uint64_t)y * f.width + x;
        uint32_t seed;
        if (sample == 0) {
            seed = rand_u32(x * y) + p.seed_offset + y;
            f.pix_sum[pix] = make_float4(0.f, 0.f, 0.f, 0.f);
        } else {
            seed = rand_u32(f.own_seed[pix]);
        }
        f.own_seed[pix] = seed;
        float o[3], d[3];
        create_ray(cam, x, y, f.width, f.height, o, d);
        f.queue[0][2 * slot] = make_float4(o[0], o[1], o[2], __uint_as_float((uint32_t)pix));
        f.queue[0][2 * slot + 1] = make_float4(d[0], d[1], d[2], 0.f);
        f.ps_seed[pix] = seed;
        // the cached tracer's shading (nrc_trace.hip) reads these back
        f.ps_through[pix] = make_float4(1.f, 1.f, 1.f, -1.f);
        f.ps_accum[pix] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// Ray source of the wavefront queue for traverse_queue().
struct QueueSource {
    const float4* __restrict__ q;
    uint4* __restrict__ hits;
    float tmin, tmax;
    __device__ __forceinline__ Ray load(uint64_t slot) const {
        const float4 o = q[2 * slot];
        const float4 d = q[2 * slot + 1];
        Ray r;
        r.ox = o.x;
        r.oy = o.y;
        r.oz = o.z;
        r.tmin = tmin;
        r.dx = d.x;
        r.dy = d.y;
        r.dz = d.z;
        r.tmax = tmax;
        return r;
    }
    __device__ __forceinline__ void store(uint64_t slot, const Hit& h) const {
        hits[slot] = make_uint4(h.leaf, __float_as_uint(h.t), __float_as_uint(h.u),
                                __float_as_uint(h.v));
    }
};

#ifndef LMB_REFILL
#define LMB_REFILL 16
#endif
constexpr int kRefill = LMB_REFILL;  // idle lanes that make a wave fetch more rays

// Blocks per CU the traversal kernel is compiled for (a register cap): at 5
// blocks of 4 waves each SIMD holds 5 waves, which needs at most 96 VGPRs.
#ifndef LMB_EXTEND_WAVES
#define LMB_EXTEND_WAVES 5
#endif

// kAlpha: the scene's geometry is not opaque (any-hit alpha test).
template <bool kAlpha>
__global__ __launch_bounds__(kPtBlock, LMB_EXTEND_WAVES) void k_extend(SceneDev s, FrameDev f, TraceParams p,
                                                     uint32_t bounce) {
    __shared__ uint2 lds_stack[kStackLds * kPtBlock];
    const uint64_t gtid = (uint64_t)blockIdx.x * kPtBlock + threadIdx.x;
    const uint32_t count = f.counters[bounce];
    if (gtid == 0) {
        atomicAdd((unsigned long long*)vertex_counter(f), (unsigned long long)count);
        atomicAdd((unsigned long long*)(f.counters + kCounterVerticesTotal),
                  (unsigned long long)count);
        // The next bounce's fetch set; its last user (bounce - 1) has finished.
        uint32_t* next = fetch_set(f.counters, bounce + 1);
        for (uint32_t k = 0; k < kFetchSegments; ++k) next[k * kFetchStride] = 0;
    }
    ShortStack<kPtBlock> st;
    st.lds = lds_stack + threadIdx.x;
    st.spill = f.spill + gtid * (kStackMax - kStackLds);
    LMB_TRI_SHARE(kPtBlock, ts);
    QueueSource src{f.queue[bounce & 1], f.hits, p.tmin, p.tmax};
    if constexpr (kAlpha) {
        AlphaScene sc = alpha_scene(s, p.quality <= kQualitySpectral);
        UntaggedSource<QueueSource> us{src};
        traverse_queue_t<kPtBlock>(sc, count, fetch_set(f.counters, bounce), (uint32_t)kRefill, st,
                                   ts, us);
    } else {
        traverse_queue<kPtBlock, kRefill>(s.nodes, s.tris, s.num_nodes, count,
                                          fetch_set(f.counters, bounce), st, ts, src);
    }
}

static void launch_extend_kernel(const SceneDev& s, const FrameDev& f, const TraceParams& p,
                                 uint32_t bounce, dim3 grid, hipStream_t stream) {
    if (s.alpha_test)
        hipLaunchKernelGGL(k_extend<true>, grid, dim3(kPtBlock), 0, stream, s, f, p, bounce);
    else
        hipLaunchKernelGGL(k_extend<false>, grid, dim3(kPtBlock), 0, stream, s, f, p, bounce);
}

// kTable: the queue's rays carry a scene index (direction .w) into `table`
// (the batched tracer); otherwise every ray belongs to `s1`.
// kGroup: f.group samples traced together (path id pixel * group + j).  At
// bounce 0 queue slot w is sample j = w % group of camera slot w / group
// (the camera queue and its hits are shared by the group's samples); a path
// that ends stores its sample colour in pix_sum[path id].
template <uint32_t Q, bool kTable, bool kGroup = false>
__global__ __launch_bounds__(kPtBlock) void k_shade_plain_t(SceneDev s1, const SceneDev* table,
                                                            FrameDev f, TraceParams p,
                                                            uint32_t bounce) {
    const uint64_t gtid = (uint64_t)blockIdx.x * kPtBlock + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * kPtBlock;
    const uint32_t group = kGroup ? f.group : 1u;
    const bool group_camera = kGroup && bounce == 0;
    const uint64_t count = group_camera ? f.npix * group : (uint64_t)f.counters[bounce];
    const uint64_t count_r = ((uint64_t)count + kPtBlock - 1) / kPtBlock * kPtBlock;
    __shared__ ReserveLds<kPtBlock, 1> rl;
    const float4* __restrict__ qin = f.queue[bounce & 1];
    float4* __restrict__ qout = f.queue[(bounce + 1) & 1];
    for (uint64_t slot = gtid; slot < count_r; slot += stride) {
        bool cont = false;
        float4 out_o = make_float4(0.f, 0.f, 0.f, 0.f), out_d = out_o;
        if (slot < count) {
            const uint64_t qs = group_camera ? slot / group : slot;
            const uint32_t j = group_camera ? (uint32_t)(slot - qs * group) : 0u;
            const float4 o4 = qin[2 * qs];
            const float4 d4 = qin[2 * qs + 1];
            const uint32_t pix = __float_as_uint(o4.w);  // path id (camera queue: pixel)
            const uint32_t pid = group_camera ? pix * group + j : pix;
            const uint4 h = f.hits[qs];
            uint32_t seed = f.ps_seed[pid];
            // every path starts here: throughput 1, nothing gathered yet
            const float4 th = bounce == 0 ? make_float4(1.f, 1.f, 1.f, -1.f) : f.ps_through[pid];
            const float4 ac = bounce == 0 ? make_float4(0.f, 0.f, 0.f, 0.f) : f.ps_accum[pid];
            V3 color = v3(th.x, th.y, th.z);
            float picked = th.w;
            V3 scol = v3(ac.x, ac.y, ac.z);
            const V3 dir = v3(d4.x, d4.y, d4.z);
            const V3 org = v3(o4.x, o4.y, o4.z);
            const uint32_t item = kTable ? __float_as_uint(d4.w) : 0u;
            const SceneDev& s = kTable ? table[item] : s1;
            if (h.x != kMissPrim) {
                const float t = __uint_as_float(h.y);
                const Isect it = make_isect(s, h.x, t, __uint_as_float(h.z),
                                            __uint_as_float(h.w), dir);
                const V3 new_pos = fma3(dir, v3s(t), org);
                if (bounce == 0 && j == 0) {  // every sample's camera hit is the same
                    f.out_normal[pix] = make_float4(it.normal.x, it.normal.y, it.normal.z, 1.f);
                    f.out_albedo[pix] = make_float4(it.color.x, it.color.y, it.color.z, 1.f);
                }
                // general.wesl:714-901
                scol = fma3(color, it.emission, scol);
                const V3 ndir = surface_scatter<Q>(it, dir, seed, color, picked);
                bool alive = bounce + 1 < p.bounces;
                if (alive && bounce >= p.rr_start) {
                    const float q = fminf(fmaxf(fmaxf(color.x, color.y), color.z), 1.0f);
                    const float rnd = rand_f32(seed);
                    seed += 1u;
                    if (rnd >= q)
                        alive = false;
                    else
                        color = color / q;
                }
                if (alive) {
                    cont = true;
                    out_o = make_float4(new_pos.x, new_pos.y, new_pos.z, __uint_as_float(pid));
                    out_d = make_float4(ndir.x, ndir.y, ndir.z, __uint_as_float(item));
                    f.ps_seed[pid] = seed;
                    f.ps_through[pid] = make_float4(color.x, color.y, color.z, picked);
                    f.ps_accum[pid] = make_float4(scol.x, scol.y, scol.z, 0.f);
                } else if (kGroup) {
                    f.pix_sum[pid] = make_float4(scol.x, scol.y, scol.z, 0.f);
                } else {
                    float4 ps = f.pix_sum[pid];
                    f.pix_sum[pid] = make_float4(ps.x + scol.x, ps.y + scol.y, ps.z + scol.z, 0.f);
                }
            } else {
                if (bounce == 0 && j == 0) {
                    f.out_normal[pix] = make_float4(0.f, 0.f, 0.f, 1.f);
                    f.out_albedo[pix] = make_float4(0.f, 0.f, 0.f, 1.f);
                }
                // general.wesl:915-938
                const V3 bg = background(s, dir);
                scol = fma3(color, bg, scol);
                if (kGroup) {
                    f.pix_sum[pid] = make_float4(scol.x, scol.y, scol.z, 0.f);
                } else {
                    float4 ps = f.pix_sum[pid];
                    f.pix_sum[pid] = make_float4(ps.x + scol.x, ps.y + scol.y, ps.z + scol.z, 0.f);
                }
            }
        }
        // Compact the survivors into the next queue: one atomic per block.
        uint32_t idx[1];
        block_reserve<kPtBlock, 1>({cont}, {f.counters + bounce + 1}, rl, idx);
        if (cont) {
            qout[2 * (uint64_t)idx[0]] = out_o;
            qout[2 * (uint64_t)idx[0] + 1] = out_d;
        }
    }
}

__global__ __launch_bounds__(kPtBlock) void k_resolve(FrameDev f, TraceParams p) {
    const uint64_t gtid = (uint64_t)blockIdx.x * kPtBlock + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * kPtBlock;
    const float ns = (float)p.samples;
    for (uint64_t pix = gtid; pix < f.npix; pix += stride) {
        const float4 s4 = f.pix_sum[pix];
        f.out_color[pix] = make_float4(fmaxf(s4.x / ns, 0.f), fmaxf(s4.y / ns, 0.f),
                                       fmaxf(s4.z / ns, 0.f), 1.f);
    }
}

FrameDev primary_view(const FrameDev& f, const TraceParams& p) {
    FrameDev f0 = f;
    if (p.samples > 1 && f.hits_primary) f0.hits = f.hits_primary;
    return f0;
}

hipError_t launch_extend(const SceneDev& s, const FrameDev& f, const TraceParams& p,
                         uint32_t bounce, hipStream_t stream) {
    const dim3 grid((uint32_t)(extend_threads(grid_paths(f)) / kPtBlock));
    launch_extend_kernel(s, f, p, bounce, grid, stream);
    return hipGetLastError();
}

hipError_t launch_generate(const CameraDev& cam, const FrameDev& f, const TraceParams& p,
                           uint32_t sample, hipStream_t stream) {
    const dim3 grid((uint32_t)(frame_threads(f.npix) / kPtBlock)), block(kPtBlock);
    hipLaunchKernelGGL(k_generate, grid, block, 0, stream, cam, f, p, sample);
    return hipGetLastError();
}

hipError_t launch_resolve(const FrameDev& f, const TraceParams& p, hipStream_t stream) {
    const dim3 grid((uint32_t)(frame_threads(f.npix) / kPtBlock)), block(kPtBlock);
    hipLaunchKernelGGL(k_resolve, grid, block, 0, stream, f, p);
    return hipGetLastError();
}

template <uint32_t Q>
static hipError_t launch_plain(const SceneDev& s, const CameraDev& cam, const FrameDev& f,
                               const TraceParams& p, hipStream_t stream, KernelTimer* tm) {
    const dim3 grid((uint32_t)(frame_threads(f.npix) / kPtBlock)), block(kPtBlock);
    const dim3 xgrid((uint32_t)(extend_threads(f.npix) / kPtBlock));
    const FrameDev f0 = primary_view(f, p);
    for (uint32_t sample = 0; sample < p.samples; ++sample) {
        hipLaunchKernelGGL(k_generate, grid, block, 0, stream, cam, f, p, sample);
        for (uint32_t b = 0; b < p.bounces; ++b) {
            const FrameDev& fb = b == 0 ? f0 : f;
            if (!(b == 0 && sample > 0 && f0.hits != f.hits)) {
                if (tm) tm->begin(kKernelExtend, stream);
                launch_extend_kernel(s, fb, p, b, xgrid, stream);
                if (tm) tm->end(kKernelExtend, stream);
            }
            if (tm) tm->begin(kKernelShade, stream);
            hipLaunchKernelGGL((k_shade_plain_t<Q, false>), grid, block, 0, stream, s,
                               (const SceneDev*)nullptr, fb, p, b);
            if (tm) tm->end(kKernelShade, stream);
        }
    }
    hipLaunchKernelGGL(k_resolve, grid, block, 0, stream, f, p);
    return hipGetLastError();
}

// ------------------------------------------------------- grouped samples
// Samples s0 .. s0 + group - 1 of every pixel start together.  A sample's
// seeds depend only on the previous sample's start seed (general.wesl:516,
// 544), so they are drawn here for the whole group, one thread per path:
// path pixel * group + j gets the start seed of sample s0 + j (own_seed is
// advanced past the group by k_group_sum).  The first group also writes
// the camera rays (traced once for all samples: no jitter) and clears
// pix_sum.
__device__ __forceinline__ uint32_t first_start_seed(const FrameDev& f, const TraceParams& p,
                                                     uint32_t s0, uint64_t pix) {
    const uint32_t x = (uint32_t)(pix % f.width), y = (uint32_t)(pix / f.width);
    return s0 == 0 ? rand_u32(x * y) + p.seed_offset + y : rand_u32(f.own_seed[pix]);
}

__global__ __launch_bounds__(kPtBlock) void k_generate_group(CameraDev cam, FrameDev f,
                                                             FrameDev fg, TraceParams p,
                                                             uint32_t s0) {
    const uint64_t gtid = (uint64_t)blockIdx.x * kPtBlock + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * kPtBlock;
    const uint32_t group = fg.group;
    if (gtid == 0) {
        f.counters[0] = (uint32_t)f.npix;
        for (uint32_t i = 1; i <= kMaxBounces; ++i) f.counters[i] = 0;
        for (uint32_t k = 0; k < 2 * kFetchSegments; ++k) f.counters[kCounterFetch + k * kFetchStride] = 0;
        if (s0 == 0) *vertex_counter(f) = 0;
    }
    const uint64_t paths = f.npix * group;
    for (uint64_t pid = gtid; pid < paths; pid += stride) {
        const uint64_t pix = pid / group;
        const uint32_t j = (uint32_t)(pid - pix * group);
        uint32_t seed = first_start_seed(f, p, s0, pix);
        for (uint32_t k = 0; k < j; ++k) seed = rand_u32(seed);
        fg.ps_seed[pid] = seed;  // throughput and gathered colour start in the shading
    }
    if (s0 != 0) return;
    for (uint64_t slot = gtid; slot < f.npix; slot += stride) {
        uint32_t x, y;
        camera_slot_pixel(slot, f.width, f.height, x, y);
        const uint64_t pix = (uint64_t)y * f.width + x;
        f.pix_sum[pix] = make_float4(0.f, 0.f, 0.f, 0.f);
        float o[3], d[3];
        create_ray(cam, x, y, f.width, f.height, o, d);
        f.queue[0][2 * slot] = make_float4(o[0], o[1], o[2], __uint_as_float((uint32_t)pix));
        f.queue[0][2 * slot + 1] = make_float4(d[0], d[1], d[2], 0.f);
    }
}

// pix_sum += the group's sample colours, in sample order (as the samples
// add them one after another when traced one at a time), and own_seed =
// the start seed of the group's last sample.
__global__ __launch_bounds__(kPtBlock) void k_group_sum(FrameDev f, const float4* __restrict__ colour,
                                                        TraceParams p, uint32_t s0,
                                                        uint32_t group) {
    const uint64_t gtid = (uint64_t)blockIdx.x * kPtBlock + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * kPtBlock;
    for (uint64_t pix = gtid; pix < f.npix; pix += stride) {
        float4 ps = f.pix_sum[pix];
        for (uint32_t j = 0; j < group; ++j) {
            const float4 c = colour[pix * group + j];
            ps = make_float4(ps.x + c.x, ps.y + c.y, ps.z + c.z, 0.f);
        }
        f.pix_sum[pix] = ps;
        uint32_t seed = first_start_seed(f, p, s0, pix);
        for (uint32_t j = 1; j < group; ++j) seed = rand_u32(seed);
        f.own_seed[pix] = seed;
    }
}

template <uint32_t Q>
static hipError_t launch_plain_grouped(const SceneDev& s, const CameraDev& cam, const FrameDev& f,
                                       const FrameDev& fg, uint32_t group, const TraceParams& p,
                                       hipStream_t stream, KernelTimer* tm) {
    const dim3 block(kPtBlock);
    const dim3 pgrid((uint32_t)(frame_threads(f.npix) / kPtBlock));
    const FrameDev f0 = primary_view(f, p);  // the camera rays' hits: hits_primary
    for (uint32_t s0 = 0; s0 < p.samples; s0 += group) {
        FrameDev g = fg;
        g.group = std::min(group, p.samples - s0);
        const uint64_t paths = f.npix * g.group;
        g.grid_paths = paths;
        const dim3 grid((uint32_t)(frame_threads(paths) / kPtBlock));
        const dim3 xgrid((uint32_t)(extend_threads(paths) / kPtBlock));
        hipLaunchKernelGGL(k_generate_group, grid, block, 0, stream, cam, f, g, p, s0);
        if (s0 == 0) {
            if (tm) tm->begin(kKernelExtend, stream);
            launch_extend_kernel(s, f0, p, 0, dim3((uint32_t)(extend_threads(f.npix) / kPtBlock)),
                                 stream);
            if (tm) tm->end(kKernelExtend, stream);
        }
        // bounce 0 reads the camera queue and hits; later bounces the group's
        FrameDev g0 = g;
        g0.queue[0] = f0.queue[0];
        g0.hits = f0.hits;
        g0.out_normal = f.out_normal;
        g0.out_albedo = f.out_albedo;
        for (uint32_t b = 0; b < p.bounces; ++b) {
            if (b > 0) {
                if (tm) tm->begin(kKernelExtend, stream);
                launch_extend_kernel(s, g, p, b, xgrid, stream);
                if (tm) tm->end(kKernelExtend, stream);
            }
            if (tm) tm->begin(kKernelShade, stream);
            hipLaunchKernelGGL((k_shade_plain_t<Q, false, true>), grid, block, 0, stream, s,
                               (const SceneDev*)nullptr, b == 0 ? g0 : g, p, b);
            if (tm) tm->end(kKernelShade, stream);
        }
        hipLaunchKernelGGL(k_group_sum, pgrid, block, 0, stream, f, (const float4*)g.pix_sum, p,
                           s0, g.group);
    }
    hipLaunchKernelGGL(k_resolve, pgrid, block, 0, stream, f, p);
    return hipGetLastError();
}

hipError_t launch_trace_grouped(const SceneDev& s, const CameraDev& cam, const FrameDev& f,
                                const FrameDev& fg, uint32_t group, const TraceParams& p,
                                hipStream_t stream, KernelTimer* timer) {
    if (f.npix == 0) return hipSuccess;
    switch (p.quality) {
        case kQualityLow:
            return launch_plain_grouped<kQualityLow>(s, cam, f, fg, group, p, stream, timer);
        case kQualityHigh:
            return launch_plain_grouped<kQualityHigh>(s, cam, f, fg, group, p, stream, timer);
        case kQualitySpectral:
            return launch_plain_grouped<kQualitySpectral>(s, cam, f, fg, group, p, stream, timer);
        default:
            return launch_plain_grouped<kQualityMedium>(s, cam, f, fg, group, p, stream, timer);
    }
}

// ---------------------------------------------------------------- batches
// Zeroes the queue and fetch counters (and, at the start of a frame, the
// frame's vertex counter) with a kernel, so a captured frame needs no memset.
__global__ __launch_bounds__(kPtBlock) void k_batch_begin(FrameDev f, uint32_t frame_start) {
    for (uint32_t i = threadIdx.x; i < kNumCounters; i += kPtBlock) {
        const bool vertices = i >= kCounterVertices && i < kCounterFetch;
        if (!vertices) f.counters[i] = 0;
    }
    if (frame_start && threadIdx.x == 0) *vertex_counter(f) = 0;
}

__device__ __forceinline__ uint32_t batch_item_of(const BatchDev& b, uint64_t tile) {
    uint32_t lo = 0, hi = b.num_items - 1;
    while (lo < hi) {  // last item whose first tile is <= tile
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (b.items[mid].tile_base <= tile)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Primary rays of the tiles [t0, t1): consecutive threads walk a
// tile_w x tile_h pixel tile, so each wavefront of the queue holds a compact
// screen tile of one image.  Seeds as general.wesl:516 with the image's own
// x, y; rays carry their image index in direction .w.  With `camera_only`
// (sample groups) only the rays are written and pix_sum cleared: the
// groups' seeds come from k_batch_group_start.
__global__ __launch_bounds__(kPtBlock) void k_batch_generate(BatchDev b, FrameDev f, TraceParams p,
                                                             uint32_t sample, uint64_t t0,
                                                             uint64_t t1, uint32_t camera_only) {
    const uint64_t gtid = (uint64_t)blockIdx.x * kPtBlock + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * kPtBlock;
    const uint32_t tile_px = b.tile_w * b.tile_h;
    const uint64_t n = (t1 - t0) * tile_px;
    const uint64_t n_r = block_uniform_count<kPtBlock>((uint32_t)n);
    const uint32_t seed_offset = *b.seed;
    __shared__ ReserveLds<kPtBlock, 1> rl;
    for (uint64_t g = gtid; g < n_r; g += stride) {
        bool valid = g < n;
        uint32_t item = 0, x = 0, y = 0;
        uint64_t pid = 0;
        if (valid) {
            // a pass holds under 2^32 pixels and an image under 2^32 tiles:
            // 32-bit division suffices
            const uint32_t g32 = (uint32_t)g;
            const uint64_t tile = t0 + g32 / tile_px;
            const uint32_t within = g32 % tile_px;
            item = batch_item_of(b, tile);
            const BatchItemDev& it = b.items[item];
            const uint32_t lt = (uint32_t)(tile - it.tile_base);
            x = (lt % it.tiles_x) * b.tile_w + within % b.tile_w;
            y = (lt / it.tiles_x) * b.tile_h + within / b.tile_w;
            valid = x < it.width && y < it.height;
            pid = it.pix_base + (uint64_t)y * it.width + x;
        }
        float4 o4 = make_float4(0.f, 0.f, 0.f, 0.f), d4 = o4;
        if (valid) {
            const BatchItemDev& it = b.items[item];
            if (sample == 0) f.pix_sum[pid] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (!camera_only) {
                const uint32_t seed = sample == 0 ? rand_u32(x * y) + seed_offset + y
                                                  : rand_u32(f.own_seed[pid]);
                f.own_seed[pid] = seed;
                f.ps_seed[pid] = seed;  // throughput and colour start in the shading
            }
            float o[3], d[3];
            create_ray(it.cam, x, y, it.width, it.height, o, d);
            o4 = make_float4(o[0], o[1], o[2], __uint_as_float((uint32_t)pid));
            d4 = make_float4(d[0], d[1], d[2], __uint_as_float(item));
        }
        uint32_t idx[1];
        block_reserve<kPtBlock, 1>({valid}, {f.counters}, rl, idx);
        if (valid) {
            f.queue[0][2 * (uint64_t)idx[0]] = o4;
            f.queue[0][2 * (uint64_t)idx[0] + 1] = d4;
        }
    }
}

__device__ __forceinline__ uint32_t batch_item_of_pixel(const BatchDev& b, uint64_t pix) {
    uint32_t lo = 0, hi = b.num_items - 1;
    while (lo < hi) {  // last item whose first pixel is <= pix
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (b.items[mid].pix_base <= pix)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// Start seeds of samples s0 .. s0 + group - 1 of every pixel of the batch,
// one thread per pixel (general.wesl:516, 544, as k_batch_generate draws
// them one sample at a time): path pixel * group + j starts sample s0 + j,
// and own_seed keeps the start seed of the group's last sample.
__global__ __launch_bounds__(kPtBlock) void k_batch_group_start(BatchDev b, FrameDev f,
                                                                FrameDev fg, uint32_t s0) {
    const uint64_t gtid = (uint64_t)blockIdx.x * kPtBlock + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * kPtBlock;
    const uint32_t group = fg.group;
    const uint32_t seed_offset = *b.seed;
    for (uint64_t pix = gtid; pix < f.npix; pix += stride) {
        uint32_t seed;
        if (s0 == 0) {
            const BatchItemDev& it = b.items[batch_item_of_pixel(b, pix)];
            const uint32_t local = (uint32_t)(pix - it.pix_base);
            const uint32_t x = local % it.width, y = local / it.width;
            seed = rand_u32(x * y) + seed_offset + y;
        } else {
            seed = rand_u32(f.own_seed[pix]);
        }
        for (uint32_t j = 0; j < group; ++j) {
            if (j > 0) seed = rand_u32(seed);
            fg.ps_seed[pix * group + j] = seed;  // the rest of the path state starts in the shading
        }
        f.own_seed[pix] = seed;
    }
}

// pix_sum += the group's sample colours, in sample order.
__global__ __launch_bounds__(kPtBlock) void k_batch_group_sum(FrameDev f,
                                                              const float4* __restrict__ colour,
                                                              uint32_t group) {
    const uint64_t gtid = (uint64_t)blockIdx.x * kPtBlock + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * kPtBlock;
    for (uint64_t pix = gtid; pix < f.npix; pix += stride) {
        float4 ps = f.pix_sum[pix];
        for (uint32_t j = 0; j < group; ++j) {
            const float4 c = colour[pix * group + j];
            ps = make_float4(ps.x + c.x, ps.y + c.y, ps.z + c.z, 0.f);
        }
        f.pix_sum[pix] = ps;
    }
}

// Camera-ray hits of a batch kept per pixel (the bounce-0 queue order
// differs between samples): scattered after the first sample's traversal,
// gathered into the queue's hit slots by later samples.
__global__ __launch_bounds__(kPtBlock) void k_hits_keep(FrameDev f, uint32_t gather) {
    const uint64_t stride = (uint64_t)gridDim.x * kPtBlock;
    const uint32_t count = f.counters[0];
    for (uint64_t slot = (uint64_t)blockIdx.x * kPtBlock + threadIdx.x; slot < count;
         slot += stride) {
        const uint32_t pid = __float_as_uint(f.queue[0][2 * slot].w);
        if (gather)
            f.hits[slot] = f.hits_primary[pid];
        else
            f.hits_primary[pid] = f.hits[slot];
    }
}

// Queue rays tagged with their image index (direction .w).
struct TaggedQueueSource {
    const float4* __restrict__ q;
    uint4* __restrict__ hits;
    float tmin, tmax;
    __device__ __forceinline__ Ray load(uint64_t slot, uint32_t* item) const {
        const float4 o = q[2 * slot];
        const float4 d = q[2 * slot + 1];
        *item = __float_as_uint(d.w);
        Ray r;
        r.ox = o.x;
        r.oy = o.y;
        r.oz = o.z;
        r.tmin = tmin;
        r.dx = d.x;
        r.dy = d.y;
        r.dz = d.z;
        r.tmax = tmax;
        return r;
    }
    __device__ __forceinline__ void store(uint64_t slot, const Hit& h) const {
        hits[slot] = make_uint4(h.leaf, __float_as_uint(h.t), __float_as_uint(h.u),
                                __float_as_uint(h.v));
    }
};

// The batch traversal binds a scene per lane: 99 VGPRs uncapped (4 waves per
// SIMD); the 96-VGPR cap of k_extend would spill.
__global__ __launch_bounds__(kPtBlock) void k_batch_extend(const SceneDev* table, FrameDev f,
                                                           TraceParams p, uint32_t bounce,
                                                           uint32_t refill) {
    __shared__ uint2 lds_stack[kStackLds * kPtBlock];
    const uint64_t gtid = (uint64_t)blockIdx.x * kPtBlock + threadIdx.x;
    const uint32_t count = f.counters[bounce];
    if (gtid == 0) {
        atomicAdd((unsigned long long*)vertex_counter(f), (unsigned long long)count);
        atomicAdd((unsigned long long*)(f.counters + kCounterVerticesTotal),
                  (unsigned long long)count);
        // The next bounce's fetch set; its last user (bounce - 1) has finished.
        uint32_t* next = fetch_set(f.counters, bounce + 1);
        for (uint32_t k = 0; k < kFetchSegments; ++k) next[k * kFetchStride] = 0;
    }
    ShortStack<kPtBlock> st;
    st.lds = lds_stack + threadIdx.x;
    st.spill = f.spill + gtid * (kStackMax - kStackLds);
    LMB_TRI_SHARE(kPtBlock, ts);
    SceneTable sc{table};
    TaggedQueueSource src{f.queue[bounce & 1], f.hits, p.tmin, p.tmax};
    traverse_queue_t<kPtBlock>(sc, count, fetch_set(f.counters, bounce), refill, st, ts, src);
}

template <uint32_t Q>
static hipError_t launch_batch_grouped(const BatchDev& b, const FrameDev& f, const FrameDev& fg,
                                       const TraceParams& p, const BatchLaunch& l,
                                       hipStream_t stream) {
    const dim3 grid(l.blocks), block(kPtBlock);
    const SceneDev none{};
    for (uint32_t s0 = 0; s0 < p.samples; s0 += fg.group) {
        FrameDev g = fg;
        g.group = std::min(fg.group, p.samples - s0);
        hipLaunchKernelGGL(k_batch_begin, dim3(1), block, 0, stream, f, (uint32_t)(s0 == 0));
        if (s0 == 0) {
            // camera rays and their hits, shared by every sample (no jitter)
            hipLaunchKernelGGL(k_batch_generate, grid, block, 0, stream, b, f, p, 0u, (uint64_t)0,
                               b.tiles_total, 1u);
            hipLaunchKernelGGL(k_batch_extend, grid, block, 0, stream, b.scenes, f, p, 0u, l.refill);
        }
        hipLaunchKernelGGL(k_batch_group_start, grid, block, 0, stream, b, f, g, s0);
        FrameDev g0 = g;  // bounce 0 reads the camera queue and hits
        g0.queue[0] = f.queue[0];
        g0.hits = f.hits;
        for (uint32_t bn = 0; bn < p.bounces; ++bn) {
            if (bn > 0)
                hipLaunchKernelGGL(k_batch_extend, grid, block, 0, stream, b.scenes, g, p, bn,
                                   l.refill);
            hipLaunchKernelGGL((k_shade_plain_t<Q, true, true>), grid, block, 0, stream, none,
                               b.scenes, bn == 0 ? g0 : g, p, bn);
        }
        hipLaunchKernelGGL(k_batch_group_sum, grid, block, 0, stream, f, (const float4*)g.pix_sum,
                           g.group);
    }
    hipLaunchKernelGGL(k_resolve, grid, block, 0, stream, f, p);
    return hipGetLastError();
}

template <uint32_t Q>
static hipError_t launch_batch_q(const BatchDev& b, const FrameDev& f, const FrameDev* fg,
                                 const TraceParams& p, const BatchLaunch& l, hipStream_t stream) {
    if (fg && fg->group > 1 && p.samples > 1 && l.pass_tiles >= b.tiles_total)
        return launch_batch_grouped<Q>(b, f, *fg, p, l, stream);
    const dim3 grid(l.blocks), block(kPtBlock);
    const SceneDev none{};
    for (uint64_t t0 = 0; t0 < b.tiles_total; t0 += l.pass_tiles) {
        const uint64_t t1 = t0 + l.pass_tiles < b.tiles_total ? t0 + l.pass_tiles : b.tiles_total;
        for (uint32_t sample = 0; sample < p.samples; ++sample) {
            hipLaunchKernelGGL(k_batch_begin, dim3(1), block, 0, stream, f,
                               (uint32_t)(t0 == 0 && sample == 0));
            hipLaunchKernelGGL(k_batch_generate, grid, block, 0, stream, b, f, p, sample, t0, t1,
                               0u);
            // camera rays carry no jitter: later samples reuse the first's hits
            const bool reuse = p.samples > 1 && f.hits_primary != nullptr;
            if (reuse && sample > 0)
                hipLaunchKernelGGL(k_hits_keep, grid, block, 0, stream, f, 1u);
            for (uint32_t bn = 0; bn < p.bounces; ++bn) {
                if (!(reuse && sample > 0 && bn == 0))
                    hipLaunchKernelGGL(k_batch_extend, grid, block, 0, stream, b.scenes, f, p, bn,
                                       l.refill);
                if (reuse && sample == 0 && bn == 0)
                    hipLaunchKernelGGL(k_hits_keep, grid, block, 0, stream, f, 0u);
                hipLaunchKernelGGL((k_shade_plain_t<Q, true>), grid, block, 0, stream, none,
                                   b.scenes, f, p, bn);
            }
        }
    }
    hipLaunchKernelGGL(k_resolve, grid, block, 0, stream, f, p);
    return hipGetLastError();
}

hipError_t launch_batch(const BatchDev& b, const FrameDev& f, const FrameDev* fg,
                        const TraceParams& p, const BatchLaunch& l, hipStream_t stream) {
    if (b.num_items == 0 || b.tiles_total == 0 || l.pass_tiles == 0 || l.blocks == 0)
        return hipSuccess;
    switch (p.quality) {
        case kQualityLow:
            return launch_batch_q<kQualityLow>(b, f, fg, p, l, stream);
        case kQualityHigh:
            return launch_batch_q<kQualityHigh>(b, f, fg, p, l, stream);
        case kQualitySpectral:
            return launch_batch_q<kQualitySpectral>(b, f, fg, p, l, stream);
        default:
            return launch_batch_q<kQualityMedium>(b, f, fg, p, l, stream);
    }
}

hipError_t launch_trace(const SceneDev& s, const CameraDev& cam, const FrameDev& f,
                        const TraceParams& p, hipStream_t stream, KernelTimer* timer) {
    if (f.npix == 0) return hipSuccess;
    switch (p.quality) {
        case kQualityLow:
            return launch_plain<kQualityLow>(s, cam, f, p, stream, timer);
        case kQualityHigh:
            return launch_plain<kQualityHigh>(s, cam, f, p, stream, timer);
        case kQualitySpectral:
            return launch_plain<kQualitySpectral>(s, cam, f, p, stream, timer);
        default:
            return launch_plain<kQualityMedium>(s, cam, f, p, stream, timer);
    }
}

}  // namespace lmb
