// nrc_trace.hip - the path tracer with the neural radiance cache.
//
// Per sample: generate -> extend / shade for every bounce -> inference.
// Shading is the plain tracer's (surface_scatter, Russian roulette) except at
// the second vertex (bounce 1):
//   * in a trace's first sample, about train_expect of the paths reaching it
//     (a hash of path id, frame and sample picks them, with the probability
//     calibrated on the previous sample's count of cache vertices) become
//     training records: the record keeps the
//     network inputs there, the throughput T and the radiance S gathered so
//     far; the path restarts its throughput at 1 and its sum at 0 and is
//     traced on to full depth; when it ends, its sum (the radiance scattered
//     at the vertex towards the previous one) is the record's target and the
//     pixel receives S + T * sum;
//   * every other path ends there with a query: the pixel receives S now and
//     T * max(cache(inputs), 0) from the inference kernel.
// The emission of the second vertex belongs to S (the cache predicts
// scattered radiance only).
// With defer_specular = n > 0, a vertex is a cache vertex only when the
// scatter before it spread the path (a diffuse lobe or one of roughness
// >= 0.1) or when it is n vertices past the second: a path seen through
// smooth glass or a mirror goes on to the vertex after its first rough
// scatter (the spread rule of neural radiance caching), so the cache's error
// there is blurred by a lobe instead of shown sharply, and specular chains
// stop n vertices late at most (their short queues are latency-bound).
// With np.stratify the first scatter (at the camera ray's hit) takes its
// direction pair and Fresnel choice from the pixel's Owen-scrambled Sobol
// sequence (dev_qmc.h), point seq_base + sample index, with seq_base the
// seed offset times the trace's sample count: the samples of a frame (and of
// frames with consecutive seed offsets) spread over the lobe instead of
// falling independently, and
// every point is still uniform, so each sample stays unbiased.  Once the
// cache ends the path at its next vertex, that scatter carries most of a
// sample's variance.
// Every path still meets at most one cache vertex; training continuations
// (rslot != 0) never meet another.  Later samples of a trace reuse the
// camera-ray hits, carry no training paths and stop after the last possible
// cache vertex.  A training frame (train_only) traces the camera rays of
// random pixels (the fraction that gives about train_expect cache
// vertices), makes every cache vertex a training record and writes no image.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dev_common.h"
#include "dev_compact.h"
#include "dev_qmc.h"
#include "dev_shade.h"
#include "lmb_frame_dev.h"
#include "lmb_nrc.h"
#include "lmb_restir.h"

namespace lmb {

constexpr int kNrcTraceBlock = 256;
constexpr uint32_t kCacheBounce = 1;  // the second path vertex: the earliest cache vertex
constexpr uint32_t kArmedBit = 0x80000000u;  // ps_accum.w: the last scatter spread the path
constexpr float kSpreadRoughness = 0.1f;

// Whether a scatter off `it` spreads the path (see defer_specular above).
template <uint32_t Q>
__device__ __forceinline__ bool scatter_spreads(const Isect& it) {
    if (it.ty == kMatDiffuse) return true;
    if (Q == kQualityLow) return false;  // Low has no roughness: mirrors and smooth glass
    if (it.ty == kMatMetallic) return fmaxf(it.rough_x, it.rough_y) >= kSpreadRoughness;
    return fmaxf(it.rough_tx, it.rough_ty) >= kSpreadRoughness;
}

// A query / training record as four 16-byte stores (the layout of NrcQuery).
__device__ __forceinline__ void store_query(NrcQuery* q, const float (&in)[kNrcInputs], uint32_t pid) {
    float4* o = reinterpret_cast<float4*>(q);
    o[0] = make_float4(in[0], in[1], in[2], in[3]);
    o[1] = make_float4(in[4], in[5], in[6], in[7]);
    o[2] = make_float4(in[8], in[9], in[10], in[11]);
    o[3] = make_float4(in[12], __uint_as_float(pid), 0.0f, 0.0f);
}

__device__ __forceinline__ float train_draw(uint32_t pid, uint32_t seed_offset, uint32_t sample) {
    const uint32_t salt = rand_u32(seed_offset * 0x9E3779B1u + sample * 0x85EBCA77u + 0x165667B1u);
    return (float)(rand_u32(pid ^ salt) >> 8) * (1.0f / 16777216.0f);
}

template <uint32_t Q>
__global__ __launch_bounds__(kNrcTraceBlock) void k_shade_nrc(SceneDev s, FrameDev f, TraceParams p,
                                                              NrcDev d, NrcTraceParams np,
                                                              uint32_t bounce) {
    const uint64_t gtid = (uint64_t)blockIdx.x * kNrcTraceBlock + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * kNrcTraceBlock;
    const uint32_t count = f.counters[bounce];
    const uint64_t count_r = block_uniform_count<kNrcTraceBlock>(count);
    const float4* __restrict__ qin = f.queue[bounce & 1];
    float4* __restrict__ qout = f.queue[(bounce + 1) & 1];
    __shared__ ReserveLds<kNrcTraceBlock, 1> rl_rec;
    __shared__ ReserveLds<kNrcTraceBlock, 2> rl_out;
    if (bounce == 0 && gtid == 0) {
        // No lane reserves queries or records at bounce 0.
        const uint32_t rec = min(d.ctl[kNrcCtlRecords], d.rec_cap);
        d.ctl[kNrcCtlQueries] = 0;
        if (np.sample == 0) d.ctl[kNrcCtlRecords] = 0;  // a new trace
        d.ctl[kNrcCtlRecordsBase] = np.sample == 0 ? 0u : rec;
    }
    float prob = 0.0f;
    if (bounce >= kCacheBounce) {
        // Selection probability for about train_expect records per sample;
        // before any sample has been counted, the paths reaching bounce 1.
        const uint32_t prev = d.ctl[kNrcCtlHitsPrev];
        const uint32_t ref = prev ? prev : max(f.counters[kCacheBounce], 1u);
        prob = np.train_only ? 1.0f : fminf(np.train_expect / (float)ref, 1.0f);
    }
    for (uint64_t slot = gtid; slot < count_r; slot += stride) {
        // ---- phase A: hit record and the cache-vertex inputs
        const bool active = slot < count;
        float4 o4 = make_float4(0.f, 0.f, 0.f, 0.f), d4 = o4;
        uint4 h = make_uint4(kMissPrim, 0u, 0u, 0u);
        uint32_t pid = 0, seed = 0, rslot = 0;
        bool armed = false;
        V3 color = v3s(0.f), scol = v3s(0.f);
        float picked = -1.0f;
        if (active && np.wide > 1u && bounce == 0) {
            // Wide group: bounce-0 slot c * wide_count + j is sample
            // sample + j of the pixel of camera-ray slot c (qin and hits are
            // the camera pass's), path id pixel * wide + j; the path starts
            // fresh with k_generate's seed for that sample.
            const uint32_t c = (uint32_t)slot / np.wide_count;
            const uint32_t j = (uint32_t)slot - c * np.wide_count;
            o4 = qin[2 * (uint64_t)c];
            d4 = qin[2 * (uint64_t)c + 1];
            h = f.hits[c];
            const uint32_t pixel = __float_as_uint(o4.w);
            pid = pixel * np.wide + j;
            const uint32_t x = pixel % f.width, y = pixel / f.width;
            seed = rand_u32(x * y) + p.seed_offset + y;  // general.wesl:516, 544
            for (uint32_t k = 0; k < np.sample + j; ++k) seed = rand_u32(seed);
            color = v3s(1.0f);
        } else if (active) {
            o4 = qin[2 * slot];
            d4 = qin[2 * slot + 1];
            pid = __float_as_uint(o4.w);
            h = f.hits[slot];
            seed = f.ps_seed[pid];
            const float4 th = f.ps_through[pid];
            const float4 ac = f.ps_accum[pid];
            color = v3(th.x, th.y, th.z);
            picked = th.w;
            scol = v3(ac.x, ac.y, ac.z);
            rslot = __float_as_uint(ac.w);
            armed = (rslot & kArmedBit) != 0u;
            rslot &= ~kArmedBit;
        }
        // the pixel of an image output (a wide group writes them from j = 0)
        const bool img_px = np.wide > 1u ? pid % np.wide == 0u : true;
        const uint32_t opix = np.wide > 1u ? pid / np.wide : pid;
        const bool hit = active && h.x != kMissPrim;
        const V3 dir = v3(d4.x, d4.y, d4.z);
        const V3 org = v3(o4.x, o4.y, o4.z);
        Isect it{};
        V3 new_pos = v3s(0.f);
        bool at_cache = false, cand = false;
        float in[kNrcInputs];
        if (hit) {
            const float t = __uint_as_float(h.y);
            it = make_isect(s, h.x, t, __uint_as_float(h.z), __uint_as_float(h.w), dir);
            new_pos = fma3(dir, v3s(t), org);
            if (bounce == 0 && !np.train_only && img_px) {
                f.out_normal[opix] = make_float4(it.normal.x, it.normal.y, it.normal.z, 1.f);
                f.out_albedo[opix] = make_float4(it.color.x, it.color.y, it.color.z, 1.f);
            }
            scol = fma3(color, it.emission, scol);  // general.wesl:714
            if (bounce >= kCacheBounce && armed && rslot == 0u) {
                at_cache = true;
                in[0] = (new_pos.x - np.lo[0]) * np.inv_ext[0];
                in[1] = (new_pos.y - np.lo[1]) * np.inv_ext[1];
                in[2] = (new_pos.z - np.lo[2]) * np.inv_ext[2];
                in[3] = 0.5f - 0.5f * dir.x;  // towards the previous vertex
                in[4] = 0.5f - 0.5f * dir.y;
                in[5] = 0.5f - 0.5f * dir.z;
                in[6] = 0.5f + 0.5f * it.normal.x;
                in[7] = 0.5f + 0.5f * it.normal.y;
                in[8] = 0.5f + 0.5f * it.normal.z;
                in[9] = saturate(it.rough_x);
                in[10] = saturate(it.color.x);
                in[11] = saturate(it.color.y);
                in[12] = saturate(it.color.z);
                cand = train_draw(pid, p.seed_offset, np.sample) < prob;
            }
        }
        // ---- phase B (whole block): training-record slots
        uint32_t r_idx[1];
        block_reserve<kNrcTraceBlock, 1>({cand}, {d.ctl + kNrcCtlRecords}, rl_rec, r_idx);
        const uint32_t r = r_idx[0];
        if (cand && r >= d.rec_cap) cand = false;
        const bool is_query = at_cache && !cand;
        // ---- phase C: scattering, continuation or the end of the path
        bool cont = false;
        float4 out_o = make_float4(0.f, 0.f, 0.f, 0.f), out_d = out_o;
        bool finish = false;
        if (active) {
            if (hit) {
                if (cand) {
                    store_query(d.rec_in + r, in, pid);
                    d.rec_aux[2 * (uint64_t)r] = make_float4(color.x, color.y, color.z, scol.x);
                    d.rec_aux[2 * (uint64_t)r + 1] = make_float4(scol.y, scol.z, 0.f, 0.f);
                    rslot = r + 1u;
                    color = v3s(1.0f);
                    scol = v3s(0.0f);
                }
                if (is_query) {
                    if (!np.train_only) {
                        const float4 ps = f.pix_sum[pid];
                        f.pix_sum[pid] = make_float4(ps.x + scol.x, ps.y + scol.y, ps.z + scol.z, 0.f);
                    }
                } else {
                    V3 ndir;
                    if (bounce == 0 && np.stratify) {
                        // the pixel's stratified sequence; sample index
                        // seq_base + sample (+ j of a wide group)
                        const uint32_t j = np.wide > 1u ? pid % np.wide : 0u;
                        float u[3];
                        qmc_point3(np.seq_base + np.sample + j, np.wide > 1u ? pid / np.wide : pid, u);
                        ndir = surface_scatter<Q, true>(it, dir, seed, color, picked, {u[0], u[1], u[2]});
                    } else {
                        ndir = surface_scatter<Q>(it, dir, seed, color, picked);
                    }
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
                        out_d = make_float4(ndir.x, ndir.y, ndir.z, 0.f);
                        f.ps_seed[pid] = seed;
                        f.ps_through[pid] = make_float4(color.x, color.y, color.z, picked);
                        const bool arm = bounce + 1u >= kCacheBounce + np.defer_specular ||
                                         scatter_spreads<Q>(it);
                        f.ps_accum[pid] = make_float4(scol.x, scol.y, scol.z,
                                                      __uint_as_float(rslot | (arm ? kArmedBit : 0u)));
                    } else {
                        finish = true;
                    }
                }
            } else {
                if (bounce == 0 && !np.train_only && img_px) {
                    f.out_normal[opix] = make_float4(0.f, 0.f, 0.f, 1.f);
                    f.out_albedo[opix] = make_float4(0.f, 0.f, 0.f, 1.f);
                }
                scol = fma3(color, background(s, dir), scol);  // general.wesl:915-938
                finish = true;
            }
            if (finish) {
                V3 pix = scol;
                if (rslot) {
                    const uint64_t rr = rslot - 1u;
                    const float4 a0 = d.rec_aux[2 * rr], a1 = d.rec_aux[2 * rr + 1];
                    d.rec_target[rr] = make_float4(scol.x, scol.y, scol.z, 0.f);
                    pix = v3(fmaf(a0.x, scol.x, a0.w), fmaf(a0.y, scol.y, a1.x),
                             fmaf(a0.z, scol.z, a1.y));
                }
                if (!np.train_only) {
                    const float4 ps = f.pix_sum[pid];
                    f.pix_sum[pid] = make_float4(ps.x + pix.x, ps.y + pix.y, ps.z + pix.z, 0.f);
                }
            }
        }
        // ---- phase D (whole block): queries and the continuation queue
        uint32_t o_idx[2];
        block_reserve<kNrcTraceBlock, 2>({is_query, cont}, {d.ctl + kNrcCtlQueries, f.counters + bounce + 1},
                                         rl_out, o_idx);
        const uint32_t qi = o_idx[0], ci = o_idx[1];
        // A training frame counts the cache vertices that found no record
        // slot but stores no query: it runs no inference, and its pass may
        // run beside a frame whose queries fill the same buffer.
        if (is_query && !np.train_only) {
            store_query(d.queries + qi, in, pid);
            d.query_weight[qi] = make_float4(color.x, color.y, color.z, 0.f);
        }
        if (cont) {
            qout[2 * (uint64_t)ci] = out_o;
            qout[2 * (uint64_t)ci + 1] = out_d;
        }
    }
}

// Camera rays of a training frame: each pixel with probability `frac`
// (about train_expect cache vertices: the fraction a full sample would
// select there), compacted into the bounce-0 queue.  Seeds as in k_generate
// for a first sample.
__global__ __launch_bounds__(kNrcTraceBlock) void k_nrc_generate_train(CameraDev cam, FrameDev f,
                                                                       TraceParams p, NrcDev d,
                                                                       NrcTraceParams np) {
    __shared__ ReserveLds<kNrcTraceBlock, 1> rl;
    const uint64_t stride = (uint64_t)gridDim.x * kNrcTraceBlock;
    // before any count: as if every pixel reached a cache vertex
    const uint32_t prev = d.ctl[kNrcCtlHitsPrev];
    const float frac = fminf(np.train_expect / (float)(prev ? prev : max(f.npix, (uint64_t)1)), 1.0f);
    if (blockIdx.x == 0 && threadIdx.x == 0) d.ctl[kNrcCtlFrac] = __float_as_uint(frac);
    const uint64_t n_r = block_uniform_count<kNrcTraceBlock>((uint32_t)f.npix);
    for (uint64_t pix = (uint64_t)blockIdx.x * kNrcTraceBlock + threadIdx.x; pix < n_r;
         pix += stride) {
        const bool sel = pix < f.npix && train_draw((uint32_t)pix, p.seed_offset, 0xFFFFu) < frac;
        uint32_t idx[1];
        block_reserve<kNrcTraceBlock, 1>({sel}, {f.counters}, rl, idx);
        if (!sel) continue;
        const uint32_t x = (uint32_t)pix % f.width, y = (uint32_t)pix / f.width;
        const uint32_t seed = rand_u32(x * y) + p.seed_offset + y;
        float o[3], dd[3];
        create_ray(cam, x, y, f.width, f.height, o, dd);
        f.queue[0][2 * (uint64_t)idx[0]] = make_float4(o[0], o[1], o[2], __uint_as_float((uint32_t)pix));
        f.queue[0][2 * (uint64_t)idx[0] + 1] = make_float4(dd[0], dd[1], dd[2], 0.f);
        f.ps_seed[pix] = seed;
        f.ps_through[pix] = make_float4(1.f, 1.f, 1.f, -1.f);
        f.ps_accum[pix] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// Start of a wide group of samples s0 .. s0 + w - 1: the bounce-0 queue
// length (the bounce-0 slots are implicit, camera slot * w + j) and the other
// queues' counters, and with the trace's first group the sums of every path
// id (pixel * stride + j, j < stride).  The path-vertex count is left to the
// camera pass.
__global__ __launch_bounds__(kNrcTraceBlock) void k_nrc_begin_wide(FrameDev f, uint32_t s0,
                                                                   uint32_t w, uint32_t stride_w) {
    const uint64_t gtid = (uint64_t)blockIdx.x * kNrcTraceBlock + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * kNrcTraceBlock;
    if (gtid == 0) {
        f.counters[0] = (uint32_t)((uint64_t)w * f.npix);
        for (uint32_t i = 1; i <= kMaxBounces; ++i) f.counters[i] = 0;
        for (uint32_t k = 0; k < 2 * kFetchSegments; ++k) f.counters[kCounterFetch + k * kFetchStride] = 0;
    }
    if (s0 != 0) return;
    const uint64_t total = (uint64_t)stride_w * f.npix;
    for (uint64_t pid = gtid; pid < total; pid += stride) f.pix_sum[pid] = make_float4(0.f, 0.f, 0.f, 0.f);
}

// The image of a wide trace: each pixel's w sums (path ids pixel * w + j)
// over the trace's samples, clamped at zero as k_resolve does.
__global__ __launch_bounds__(kNrcTraceBlock) void k_nrc_resolve_wide(FrameDev f, uint32_t samples,
                                                                     uint32_t w) {
    const uint64_t gtid = (uint64_t)blockIdx.x * kNrcTraceBlock + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * kNrcTraceBlock;
    const float ns = (float)samples;
    for (uint64_t pix = gtid; pix < f.npix; pix += stride) {
        float4 t = f.pix_sum[pix * w];
        for (uint32_t j = 1; j < w; ++j) {
            const float4 u = f.pix_sum[pix * w + j];
            t = make_float4(t.x + u.x, t.y + u.y, t.z + u.z, 0.f);
        }
        f.out_color[pix] = make_float4(fmaxf(t.x / ns, 0.f), fmaxf(t.y / ns, 0.f),
                                       fmaxf(t.z / ns, 0.f), 1.f);
    }
}

// End of a sample (of `per` samples traced together): its cache vertices are
// its queries plus the records it stored (a record that found no room became
// a query); a training frame's count is scaled to the whole image by its
// pixel fraction.  The count per sample sets the selection probability of
// the next training paths.
__global__ void k_nrc_end_sample(NrcDev d, uint32_t train_only, uint32_t per) {
    const uint32_t rec = min(d.ctl[kNrcCtlRecords], d.rec_cap);
    const uint32_t hits = d.ctl[kNrcCtlQueries] + (rec - min(d.ctl[kNrcCtlRecordsBase], rec));
    float full = (float)hits;
    if (train_only) {
        const float frac = __uint_as_float(d.ctl[kNrcCtlFrac]);
        full = frac > 0.0f ? fminf((float)hits / frac, 4.0e9f) : 0.0f;
    }
    full /= (float)per;
    if (full >= 1.0f) d.ctl[kNrcCtlHitsPrev] = (uint32_t)full;
}

template <uint32_t Q>
static hipError_t launch_nrc_train(const SceneDev& s, const CameraDev& cam, const FrameDev& f,
                                   const TraceParams& p, const NrcDev& d, NrcTraceParams np,
                                   hipStream_t stream, KernelTimer* tm) {
    const dim3 grid((uint32_t)(frame_threads(grid_paths(f)) / kNrcTraceBlock)), block(kNrcTraceBlock);
    // queue lengths, this trace's vertex count and both traversal fetch sets
    hipError_t e = launch_fill_zero(f.counters, (kMaxBounces + 1) * 4, stream);
    if (e == hipSuccess) e = launch_fill_zero(f.counters + kCounterVertices, 8, stream);
    if (e == hipSuccess) e = launch_fill_zero(f.counters + kCounterFetch, 2 * kFetchWords * 4, stream);
    if (e != hipSuccess) return e;
    np.sample = 0;
    hipLaunchKernelGGL(k_nrc_generate_train, grid, block, 0, stream, cam, f, p, d, np);
    for (uint32_t b = 0; b < p.bounces; ++b) {
        if (tm) tm->begin(kKernelExtend, stream);
        e = launch_extend(s, f, p, b, stream);
        if (tm) tm->end(kKernelExtend, stream);
        if (e != hipSuccess) return e;
        if (tm) tm->begin(kKernelShade, stream);
        hipLaunchKernelGGL(k_shade_nrc<Q>, grid, block, 0, stream, s, f, p, d, np, b);
        if (tm) tm->end(kKernelShade, stream);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_nrc_end_sample, dim3(1), dim3(1), 0, stream, d, 1u, 1u);
    return hipGetLastError();
}

template <uint32_t Q>
static hipError_t launch_nrc_q(const SceneDev& s, const CameraDev& cam, const FrameDev& f,
                               const TraceParams& p, const NrcDev& d, NrcTraceParams np,
                               hipStream_t stream, KernelTimer* tm) {
    const dim3 grid((uint32_t)(frame_threads(f.npix) / kNrcTraceBlock)), block(kNrcTraceBlock);
    hipError_t e;
    const FrameDev f0 = primary_view(f, p);
    const float train_expect = np.train_expect;
    // Without training paths every path has ended by the bounce after its
    // latest possible cache vertex.
    const uint32_t short_bounces = std::min(p.bounces, kCacheBounce + np.defer_specular + 1u);
    for (uint32_t sample = 0; sample < p.samples; ++sample) {
        // Training paths only in the first sample: later samples end at
        // their cache vertices and skip the long latency-bound tail of the
        // training paths.
        np.sample = sample;
        np.train_expect = sample == 0 ? train_expect : 0.0f;
        e = launch_generate(cam, f, p, sample, stream);
        if (e != hipSuccess) return e;
        const uint32_t bounces = np.train_expect > 0.0f ? p.bounces : short_bounces;
        for (uint32_t b = 0; b < bounces; ++b) {
            const FrameDev& fb = b == 0 ? f0 : f;
            if (!(b == 0 && sample > 0 && f0.hits != f.hits)) {
                if (tm) tm->begin(kKernelExtend, stream);
                e = launch_extend(s, fb, p, b, stream);
                if (tm) tm->end(kKernelExtend, stream);
                if (e != hipSuccess) return e;
            }
            if (tm) tm->begin(kKernelShade, stream);
            hipLaunchKernelGGL(k_shade_nrc<Q>, grid, block, 0, stream, s, fb, p, d, np, b);
            if (tm) tm->end(kKernelShade, stream);
            e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k_nrc_end_sample, dim3(1), dim3(1), 0, stream, d, 0u, 1u);
        if (p.bounces > kCacheBounce) {
            if (tm) tm->begin(kKernelCacheInfer, stream);
            e = nrc_launch_infer_accumulate(d, f.pix_sum, stream);
            if (tm) tm->end(kKernelCacheInfer, stream);
            if (e != hipSuccess) return e;
        }
    }
    return launch_resolve(f, p, stream);
}

// A trace whose samples are traced w at a time (no training paths): the
// camera rays once into the camera-ray hits (f_img, bounce 0), then per group
// of w samples one set of bounces and one inference over all of them in fw
// (state, queues, hits and sums of w * npix path ids), so the latency-bound
// tail of each bounce (the deferred specular paths) is paid once per group
// instead of once per sample.  The samples of a pixel have adjacent slots and
// path ids: their camera ray, hit and surface data are read together, and
// their secondary rays leave one point together.
template <uint32_t Q>
static hipError_t launch_nrc_wide_q(const SceneDev& s, const CameraDev& cam, const FrameDev& f_img,
                                    const FrameDev& fw, const TraceParams& p, const NrcDev& d,
                                    NrcTraceParams np, uint32_t w, hipStream_t stream,
                                    KernelTimer* tm) {
    const dim3 grid((uint32_t)(frame_threads(fw.npix) / kNrcTraceBlock)), block(kNrcTraceBlock);
    const FrameDev f0 = primary_view(f_img, p);
    hipError_t e = launch_generate(cam, f_img, p, 0, stream);
    if (e != hipSuccess) return e;
    if (tm) tm->begin(kKernelExtend, stream);
    e = launch_extend(s, f0, p, 0, stream);
    if (tm) tm->end(kKernelExtend, stream);
    if (e != hipSuccess) return e;
    FrameDev fw0 = fw;  // bounce 0: the camera rays and their hits, by camera slot
    fw0.queue[0] = f0.queue[0];
    fw0.hits = f0.hits;
    const uint32_t bounces = std::min(p.bounces, kCacheBounce + np.defer_specular + 1u);
    np.train_expect = 0.0f;
    np.wide = w;
    for (uint32_t s0 = 0; s0 < p.samples; s0 += w) {
        const uint32_t wc = std::min(w, p.samples - s0);
        hipLaunchKernelGGL(k_nrc_begin_wide, grid, block, 0, stream, fw, s0, wc, w);
        np.sample = s0;
        np.wide_count = wc;
        for (uint32_t b = 0; b < bounces; ++b) {
            const FrameDev& fb = b == 0 ? fw0 : fw;
            if (b > 0) {
                if (tm) tm->begin(kKernelExtend, stream);
                e = launch_extend(s, fb, p, b, stream);
                if (tm) tm->end(kKernelExtend, stream);
                if (e != hipSuccess) return e;
            }
            if (tm) tm->begin(kKernelShade, stream);
            hipLaunchKernelGGL(k_shade_nrc<Q>, grid, block, 0, stream, s, fb, p, d, np, b);
            if (tm) tm->end(kKernelShade, stream);
            e = hipGetLastError();
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL(k_nrc_end_sample, dim3(1), dim3(1), 0, stream, d, 0u, wc);
        if (p.bounces > kCacheBounce) {
            if (tm) tm->begin(kKernelCacheInfer, stream);
            e = nrc_launch_infer_accumulate(d, fw.pix_sum, stream);
            if (tm) tm->end(kKernelCacheInfer, stream);
            if (e != hipSuccess) return e;
        }
    }
    hipLaunchKernelGGL(k_nrc_resolve_wide, grid, block, 0, stream, fw, p.samples,
                       std::min(w, p.samples));
    return hipGetLastError();
}

hipError_t launch_trace_nrc_wide(const SceneDev& s, const CameraDev& cam, const FrameDev& f_img,
                                 const FrameDev& fw, const TraceParams& p, const NrcDev& d,
                                 NrcTraceParams np, uint32_t w, hipStream_t stream,
                                 KernelTimer* timer) {
    if (f_img.npix == 0) return hipSuccess;
    switch (p.quality) {
        case kQualityLow:
            return launch_nrc_wide_q<kQualityLow>(s, cam, f_img, fw, p, d, np, w, stream, timer);
        case kQualityHigh:
            return launch_nrc_wide_q<kQualityHigh>(s, cam, f_img, fw, p, d, np, w, stream, timer);
        case kQualitySpectral:
            return launch_nrc_wide_q<kQualitySpectral>(s, cam, f_img, fw, p, d, np, w, stream, timer);
        default:
            return launch_nrc_wide_q<kQualityMedium>(s, cam, f_img, fw, p, d, np, w, stream, timer);
    }
}

hipError_t launch_trace_nrc(const SceneDev& s, const CameraDev& cam, const FrameDev& f,
                            const TraceParams& p, const NrcDev& d, NrcTraceParams np,
                            hipStream_t stream, KernelTimer* timer) {
    if (f.npix == 0) return hipSuccess;
    if (np.train_only) {
        switch (p.quality) {
            case kQualityLow:
                return launch_nrc_train<kQualityLow>(s, cam, f, p, d, np, stream, timer);
            case kQualityHigh:
                return launch_nrc_train<kQualityHigh>(s, cam, f, p, d, np, stream, timer);
            case kQualitySpectral:
                return launch_nrc_train<kQualitySpectral>(s, cam, f, p, d, np, stream, timer);
            default:
                return launch_nrc_train<kQualityMedium>(s, cam, f, p, d, np, stream, timer);
        }
    }
    switch (p.quality) {
        case kQualityLow:
            return launch_nrc_q<kQualityLow>(s, cam, f, p, d, np, stream, timer);
        case kQualityHigh:
            return launch_nrc_q<kQualityHigh>(s, cam, f, p, d, np, stream, timer);
        case kQualitySpectral:
            return launch_nrc_q<kQualitySpectral>(s, cam, f, p, d, np, stream, timer);
        default:
            return launch_nrc_q<kQualityMedium>(s, cam, f, p, d, np, stream, timer);
    }
}

}  // namespace lmb
