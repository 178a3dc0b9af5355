// guided.hip - the reference's full path (Markov-chain guiding, world hash
// grid, ReSTIR reservoir per pixel) as a wavefront pipeline.
//
// rt_main (general.wesl:506-610) runs SAMPLES x rt_sample
// (general.wesl:643-975) per pixel in one thread.  Here each sample is:
//
//   k_g_generate   seed, tentative chain at sample start, screen-space
//                  candidate draws (fill_mc_screen_space_states, :367-385),
//                  primary ray
//   per bounce:
//     k_extend     closest hits (pathtrace.hip)
//     k_g_shade    hit record, world-grid training of the previous vertex
//                  (:674-685 / :926-935), emission, quality hooks, Fresnel
//     k_g_guide    world-space candidates (:389-400), 21-candidate
//                  resampling (:404-425), scattering with the cosine / vMF
//                  mixture (:784-901); survivors compacted into the next
//                  queue
//   k_g_finish     final world-grid update (:944-948) and rt_main's per-
//                  sample tail (:535-581)
//   k_g_shuffle    subgroupShuffle of the tentative chains (:349-361, 582)
// and after the samples k_g_resolve writes info, chains, images and the
// reservoir (:585-606).
//
// A logical subgroup is `kSubgroup` consecutive pixels of a row (the
// reference's 64x1 workgroup on a wave64 device).  World-grid updates are
// racy exactly as in the reference (relaxed atomics, CAS key claim,
// try-lock); the order in which paths touch the grid differs from any
// particular run of the megakernel.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "dev_common.h"
#include "dev_compact.h"
#include "dev_guide.h"
#include "dev_restir.h"
#include "dev_shade.h"
#include "lmb_frame_dev.h"
#include "lmb_guided.h"

namespace lmb {

constexpr int kGBlock = 256;
constexpr uint32_t kSubgroup = 64;
constexpr uint32_t kFlagNotVmf = 1u;
constexpr uint32_t kFlagDone = 2u;

static uint32_t g_blocks(uint64_t n) {
    uint64_t b = (n + kGBlock - 1) / kGBlock;
    uint64_t cap = (uint64_t)device_cu_count() * 8;
    return (uint32_t)std::max<uint64_t>(1, std::min(b, cap));
}

__device__ __forceinline__ float4 f4(V3 v, float w) { return make_float4(v.x, v.y, v.z, w); }
__device__ __forceinline__ V3 xyz(float4 a) { return v3(a.x, a.y, a.z); }

__device__ __forceinline__ MarkovChain load_mc_or_zero(const MarkovChain* p, uint64_t count,
                                                       uint64_t idx) {
    if (idx < count) return p[idx];
    MarkovChain z{};
    return z;
}

__global__ __launch_bounds__(kGBlock) void k_g_generate(CameraDev cam, FrameDev f, GuidedDev g,
                                                        TraceParams p, uint32_t sample) {
    const uint64_t gtid = (uint64_t)blockIdx.x * kGBlock + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * kGBlock;
    if (gtid == 0) {
        f.counters[0] = (uint32_t)f.npix;
        for (uint32_t i = 1; i <= kMaxBounces; ++i) f.counters[i] = 0;
        for (uint32_t k = 0; k < 2 * kFetchSegments; ++k) f.counters[kCounterFetch + k * kFetchStride] = 0;
        if (sample == 0) *reinterpret_cast<uint64_t*>(f.counters + kCounterVertices) = 0;
    }
    for (uint64_t pix = gtid; pix < f.npix; pix += stride) {
        const uint32_t x = (uint32_t)pix % f.width, y = (uint32_t)pix / f.width;
        uint32_t own_seed;
        if (sample == 0) {
            own_seed = rand_u32(x * y) + p.seed_offset + y;
            g.tent_snap[pix] = load_mc_or_zero(f.mc_old, f.mc_count, pix);
            g.sum_mc[pix] = 0.0f;
            f.pix_sum[pix] = make_float4(0.f, 0.f, 0.f, 0.f);
            g.out_rad_sum[pix] = make_float4(0.f, 0.f, 0.f, 0.f);
            // The running reservoir lives in HBM only when every sample
            // adds to it (per-sample reuse); the per-pixel sample is written
            // whole by k_g_finish before k_g_resolve reads it.
            if (p.restir == 1u) g.pix_res[pix] = Reservoir{};
        } else {
            own_seed = f.own_seed[pix];
        }
        g.tent[pix] = g.tent_snap[pix];
        g.cand_seed[pix] = own_seed;
        const ScreenCands sc = screen_candidates(own_seed, x, y, f.width, kSubgroup);
        own_seed = sc.seed_after;
        f.own_seed[pix] = own_seed;
        float o[3], d[3];
        create_ray(cam, x, y, f.width, f.height, o, d);
        f.queue[0][2 * pix] = make_float4(o[0], o[1], o[2], __uint_as_float((uint32_t)pix));
        f.queue[0][2 * pix + 1] = make_float4(d[0], d[1], d[2], 0.f);
        f.ps_seed[pix] = own_seed;
        g.color[pix] = make_float4(1.f, 1.f, 1.f, -1.f);
        g.scol[pix] = make_float4(0.f, 0.f, 0.f, 0.f);
        g.orc[pix] = make_float4(1.f, 1.f, 1.f, 0.f);
        g.sout[pix] = make_float4(0.f, 0.f, 0.f, 0.f);
        g.aem[pix] = make_float4(0.f, 0.f, 0.f, 0.f);
        g.aalb[pix] = make_float4(1.f, 1.f, 1.f, 0.f);
        // SampleReturn() with ty = SKY and pdf = 1 (general.wesl:649-652)
        g.s_point[pix] = make_float4(0.f, 0.f, 0.f, __uint_as_float(kMatSky));
        g.s_normal[pix] = make_float4(0.f, 0.f, 0.f, 1.f);
        g.s_albedo[pix] = make_float4(0.f, 0.f, 0.f, 0.f);
        g.s_emission[pix] = make_float4(0.f, 0.f, 0.f, 0.f);
        g.s_vis_point[pix] = make_float4(0.f, 0.f, 0.f, 0.f);
        g.s_vis_normal[pix] = make_float4(0.f, 0.f, 0.f, 0.f);
        g.s_start_dir[pix] = make_float4(d[0], d[1], d[2], 0.f);
    }
}

// World-grid training of the previous vertex when the path reaches
// `new_pos` carrying radiance_out (+ reflected estimate).
__device__ __forceinline__ void train_previous(const FrameDev& f, const GuidedDev& g,
                                               V3 camera_loc, uint32_t pid, V3 prev_pos,
                                               V3 new_pos, V3 total, V3 radiance_out,
                                               uint32_t bounce, bool last_not_vmf,
                                               uint32_t* seed) {
    const float4 ln = g.lastn[pid], ld = g.lastd[pid];
    const GridKey kp = grid_key(prev_pos, xyz(ln), xyz(ld), camera_loc, f.world_size);
    const float weight = length3(total);
    const float sum_mc = g.sum_mc[pid];
    // grid_add_radiance, then for weight > 0 grid_load / update / grid_store
    grid_add_and_train(f.wmc, f.wmc_old, kp, total, weight > 0.0f, [&](const Chain& wc) {
        return update_markov_state(prev_pos, new_pos, weight, *seed, false, wc, 1.0f, sum_mc);
    });
    if (weight > 0.0f) {
        const Chain t = chain_load(g.tent[pid]);
        g.tent[pid] = chain_store(update_markov_state(prev_pos, new_pos, length3(radiance_out),
                                                      *seed, last_not_vmf, t,
                                                      bounce == 1 ? 1.0f : 0.5f, sum_mc));
        *seed += 1u;
    }
}

// At most 128 VGPRs (4 waves per SIMD, one 8-byte spill for the spectral
// and high qualities): 3 waves at 142 VGPRs were slower.
template <uint32_t Q>
__global__ __launch_bounds__(kGBlock, 4) void k_g_shade(SceneDev s, FrameDev f, GuidedDev g,
                                                        TraceParams p, uint32_t bounce,
                                                        V3 camera_loc) {
    const uint64_t gtid = (uint64_t)blockIdx.x * kGBlock + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * kGBlock;
    const uint32_t count = f.counters[bounce];
    float4* __restrict__ q = f.queue[bounce & 1];
    for (uint64_t slot = gtid; slot < count; slot += stride) {
        const float4 o4 = q[2 * slot];
        const float4 d4 = q[2 * slot + 1];
        const uint32_t pid = __float_as_uint(o4.w);
        const uint4 h = f.hits[slot];
        uint32_t seed = f.ps_seed[pid];
        const float4 c4 = g.color[pid];
        V3 color = xyz(c4);
        float picked = c4.w;
        const float4 sc4 = g.scol[pid];
        V3 scol = xyz(sc4);
        const float cached_pdf = sc4.w;
        const float4 orc4 = g.orc[pid];
        V3 orc = xyz(orc4);
        const float cached_brdf = orc4.w;
        const float4 so4 = g.sout[pid];
        V3 sout = xyz(so4);
        const bool last_not_vmf = (__float_as_uint(so4.w) & kFlagNotVmf) != 0u;
        const V3 dir = xyz(d4);
        const V3 org = xyz(o4);
        if (h.x != kMissPrim) {
            const float t = __uint_as_float(h.y);
            const Isect it =
                make_isect(s, h.x, t, __uint_as_float(h.z), __uint_as_float(h.w), dir);
            const V3 new_pos = fma3(dir, v3s(t), org);
            if (bounce != 0) {
                const GridKey kn = grid_key(new_pos, it.normal, dir, camera_loc, f.world_size);
                const V3 est = grid_radiance(f.wmc, f.wmc_old, kn);
                const V3 reflected = strip_nan3(((it.color * est) * cached_brdf) / cached_pdf);
                const V3 rout = (it.emission * cached_brdf) / cached_pdf;
                train_previous(f, g, camera_loc, pid, org, new_pos, rout + reflected, rout, bounce,
                               last_not_vmf, &seed);
            }
            g.lastn[pid] = f4(it.normal, 0.f);
            g.lastd[pid] = f4(dir, 0.f);
            V3 aem = fma3(it.emission, xyz(g.aalb[pid]), xyz(g.aem[pid]));
            V3 aalb = xyz(g.aalb[pid]) * it.color;
            g.aem[pid] = f4(aem, 0.f);
            g.aalb[pid] = f4(aalb, 0.f);
            if (bounce == 0) {
                g.s_point[pid] = f4(new_pos, __uint_as_float(it.ty));
                g.s_emission[pid] = f4(aem, 0.f);
                g.s_normal[pid] = f4(it.normal, g.s_normal[pid].w);
                g.s_albedo[pid] = f4(aalb, 0.f);
            } else if (bounce == 1) {
                g.s_vis_point[pid] = f4(new_pos, 0.f);
                g.s_vis_normal[pid] = f4(it.normal, 0.f);
            }
            scol = fma3(color, it.emission, scol);
            if (bounce != 0) sout = fma3(orc, it.emission, sout);
            const V3 assign_normal =
                get_normal<Q>(dir, it.normal, it.tangent, seed, it.rough_x, it.rough_y);
            const V3 transparent_normal =
                get_normal<Q>(dir, it.normal, it.tangent, seed, it.rough_tx, it.rough_ty);
            V3 normal = transparent_normal;
            bool should_reflect = false;
            seed += 1u;
            V3 ccol;
            float ri = pick_ref_idx<Q>(&picked, it.ior_low, it.ior_high, seed, &ccol);
            seed += 1u;
            if (!(!it.front_face && it.ty == kMatTransparent)) ri = 1.0f / ri;
            color = color * ccol;
            orc = orc * ccol;
            if (ri != 1.0f) {
                const float cos_theta = fminf(dot3(-dir, normal), 1.0f);
                const float sin_theta = sqrtf(1.0f - cos_theta * cos_theta);
                const bool cannot_refract = ri * sin_theta > 1.0f;
                const float pdf = reflectance(cos_theta, ri);
                if (cannot_refract || pdf > rand_f32(seed)) should_reflect = true;
                seed += 1u;
            }
            if (!should_reflect || it.ty == kMatTransparent) {
                color = color * it.color;
                if (bounce != 0) orc = orc * it.color;
            }
            if (!should_reflect) {
                normal = assign_normal;
                seed += 1u;
            }
            q[2 * slot] = f4(new_pos, o4.w);
            g.gr0[slot] = f4(it.normal, it.rough_x);
            g.gr1[slot] = f4(it.tangent, ri);
            g.gr2[slot] = f4(normal, __uint_as_float(it.ty | (should_reflect ? 0x100u : 0u)));
            g.gr3[slot] = f4(it.color, 0.f);
            f.ps_seed[pid] = seed;
            g.color[pid] = f4(color, picked);
            g.scol[pid] = f4(scol, cached_pdf);
            g.orc[pid] = f4(orc, cached_brdf);
            g.sout[pid] = f4(sout, __uint_as_float(0u));
        } else {
            const V3 bg = background(s, dir);
            const V3 new_pos = fma3(dir, v3s(p.tmax), org);
            if (bounce == 1) {
                g.s_vis_point[pid] = f4(new_pos, 0.f);
                g.s_vis_normal[pid] = f4(-dir, 0.f);
            }
            if (bounce != 0) {
                sout = fma3(orc, bg, sout);
                const V3 rout = (bg * cached_brdf) / cached_pdf;
                train_previous(f, g, camera_loc, pid, org, new_pos, rout, rout, bounce, false, &seed);
            }
            scol = fma3(color, bg, scol);
            g.gr2[slot] = make_float4(0.f, 0.f, 0.f, __uint_as_float(kFlagDone << 16));
            f.ps_seed[pid] = seed;
            g.scol[pid] = f4(scol, cached_pdf);
            g.sout[pid] = f4(sout, __uint_as_float(0u));
        }
    }
}

template <uint32_t Q>
__global__ __launch_bounds__(kGBlock) void k_g_guide(SceneDev s, FrameDev f, GuidedDev g,
                                                     TraceParams p, uint32_t bounce,
                                                     V3 camera_loc) {
    const uint64_t gtid = (uint64_t)blockIdx.x * kGBlock + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * kGBlock;
    const uint32_t count = f.counters[bounce];
    const uint64_t count_r = block_uniform_count<kGBlock>(count);
    const float4* __restrict__ qin = f.queue[bounce & 1];
    float4* __restrict__ qout = f.queue[(bounce + 1) & 1];
    __shared__ ReserveLds<kGBlock, 1> rl;
    const float jitter_size = kGridSize * 1.5f;  // general.wesl:387
    for (uint64_t slot = gtid; slot < count_r; slot += stride) {
        bool cont = false;
        float4 out_o = make_float4(0.f, 0.f, 0.f, 0.f), out_d = out_o;
        if (slot < count) {
            const float4 g2 = g.gr2[slot];
            const uint32_t meta = __float_as_uint(g2.w);
            const float4 o4 = qin[2 * slot];
            const uint32_t pid = __float_as_uint(o4.w);
            if (((meta >> 16) & kFlagDone) == 0u) {
                const V3 origin = xyz(o4);
                const V3 dir = xyz(qin[2 * slot + 1]);
                const float4 g0 = g.gr0[slot], g1 = g.gr1[slot], g3 = g.gr3[slot];
                const V3 inormal = xyz(g0), tangent = xyz(g1), normal = xyz(g2), hcolor = xyz(g3);
                const float rough_x = g0.w, ri = g1.w;
                const uint32_t ty = meta & 0xFFu;
                const bool should_reflect = (meta & 0x100u) != 0u;
                uint32_t seed = f.ps_seed[pid];
                // fill_mc_world_space_states (general.wesl:389-400): the
                // candidates' keys first, then their cells together
                Chain wc[kWorldResamples];
                GridKey wk[kWorldResamples];
                bool wforce[kWorldResamples];
#pragma unroll
                for (uint32_t i = 0; i < kWorldResamples; ++i) {
                    const bool rand_candidate = rand_f32(seed) < 0.7f;
                    const float jitter_mul = rand_candidate ? 0.75f : 1.0f;
                    seed = rand_u32(seed);
                    const float jx = (float)seed * 2.3283064365386963e-10f;
                    seed = rand_u32(seed);
                    const float jz = (float)seed * 2.3283064365386963e-10f;
                    const V3 jitter =
                        v3(jx - floorf(jx), 0.0f, jz - floorf(jz)) * jitter_size * jitter_mul;
                    const V3 pos = origin + from_onb_no_tangent(jitter, inormal);
                    wk[i] = grid_key(pos, inormal, dir, camera_loc, f.world_size);
                    wforce[i] = rand_candidate;
                    seed += 1u;
                }
                // k_g_guide writes no key or chain of the grid: plain reads
                grid_load_n<true>(f.wmc, f.wmc_old, wk, wforce, wc);
                const Chain tent = chain_load(g.tent[pid]);
                const uint32_t x = pid % f.width, y = pid / f.width;
                const ScreenCands sc = screen_candidates(g.cand_seed[pid], x, y, f.width, kSubgroup);
                const Chain lane_chain = sc.lane_ok ? chain_load(g.tent_snap[sc.lane_pix]) : chain_zero();
                // resample_mcmm (general.wesl:404-425).  Candidate i: the
                // shuffled lane chain (i < 4), last frame's screen chains
                // (12), the world-space chains (4), the tentative chain; the
                // loops run in that order with constant array indices, so
                // the candidate arrays stay in registers.
                Chain sel = chain_zero();
                float sum_mc = 0.0f;
                auto resample = [&](const Chain& cand) {
                    if (resample_step(cand, seed, sum_mc)) sel = cand;
                };
                for (uint32_t i = 0; i < kScreenResamples; ++i) resample(lane_chain);
#pragma unroll
                for (uint32_t i = 0; i < kScreenLastFrame; ++i)
                    resample(chain_load(load_mc_or_zero(f.mc_old, f.mc_count, sc.idx[i])));
#pragma unroll
                for (uint32_t i = 0; i < kWorldResamples; ++i) resample(wc[i]);
                resample(tent);
                static_assert(kNumResamples == kScreenResamples + kScreenLastFrame + kWorldResamples + 1,
                              "candidate groups");
                g.sum_mc[pid] = sum_mc;
                float4 c4 = g.color[pid];
                V3 color = xyz(c4);
                float4 orc4 = g.orc[pid];
                V3 orc = xyz(orc4);
                float cached_pdf = g.scol[pid].w;
                float cached_brdf = orc4.w;
                uint32_t flags = 0u;
                V3 ndir = dir;
                const bool reflecting = should_reflect || ty == kMatMetallic;
                if (!reflecting) {
                    if (ty == kMatDiffuse) {
                        const V3 light_dir = get_light_dir(sel, origin);
                        const V3 mean = normalize3(light_dir);
                        const float kappa = get_kappa(sel);
                        const V3 dir_cos = cosine_weighted_hemisphere(seed, inormal, tangent);
                        const float cos_pdf = fmaxf(dot3(dir_cos, inormal) / kPi, 0.0f);
                        const V3 Xs = vmf_sample(mean, kappa, seed);
                        seed += 2u;
                        const V3 dir_markov = normalize3(Xs);
                        float c0 = 0.2f, c1 = 0.8f;
                        if (!(sel.weight_sum > 0.0f) || near_zero(length3(light_dir)) ||
                            near_zero(sel.weight_sum) || near_zero(length3(Xs))) {
                            c0 = 1.0f;
                            c1 = 0.0f;
                        }
                        const float csum = c0 + c1;
                        const float conf0 = c0 / csum, conf1 = c1 / csum;
                        const float rnd = rand_f32(seed);
                        seed += 1u;
                        V3 d;
                        if (rnd < conf0) {
                            d = dir_cos;
                            flags |= kFlagNotVmf;
                        } else {
                            d = dir_markov;
                        }
                        // all_vmf_pdfs (general.wesl:984-998) over the same candidates
                        float mix_pdf = 0.0f;
                        if (conf1 != 0.0f && sum_mc > 0.0f) {
                            // the candidates in resample_mcmm's order
                            auto add_pdf = [&](const Chain& cand) {
                                mix_pdf += mixture_term(cand, origin, d, sum_mc);
                            };
                            for (uint32_t i = 0; i < kScreenResamples; ++i) add_pdf(lane_chain);
#pragma unroll
                            for (uint32_t i = 0; i < kScreenLastFrame; ++i)
                                add_pdf(chain_load(load_mc_or_zero(f.mc_old, f.mc_count, sc.idx[i])));
#pragma unroll
                            for (uint32_t i = 0; i < kWorldResamples; ++i) add_pdf(wc[i]);
                            add_pdf(tent);
                        }
                        const float avg_pdf =
                            conf0 * cos_pdf + (conf1 == 0.0f ? 0.0f : conf1 * fmaxf(mix_pdf, 0.0f));
                        const V3 diffuse =
                            make_diffuse<Q>(hcolor, dir, inormal, tangent, rough_x, d);
                        cached_pdf = avg_pdf;
                        cached_brdf = fmaxf(dot3(d, inormal) / kPi, 0.0f);
                        const float nf = fabsf(cached_brdf) / fmaxf(avg_pdf, 0.001f);
                        color = (color * nf) * diffuse;
                        orc = (orc * nf) * diffuse;
                        ndir = d;
                        if (bounce == 0) {
                            float4 sn = g.s_normal[pid];
                            sn.w = avg_pdf;
                            g.s_normal[pid] = sn;
                        }
                    } else if (ty == kMatTransparent) {
                        cached_pdf = 1.0f;
                        cached_brdf = 1.0f;
                        ndir = refract3(dir, normal, ri);
                    }
                } else {
                    cached_pdf = 1.0f;
                    cached_brdf = 1.0f;
                    ndir = reflect3(dir, normal);
                    if (dot3(ndir, inormal) < 0.0f) ndir = reflect3(ndir, inormal);
                }
                bool alive = bounce + 1 < p.bounces;
                if (alive && bounce >= p.rr_start) {
                    const float qq = fminf(fmaxf(fmaxf(color.x, color.y), color.z), 1.0f);
                    const float rnd = rand_f32(seed);
                    seed += 1u;
                    if (rnd >= qq) {
                        alive = false;
                    } else {
                        color = color / qq;
                        orc = orc / qq;
                    }
                }
                f.ps_seed[pid] = seed;
                g.color[pid] = f4(color, c4.w);
                g.orc[pid] = f4(orc, cached_brdf);
                float4 s4 = g.scol[pid];
                s4.w = cached_pdf;
                g.scol[pid] = s4;
                float4 so4 = g.sout[pid];
                so4.w = __uint_as_float(flags);
                g.sout[pid] = so4;
                if (alive) {
                    cont = true;
                    out_o = o4;
                    out_d = f4(ndir, 0.f);
                }
            }
        }
        // Compact the survivors into the next queue: one atomic per block.
        uint32_t idx[1];
        block_reserve<kGBlock, 1>({cont}, {f.counters + bounce + 1}, rl, idx);
        if (cont) {
            qout[2 * (uint64_t)idx[0]] = out_o;
            qout[2 * (uint64_t)idx[0] + 1] = out_d;
        }
    }
}

// Final world-grid update of the sample (general.wesl:944-948) and the
// per-sample tail of rt_main (general.wesl:535-581).
//
// Writers of a shared cell.  In the reference's lock-step execution, the
// lanes of a wave that write one cell have one lane's stores land and one
// lane take the cell's try-lock: one write per wave and key.  That is kept
// for every cell but one: the cell of the pixels whose camera ray missed
// (point and normal zero, general.wesl:649-652), which every sky pixel of the
// frame shares; lane by lane its same-address atomics queue in L2 (7.3 ms of
// a 1080p frame).  For that cell the waves of a pixel tile claim it in a
// small LDS table (the 64-bit key, 0 = empty, one probe), so it is written
// once per tile instead of once per wave.
constexpr uint32_t kClaimSlots = 256;
// A pass of k_g_finish covers a square tile of pixels in 16 x 16 sub-tiles,
// one lane per pixel, with one claim table for the tile.
constexpr uint32_t kFinishSub = 16;
constexpr uint32_t kFinishTile = 32;
static_assert(kFinishSub * kFinishSub == kGBlock, "one lane per pixel of a finish sub-tile");
static_assert(kFinishTile % kFinishSub == 0, "whole sub-tiles per finish tile");

template <uint32_t Q>
__global__ __launch_bounds__(kGBlock) void k_g_finish(FrameDev f, GuidedDev g, TraceParams p,
                                                      V3 camera_loc) {
    __shared__ unsigned long long claim[kClaimSlots];
    const uint32_t tiles_x = (f.width + kFinishTile - 1) / kFinishTile;
    const uint32_t tiles = tiles_x * ((f.height + kFinishTile - 1) / kFinishTile);
    for (uint32_t tile = blockIdx.x; tile < tiles; tile += gridDim.x) {
        for (uint32_t i = threadIdx.x; i < kClaimSlots; i += kGBlock) claim[i] = 0ull;
        __syncthreads();
        constexpr uint32_t kSubs = kFinishTile / kFinishSub;
        for (uint32_t sub = 0; sub < kSubs * kSubs; ++sub) {
            const uint32_t x = (tile % tiles_x) * kFinishTile + (sub % kSubs) * kFinishSub +
                               threadIdx.x % kFinishSub;
            const uint32_t y = (tile / tiles_x) * kFinishTile + (sub / kSubs) * kFinishSub +
                               threadIdx.x / kFinishSub;
            const bool live = x < f.width && y < f.height;
            const uint64_t pix = live ? (uint64_t)y * f.width + x : 0;
            const float4 sp0 = g.s_point[pix];
            const GridKey k = grid_key(xyz(sp0), xyz(g.s_normal[pix]), xyz(g.s_start_dir[pix]),
                                       camera_loc, f.world_size);
            bool writer = false;
            if (live && first_writer_of_key(k)) {
                if (__float_as_uint(sp0.w) == kMatSky) {
                    const unsigned long long kw = key_word(k);
                    const unsigned long long prev = atomicCAS(&claim[k.idx % kClaimSlots], 0ull, kw);
                    writer = prev == 0ull || prev != kw;  // another key in the slot: write anyway
                } else {
                    writer = true;
                }
            }
            if (!live) continue;
            uint32_t seed = f.ps_seed[pix];
            const float4 sp = g.s_point[pix], sn = g.s_normal[pix];
            const V3 point = xyz(sp), normal = xyz(sn);
            const V3 vis_point = xyz(g.s_vis_point[pix]);
            const V3 sout = xyz(g.sout[pix]);
            if (writer) {
                // Only the writer's chain update lands, so only it loads
                // the chain; another lane's load would differ in nothing
                // but a timeout write that the writer's store repeats.
                const Chain wc = grid_load(f.wmc, f.wmc_old, k, false);
                const Chain st = update_markov_state(point, vis_point, length3(sout), seed, false, wc,
                                                     1.0f, g.sum_mc[pix]);
                grid_store(f.wmc, k, st);
                grid_add_radiance(f.wmc, k, sout);
            }
            // rt_main
            const float4 ps = f.pix_sum[pix];
            const V3 scol = xyz(g.scol[pix]);
            f.pix_sum[pix] = make_float4(ps.x + scol.x, ps.y + scol.y, ps.z + scol.z, 0.f);
            const float4 ors = g.out_rad_sum[pix];
            const V3 out_rad = xyz(ors) + sout;
            g.out_rad_sum[pix] = f4(out_rad, 0.f);
            uint32_t own_seed = rand_u32(f.own_seed[pix]);
            const uint32_t ty = __float_as_uint(sp.w);
            if (p.restir == 2u) {  // RESTIR_OPTION_PER_PIXEL: remember the last sample
                Reservoir smp{};
                const V3 vp = vis_point, vn = xyz(g.s_vis_normal[pix]);
                smp.visible_point[0] = vp.x;
                smp.visible_point[1] = vp.y;
                smp.visible_point[2] = vp.z;
                smp.visible_normal = normalised_to_u32(vn);
                smp.sample_point[0] = point.x;
                smp.sample_point[1] = point.y;
                smp.sample_point[2] = point.z;
                smp.sample_normal = normalised_to_u32(normal);
                const V3 L = out_rad / (float)p.samples;
                smp.out_radiance[0] = L.x;
                smp.out_radiance[1] = L.y;
                smp.out_radiance[2] = L.z;
                smp.ty = ty;
                smp.roughness = conv_rough<Q>(0.0f);
                smp.packed_confidence_valid = 1u;  // sample_valid
                g.pix_sample[pix] = smp;  // rt_main's pdf sum (:561) is never read
            } else if (p.restir == 1u) {  // RESTIR_OPTION_PER_SAMPLE
                Reservoir R = g.pix_res[pix];
                if (ty != kMatSky) {
                    Reservoir smp{};
                    smp.visible_point[0] = vis_point.x;
                    smp.visible_point[1] = vis_point.y;
                    smp.visible_point[2] = vis_point.z;
                    smp.visible_normal = normalised_to_u32(xyz(g.s_vis_normal[pix]));
                    smp.sample_point[0] = point.x;
                    smp.sample_point[1] = point.y;
                    smp.sample_point[2] = point.z;
                    smp.sample_normal = normalised_to_u32(normal);
                    smp.out_radiance[0] = sout.x;
                    smp.out_radiance[1] = sout.y;
                    smp.out_radiance[2] = sout.z;
                    smp.ty = ty;
                    smp.roughness = conv_rough<Q>(0.0f);
                    const float w = p_hat(smp);
                    R.w = R.w + w;
                    R.packed_confidence_valid =
                        ((R.packed_confidence_valid + 1u) & kMaxConfidence) |
                        (R.packed_confidence_valid & ~kMaxConfidence);
                    if (rand_f32(own_seed) * R.w < w) assign_sample(R, smp);
                    own_seed = rand_u32(own_seed);
                    R.W = safe_div(R.w, (float)confidence_of(R.packed_confidence_valid) * p_hat(R));
                } else {
                    R.ty = kMatSky;
                }
                g.pix_res[pix] = R;
            }
            f.own_seed[pix] = own_seed;
            g.pix_last_normal[pix] = f4(normal, __uint_as_float(ty));
            g.pix_last_albedo[pix] = g.s_albedo[pix];
            g.pix_last_emission[pix] = g.s_emission[pix];
        }
        __syncthreads();  // every claim of this tile made before the table is cleared
    }
}

// subgroupShuffle(tentative, (lane + 1) % sub_size) over the active lanes.
__global__ __launch_bounds__(kGBlock) void k_g_shuffle(FrameDev f, GuidedDev g) {
    const uint64_t stride = (uint64_t)gridDim.x * kGBlock;
    for (uint64_t pix = (uint64_t)blockIdx.x * kGBlock + threadIdx.x; pix < f.npix;
         pix += stride) {
        const uint32_t x = (uint32_t)pix % f.width, y = (uint32_t)pix / f.width;
        const uint32_t base = (x / kSubgroup) * kSubgroup;
        const uint32_t active = min(kSubgroup, f.width - base);
        const uint32_t src = base + ((x - base + 1u) % active);
        g.tent_snap[pix] = g.tent[(uint64_t)y * f.width + src];
    }
}

// general.wesl:585-606
__global__ __launch_bounds__(kGBlock) void k_g_resolve(FrameDev f, GuidedDev g, TraceParams p,
                                                       V3 camera_loc) {
    const uint64_t stride = (uint64_t)gridDim.x * kGBlock;
    const float ns = (float)p.samples;
    for (uint64_t pix = (uint64_t)blockIdx.x * kGBlock + threadIdx.x; pix < f.npix;
         pix += stride) {
        const float4 ln = g.pix_last_normal[pix];
        const float4 la = g.pix_last_albedo[pix];
        const float4 le = g.pix_last_emission[pix];
        if (p.restir != 0u) {
            CompressedInfo info;
            info.arr[0] = camera_loc.x;
            info.arr[1] = camera_loc.y;
            info.arr[2] = camera_loc.z;
            info.arr[3] = le.x;
            info.arr[4] = le.y;
            info.arr[5] = le.z;
            info.arr[6] = la.x;
            info.arr[7] = la.y;
            info.arr[8] = la.z;
            f.info[pix] = info;
        }
        if (pix < f.mc_count) f.mc[pix] = g.tent_snap[pix];
        f.out_normal[pix] = make_float4(ln.x, ln.y, ln.z, 1.f);
        f.out_albedo[pix] = make_float4(la.x, la.y, la.z, 1.f);
        const float4 s4 = f.pix_sum[pix];
        f.out_color[pix] = make_float4(fmaxf(s4.x / ns, 0.f), fmaxf(s4.y / ns, 0.f),
                                       fmaxf(s4.z / ns, 0.f), 1.f);
        Reservoir R = p.restir == 1u ? g.pix_res[pix] : Reservoir{};
        if (p.restir == 2u) {
            const Reservoir smp = g.pix_sample[pix];
            if (smp.ty != kMatSky && smp.packed_confidence_valid != 0u) {
                uint32_t own_seed = f.own_seed[pix];
                const float w = p_hat(smp);
                R.w = R.w + w;
                R.packed_confidence_valid = ((R.packed_confidence_valid + 1u) & kMaxConfidence) |
                                            (R.packed_confidence_valid & ~kMaxConfidence);
                if (rand_f32(own_seed) * R.w < w) assign_sample(R, smp);
                own_seed = rand_u32(own_seed);
                f.own_seed[pix] = own_seed;
                R.W = safe_div(R.w, (float)confidence_of(R.packed_confidence_valid) * p_hat(R));
            } else if (smp.ty == kMatSky) {
                R.ty = kMatSky;
            }
        }
        if (pix < f.gi_count) f.gi[pix] = R;
    }
}

template <uint32_t Q>
static hipError_t launch_guided_q(const SceneDev& s, const CameraDev& cam, const FrameDev& f,
                                  const GuidedDev& g, const TraceParams& p, hipStream_t stream,
                                  KernelTimer* tm) {
    const dim3 grid(g_blocks(f.npix)), block(kGBlock);
    const V3 camera_loc = V3{cam.view_inverse[12], cam.view_inverse[13], cam.view_inverse[14]};
    for (uint32_t sample = 0; sample < p.samples; ++sample) {
        hipLaunchKernelGGL(k_g_generate, grid, block, 0, stream, cam, f, g, p, sample);
        for (uint32_t b = 0; b < p.bounces; ++b) {
            if (tm) tm->begin(kKernelExtend, stream);
            hipError_t e = launch_extend(s, f, p, b, stream);
            if (e != hipSuccess) return e;
            if (tm) tm->end(kKernelExtend, stream);
            if (tm) tm->begin(kKernelShade, stream);
            hipLaunchKernelGGL(k_g_shade<Q>, grid, block, 0, stream, s, f, g, p, b, camera_loc);
            hipLaunchKernelGGL(k_g_guide<Q>, grid, block, 0, stream, s, f, g, p, b, camera_loc);
            if (tm) tm->end(kKernelShade, stream);
        }
        hipLaunchKernelGGL(k_g_finish<Q>, grid, block, 0, stream, f, g, p, camera_loc);
        hipLaunchKernelGGL(k_g_shuffle, grid, block, 0, stream, f, g);
    }
    hipLaunchKernelGGL(k_g_resolve, grid, block, 0, stream, f, g, p, camera_loc);
    return hipGetLastError();
}

hipError_t launch_guided(const SceneDev& s, const CameraDev& cam, const FrameDev& f,
                         const GuidedDev& g, const TraceParams& p, hipStream_t stream,
                         KernelTimer* timer) {
    if (f.npix == 0) return hipSuccess;
    if (f.world_size == 0) return hipErrorInvalidValue;
    switch (p.quality) {
        case kQualityLow:
            return launch_guided_q<kQualityLow>(s, cam, f, g, p, stream, timer);
        case kQualityHigh:
            return launch_guided_q<kQualityHigh>(s, cam, f, g, p, stream, timer);
        case kQualitySpectral:
            return launch_guided_q<kQualitySpectral>(s, cam, f, g, p, stream, timer);
        default:
            return launch_guided_q<kQualityMedium>(s, cam, f, g, p, stream, timer);
    }
}

}  // namespace lmb
