// dev_guide.h - device restatement of the reference's Markov-chain (vMF)
// path guiding and its world-space hash grid / radiance cache:
//
//   strip_NaN, hashes, grid_loc     general.wesl:117-188
//   grid_load / store / radiance    general.wesl:192-311 (relaxed atomics,
//                                   CAS key claim, try-lock: racy by design)
//   update_marcov_state             general.wesl:464-492
//   get_light_dir                   general.wesl:495-498
//   get_kappa / vmf_pdf / expm1     general.wesl:1000-1047
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dev_common.h"
#include "dev_restir.h"
#include "dev_shade.h"
#include "lmb_layout.h"

namespace lmb {

// bindings.wesl:3-9 MarcovChainState
struct Chain {
    V3 ls;
    float mean_cosine;
    float weight_sum;
    uint32_t num_samples;
    float score;
};

__device__ __forceinline__ Chain chain_zero() {
    Chain c;
    c.ls = v3s(0.0f);
    c.mean_cosine = 0.0f;
    c.weight_sum = 0.0f;
    c.num_samples = 0u;
    c.score = 0.0f;
    return c;
}
__device__ __forceinline__ Chain chain_load(const MarkovChain& m) {
    Chain c;
    c.ls = v3(m.light_source[0], m.light_source[1], m.light_source[2]);
    c.mean_cosine = m.mean_cosine;
    c.weight_sum = m.weight_sum;
    c.num_samples = m.num_samples;
    c.score = m.score;
    return c;
}
__device__ __forceinline__ MarkovChain chain_store(const Chain& c) {
    MarkovChain m;
    m.light_source[0] = c.ls.x;
    m.light_source[1] = c.ls.y;
    m.light_source[2] = c.ls.z;
    m.mean_cosine = c.mean_cosine;
    m.weight_sum = c.weight_sum;
    m.num_samples = c.num_samples;
    m.score = c.score;
    m.pad = 0u;
    return m;
}

constexpr uint32_t kNumResamples = 21;     // general.wesl:341-345
constexpr uint32_t kScreenResamples = 4;
constexpr uint32_t kScreenLastFrame = 12;
constexpr uint32_t kWorldResamples = 4;
constexpr uint32_t kResampleSize = 91;     // general.wesl:363
constexpr uint32_t kChainSampleClamp = 1024;
constexpr uint32_t kCacheTimeout = 30;
constexpr uint32_t kMaxNumSamples = 0x0FFFFFFFu;
constexpr float kGridSize = 0.1f;

// general.wesl:120-131: +-Inf -> 10000, NaN -> 0
__device__ __forceinline__ float strip_nan(float f) {
    uint32_t b = __float_as_uint(f);
    if ((b & 0x7F800000u) == 0x7F800000u) return (b & 0x007FFFFFu) == 0u ? 10000.0f : 0.0f;
    return f;
}
__device__ __forceinline__ V3 strip_nan3(V3 v) {
    return v3(strip_nan(v.x), strip_nan(v.y), strip_nan(v.z));
}

// general.wesl:148-156
__device__ __forceinline__ uint32_t incoming_hash(V3 v) {
    return (v.x > 0.0f ? 1u : 0u) | (v.y > 0.0f ? 2u : 0u) | (v.z > 0.0f ? 4u : 0u);
}

struct GridKey {
    uint32_t idx;        // primary hash % table size
    uint32_t secondary;  // secondary hash (never 0)
};

// f32 -> i32 toward zero, saturating, NaN -> 0.
__device__ __forceinline__ int32_t f2i_sat(float f) {
    if (!(f == f)) return 0;
    if (f >= 2147483520.0f) return 2147483647;
    if (f <= -2147483648.0f) return (int32_t)0x80000000;
    return (int32_t)f;
}

// general.wesl:182-188 grid_loc.  num_shifts = u32(log2(max(len, 1))) is
// taken exactly from the float exponent (floor(log2 x) for x >= 1).
__device__ __forceinline__ void grid_loc(V3 pos, V3 camera_loc, int32_t loc[3]) {
    V3 d = pos - camera_loc;
    float len = fmaxf(fabsf(d.x), fmaxf(fabsf(d.y), fabsf(d.z)));
    float x = fmaxf(len / 1.0f, 1.0f);
    uint32_t shifts;
    if (!(x < 3.0e38f)) {
        shifts = x == x ? 127u : 0u;
    } else {
        shifts = (__float_as_uint(x) >> 23) - 127u;
    }
    uint32_t mask = 0xFFFFFFFFu << (shifts < 30u ? shifts : 30u);
    loc[0] = f2i_sat(pos.x / kGridSize) & (int32_t)mask;
    loc[1] = f2i_sat(pos.y / kGridSize) & (int32_t)mask;
    loc[2] = f2i_sat(pos.z / kGridSize) & (int32_t)mask;
}

// general.wesl:161-178 primary_hasher / secondary_hasher
__device__ __forceinline__ GridKey grid_key(V3 pos, V3 normal, V3 incoming, V3 camera_loc,
                                            uint32_t table_size) {
    int32_t loc[3];
    grid_loc(pos, camera_loc, loc);
    const uint32_t n = normalised_to_u32(normal) & 0xFF00FF00u;
    const uint32_t inc = rand_u32(incoming_hash(incoming));
    const uint32_t h1 = ((((uint32_t)loc[0] * 1u) ^ ((uint32_t)loc[1] * 687603047u)) ^
                         ((uint32_t)loc[2] * 571899763u)) ^
                        (n * 894271171u);
    const uint32_t primary = h1 ^ (inc * 840057787u);
    const uint32_t s3 = (((uint32_t)loc[0] * 1699644091u) ^ ((uint32_t)loc[1] * 550105817u)) ^
                        ((uint32_t)loc[2] * 956114303u);
    const uint32_t secondary = ((rand_u32(s3) ^ (n * 155096983u)) ^ (0u * 284005801u)) | 1u;
    GridKey k;
    k.idx = primary % table_size;
    k.secondary = secondary;
    return k;
}

__device__ __forceinline__ uint32_t aload(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void astore(uint32_t* p, uint32_t v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Two adjacent words of a cell in one relaxed 8-byte atomic access (the
// cell's words are 8-byte aligned in pairs from its start: 56-B cells in an
// array aligned to 256 B).  Reading or writing two of the reference's
// separately atomic words at one instant is one of the interleavings its
// racy grid allows, and it halves the device-scope (L2-bypassing) accesses.
__device__ __forceinline__ uint64_t aload2(const uint32_t* p) {
    return __hip_atomic_load(reinterpret_cast<const uint64_t*>(p), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void astore2(uint32_t* p, uint32_t lo, uint32_t hi) {
    __hip_atomic_store(reinterpret_cast<uint64_t*>(p), ((uint64_t)hi << 32) | lo, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t lo32(uint64_t v) { return (uint32_t)v; }
__device__ __forceinline__ uint32_t hi32(uint64_t v) { return (uint32_t)(v >> 32); }

// Plain (L2-cached) reads of cells that no thread writes while the kernel
// runs: last frame's table, which only advance_frame writes, and in
// k_g_guide the current table too (its writers are k_g_shade and
// k_g_finish, other launches on the same stream; k_g_guide itself writes
// only timeouts).  They return what the relaxed atomic loads would, and
// neighbouring vertices that read one cell share its L2 line instead of
// each reading it from memory (device-scope atomics bypass the L2).
__device__ __forceinline__ Chain world_chain_read_plain(const WorldMarkovChain* w) {
    const uint2 a = *reinterpret_cast<const uint2*>(&w->chain_light_source[0]);
    const uint2 b = *reinterpret_cast<const uint2*>(&w->chain_light_source[2]);
    const uint2 c2 = *reinterpret_cast<const uint2*>(&w->chain_weight_sum);
    const uint32_t d = w->chain_score;
    Chain c;
    c.ls = v3(__uint_as_float(a.x), __uint_as_float(a.y), __uint_as_float(b.x));
    c.mean_cosine = __uint_as_float(b.y);
    c.weight_sum = __uint_as_float(c2.x);
    c.num_samples = c2.y;
    c.score = __uint_as_float(d);
    return c;
}

__device__ __forceinline__ Chain world_chain_read(const WorldMarkovChain* w) {
    // words 6..13: light source (3), mean cosine, weight sum, samples, score,
    // timeout
    const uint64_t a = aload2(&w->chain_light_source[0]);
    const uint64_t b = aload2(&w->chain_light_source[2]);
    const uint64_t c2 = aload2(&w->chain_weight_sum);
    const uint64_t d = aload2(&w->chain_score);
    Chain c;
    c.ls = v3(__uint_as_float(lo32(a)), __uint_as_float(hi32(a)), __uint_as_float(lo32(b)));
    c.mean_cosine = __uint_as_float(hi32(b));
    c.weight_sum = __uint_as_float(lo32(c2));
    c.num_samples = hi32(c2);
    c.score = __uint_as_float(lo32(d));
    return c;
}

// general.wesl:192-226 grid_load.  The reference reads the current cell's
// chain before testing its key, and a forced or missed key then replaces it
// with last frame's chain or zero: the current chain is read here only when
// it is returned (the world candidates force last frame's table 70% of the
// time, and each skipped read is one or two 128-B lines).  The relaxed loads
// of the key and the chain are then made in the other order, which is one
// of the interleavings the reference's racy grid allows.
__device__ __forceinline__ Chain grid_load(WorldMarkovChain* cur, const WorldMarkovChain* old,
                                           const GridKey& k, bool force_old) {
    Chain s;
    if (force_old || aload(&cur[k.idx].secondary_hash) != k.secondary) {
        if (old[k.idx].secondary_hash == k.secondary)  // last frame's table: plain reads
            s = world_chain_read_plain(old + k.idx);
        else
            s = chain_zero();
    } else {
        s = world_chain_read(cur + k.idx);
        cur[k.idx].timeout = kCacheTimeout;
    }
    s.ls = strip_nan3(s.ls);
    s.mean_cosine = strip_nan(s.mean_cosine);
    s.weight_sum = strip_nan(s.weight_sum);
    s.score = strip_nan(s.score);
    return s;
}

// N grid_loads at once (fill_mc_world_space_states' candidates): the key
// words of both tables for every candidate are loaded together, then the
// chains of the cells whose keys matched, so a lane waits for two memory
// round trips instead of two per candidate.  Each candidate returns what
// grid_load returns for it; last frame's key is also read where the current
// cell matched (a relaxed load with no effect on the result).
// kCurStill: no thread writes the current table's keys or chains while
// the kernel runs (k_g_guide), so its reads are plain too.
template <bool kCurStill, int N>
__device__ __forceinline__ void grid_load_n(WorldMarkovChain* cur, const WorldMarkovChain* old,
                                            const GridKey (&k)[N], const bool (&force_old)[N],
                                            Chain (&out)[N]) {
    uint32_t kc[N], ko[N];
#pragma unroll
    for (int i = 0; i < N; ++i) {
        kc[i] = force_old[i] ? 0u
                : kCurStill  ? cur[k[i].idx].secondary_hash
                             : aload(&cur[k[i].idx].secondary_hash);
        ko[i] = old[k[i].idx].secondary_hash;
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const bool use_cur = !force_old[i] && kc[i] == k[i].secondary;  // the key is never 0
        if (use_cur) {
            out[i] = kCurStill ? world_chain_read_plain(cur + k[i].idx) : world_chain_read(cur + k[i].idx);
            cur[k[i].idx].timeout = kCacheTimeout;
        } else if (ko[i] == k[i].secondary) {
            out[i] = world_chain_read_plain(old + k[i].idx);
        } else {
            out[i] = chain_zero();
        }
    }
#pragma unroll
    for (int i = 0; i < N; ++i) {
        out[i].ls = strip_nan3(out[i].ls);
        out[i].mean_cosine = strip_nan(out[i].mean_cosine);
        out[i].weight_sum = strip_nan(out[i].weight_sum);
        out[i].score = strip_nan(out[i].score);
    }
}

// general.wesl:228-246 grid_store_markov_chain
__device__ __forceinline__ void grid_store(WorldMarkovChain* cur, const GridKey& k,
                                           const Chain& s) {
    WorldMarkovChain* w = cur + k.idx;
    if (aload(&w->secondary_hash) != k.secondary) {
        if (atomicCAS(&w->secondary_hash, 0u, k.secondary) != 0u) return;
    }
    astore2(&w->chain_light_source[0], __float_as_uint(s.ls.x), __float_as_uint(s.ls.y));
    astore2(&w->chain_light_source[2], __float_as_uint(s.ls.z), __float_as_uint(s.mean_cosine));
    astore2(&w->chain_weight_sum, __float_as_uint(s.weight_sum), s.num_samples);
    astore2(&w->chain_score, __float_as_uint(s.score), kCacheTimeout);  // and the timeout
}

// general.wesl:248-271 grid_load_radiance_estimate
__device__ __forceinline__ V3 grid_radiance(WorldMarkovChain* cur, const WorldMarkovChain* old,
                                            const GridKey& k) {
    // words 0..4: secondary hash, samples, luminance (3); the key read with
    // the samples (the reference reads it after the radiance)
    const WorldMarkovChain* wc = cur + k.idx;
    uint64_t h = aload2(&wc->secondary_hash);
    uint64_t l01 = aload2(&wc->luminance[0]);
    uint64_t l2 = aload2(&wc->luminance[2]);
    if (lo32(h) != k.secondary) {
        const WorldMarkovChain* wo = old + k.idx;  // last frame's table: plain reads
        h = *reinterpret_cast<const uint64_t*>(&wo->secondary_hash);
        if (lo32(h) == k.secondary) {
            l01 = *reinterpret_cast<const uint64_t*>(&wo->luminance[0]);
            l2 = *reinterpret_cast<const uint64_t*>(&wo->luminance[2]);
        } else {
            return v3s(0.0f);
        }
    } else {
        cur[k.idx].timeout = kCacheTimeout;
    }
    const uint32_t n = hi32(h);
    const V3 rad = v3(__uint_as_float(lo32(l01)), __uint_as_float(hi32(l01)), __uint_as_float(lo32(l2)));
    if (n == 0u) return v3s(0.0f);
    V3 r = rad / (float)n;
    return v3(fminf(fmaxf(r.x, 0.0f), 1000.0f), fminf(fmaxf(r.y, 0.0f), 1000.0f),
              fminf(fmaxf(r.z, 0.0f), 1000.0f));
}

// general.wesl:273-311 grid_add_radiance (try-lock: gives up when held)
__device__ __forceinline__ void grid_add_radiance(WorldMarkovChain* cur, const GridKey& k,
                                                  V3 radiance) {
    WorldMarkovChain* w = cur + k.idx;
    V3 c = v3(fminf(fmaxf(radiance.x, 0.0f), 100.0f), fminf(fmaxf(radiance.y, 0.0f), 100.0f),
              fminf(fmaxf(radiance.z, 0.0f), 100.0f));
    if (aload(&w->secondary_hash) != k.secondary) {
        if (atomicCAS(&w->secondary_hash, 0u, k.secondary) != 0u) return;
    }
    if (atomicExch(&w->lock, 1u) == 1u) return;
    // With the lock held: the sample count and the three sums read together
    // (words 1..5), the sums written back as a pair and a pair with the
    // lock's release (words 2..5).
    const uint32_t n = aload(&w->num_samples);
    const uint64_t l01 = aload2(&w->luminance[0]);
    const uint32_t l2 = aload(&w->luminance[2]);
    if (n >= kMaxNumSamples) {
        astore(&w->lock, 0u);
        w->timeout = kCacheTimeout;
        return;
    }
    atomicAdd(&w->num_samples, 1u);
    const float r = strip_nan(c.x + __uint_as_float(lo32(l01)));
    const float g = strip_nan(c.y + __uint_as_float(hi32(l01)));
    const float b = strip_nan(c.z + __uint_as_float(l2));
    astore2(&w->luminance[0], __float_as_uint(fmaxf(r, 0.0f)), __float_as_uint(fmaxf(g, 0.0f)));
    astore2(&w->luminance[2], __float_as_uint(fmaxf(b, 0.0f)), 0u);  // and the lock's release
    w->timeout = kCacheTimeout;
}

// grid_add_radiance, then (when `update`) grid_load, update and grid_store
// of the same key, as the training of a path vertex performs them
// (general.wesl:926-935), with the cell's key read once.  Within a launch a
// key only changes from 0 to a claimed value, so the add's key check and
// claim settle what the load's and the store's key checks would see: the
// cell is this key's if it held it or the claim found 0 or this key (a
// claim that found this key makes the add give up, as in grid_add_radiance,
// while the load and store still see the key).  The chain is read before the
// lock is taken, so its round trip overlaps the lock's; the chain words are
// not written by the add, and reading them before the add's stores is one
// of the orders the relaxed atomics allow.  upd(chain) returns the chain
// to store; it sees the loaded chain whenever `update` is set.
template <class Upd>
__device__ __forceinline__ void grid_add_and_train(WorldMarkovChain* cur, const WorldMarkovChain* old,
                                                   const GridKey& k, V3 radiance, bool update,
                                                   Upd&& upd) {
    WorldMarkovChain* w = cur + k.idx;
    bool ours = true, add = true;
    if (aload(&w->secondary_hash) != k.secondary) {
        const uint32_t prev = atomicCAS(&w->secondary_hash, 0u, k.secondary);
        ours = prev == 0u || prev == k.secondary;
        add = prev == 0u;
    }
    Chain s = chain_zero();
    if (update) {
        if (ours) {
            s = world_chain_read(w);
            w->timeout = kCacheTimeout;
        } else if (old[k.idx].secondary_hash == k.secondary) {  // last frame's table: plain reads
            s = world_chain_read_plain(old + k.idx);
        }
        s.ls = strip_nan3(s.ls);
        s.mean_cosine = strip_nan(s.mean_cosine);
        s.weight_sum = strip_nan(s.weight_sum);
        s.score = strip_nan(s.score);
    }
    if (add && atomicExch(&w->lock, 1u) != 1u) {
        const V3 c = v3(fminf(fmaxf(radiance.x, 0.0f), 100.0f), fminf(fmaxf(radiance.y, 0.0f), 100.0f),
                        fminf(fmaxf(radiance.z, 0.0f), 100.0f));
        const uint32_t n = aload(&w->num_samples);
        const uint64_t l01 = aload2(&w->luminance[0]);
        const uint32_t l2 = aload(&w->luminance[2]);
        if (n >= kMaxNumSamples) {
            astore(&w->lock, 0u);
            w->timeout = kCacheTimeout;
        } else {
            atomicAdd(&w->num_samples, 1u);
            const float r = strip_nan(c.x + __uint_as_float(lo32(l01)));
            const float g = strip_nan(c.y + __uint_as_float(hi32(l01)));
            const float b = strip_nan(c.z + __uint_as_float(l2));
            astore2(&w->luminance[0], __float_as_uint(fmaxf(r, 0.0f)), __float_as_uint(fmaxf(g, 0.0f)));
            astore2(&w->luminance[2], __float_as_uint(fmaxf(b, 0.0f)), 0u);  // and the lock's release
            w->timeout = kCacheTimeout;
        }
    }
    if (update) {
        const Chain st = upd(s);  // as grid_store, stored only when the cell is this key's
        if (!ours) return;
        astore2(&w->chain_light_source[0], __float_as_uint(st.ls.x), __float_as_uint(st.ls.y));
        astore2(&w->chain_light_source[2], __float_as_uint(st.ls.z), __float_as_uint(st.mean_cosine));
        astore2(&w->chain_weight_sum, __float_as_uint(st.weight_sum), st.num_samples);
        astore2(&w->chain_score, __float_as_uint(st.score), kCacheTimeout);  // and the timeout
    }
}

// general.wesl:495-498
__device__ __forceinline__ V3 get_light_dir(const Chain& s, V3 current) {
    return safe_div_v3(s.ls, s.weight_sum) - current;
}

__device__ __forceinline__ V3 mix3(V3 a, V3 b, float t) { return a * (1.0f - t) + b * t; }
__device__ __forceinline__ float mix1(float a, float b, float t) { return a * (1.0f - t) + b * t; }

// general.wesl:464-492 update_marcov_state
__device__ __forceinline__ Chain update_markov_state(V3 x1, V3 x2, float p_hat, uint32_t seed,
                                                     bool not_vmf, Chain state, float confidence,
                                                     float sum_mc) {
    if (not_vmf) state = chain_zero();
    const float score = p_hat;
    const float rnd = rand_f32(seed);
    if (rnd * sum_mc < p_hat * (float)kNumResamples) {
        if (near_zero(length3(get_light_dir(state, x1)))) return state;
        state.num_samples = min(state.num_samples + 1u, kChainSampleClamp);
        const float alpha = state.num_samples == 1u
                                ? 1.0f
                                : fmaxf(1.0f / (float)state.num_samples, 0.01f) * confidence;
        state.weight_sum = mix1(state.weight_sum, score, alpha);
        state.ls = mix3(state.ls, score * x2, alpha);
        const V3 mean = normalize3(get_light_dir(state, x1));
        state.mean_cosine = mix1(state.mean_cosine, score * dot3(normalize3(x2 - x1), mean), alpha);
    }
    return state;
}

// general.wesl:1000-1016
__device__ __forceinline__ float get_kappa(const Chain& s) {
    const float np_ = 0.2f, rp = 0.99f;
    const float n2 = (float)(s.num_samples * s.num_samples);
    const float r = ((n2 * (s.mean_cosine / s.weight_sum)) + (np_ * rp)) / (n2 + np_);
    const float k = ((3.0f * r) - ((r * r) * r)) / (1.0f - (r * r));
    return fminf(fmaxf(k, 0.001f), 8.50705867e37f);
}
// general.wesl:1025-1047
__device__ __forceinline__ float exp_minus_one(float x) {
    float u = expf(x);
    if (u == 1.0f) return x;
    float y = u - 1.0f;
    if (fabsf(x) < 1.0f) return (x * y) / logf(u);
    return y;
}
__device__ __forceinline__ float x_div_exp_minus_one(float x) {
    float u = expf(x);
    if (u == 1.0f) return 1.0f;
    float y = u - 1.0f;
    if (fabsf(x) < 1.0f) return logf(u) / y;
    return x / y;
}
// general.wesl:1018-1023
__device__ __forceinline__ float vmf_pdf(V3 mean, float kappa, V3 dir) {
    const float fraction = x_div_exp_minus_one(-2.0f * kappa) / (4.0f * kPi);
    const float c = dot3(dir, mean);
    return fmaxf(fraction * expf(kappa * (c - 1.0f)), 0.0f);
}

// general.wesl:797-808: Jakob (2012) vMF sample about `mean` from the draws
// at seed (azimuth) and seed + 1 (cosine); the unnormalised Xs, whose length
// the caller tests.  The caller advances its seed by 2.
__device__ __forceinline__ V3 vmf_sample(V3 mean, float kappa, uint32_t seed) {
    const float theta = rand_f32(seed) * kTau;
    const float vx = cosf(theta), vy = sinf(theta);
    const float xi = rand_f32(seed + 1u);
    const float m2k = -2.0f * kappa;
    const float wm1 = logf(1.0f + xi * exp_minus_one(m2k)) / kappa;
    const float W = 1.0f + wm1;
    const float norm_v = sqrtf(-fmaf(wm1, wm1, 2.0f * wm1));
    return from_onb_no_tangent(v3(strip_nan(vx * norm_v), W, strip_nan(vy * norm_v)), mean);
}

// One candidate of resample_mcmm (general.wesl:408-421): advances the seed
// twice and the running weight sum; true when the candidate replaces the
// selection.
__device__ __forceinline__ bool resample_step(const Chain& cand, uint32_t& seed, float& sum_mc) {
    seed = rand_u32(seed);
    sum_mc += cand.weight_sum;
    const bool take = rand_f32(seed) * sum_mc < cand.weight_sum;
    seed = rand_u32(seed);
    return take;
}

// One candidate's term of all_vmf_pdfs (general.wesl:989-995).
__device__ __forceinline__ float mixture_term(const Chain& cand, V3 origin, V3 d, float sum_mc) {
    if (!(cand.weight_sum > 0.0f)) return 0.0f;
    return vmf_pdf(normalize3(get_light_dir(cand, origin)), get_kappa(cand), d) *
           (cand.weight_sum / sum_mc);
}

// Screen-space candidates of general.wesl:367-385 drawn from the seed the
// fill starts with: the lane picked 4x (the first loop does not advance the
// seed) and the 12 previous-frame neighbours in a 91 x 91 window.  A
// logical subgroup is `subgroup` consecutive pixels of a row.
struct ScreenCands {
    uint32_t lane_pix;  // pixel whose sample-start tentative chain is shuffled in
    bool lane_ok;
    uint32_t idx[kScreenLastFrame];
    uint32_t seed_after;
};

__device__ __forceinline__ ScreenCands screen_candidates(uint32_t seed, uint32_t x, uint32_t y,
                                                         uint32_t width, uint32_t subgroup) {
    ScreenCands c;
    const uint32_t base = (x / subgroup) * subgroup;
    const uint32_t lane = rand_u32(seed) % subgroup;
    c.lane_pix = y * width + base + lane;
    c.lane_ok = base + lane < width;
#pragma unroll
    for (uint32_t i = 0; i < kScreenLastFrame; ++i) {
        const uint32_t xr = rand_u32(seed);
        seed = xr;
        const uint32_t yr = rand_u32(seed);
        seed = yr;
        const int32_t dx = (int32_t)(xr % kResampleSize) - (int32_t)(kResampleSize / 2);
        const int32_t dy = (int32_t)(yr % kResampleSize) - (int32_t)(kResampleSize / 2);
        const uint32_t nx = (uint32_t)((int32_t)x + dx), ny = (uint32_t)((int32_t)y + dy);
        c.idx[i] = nx + ny * width;  // u32 wrap-around as in the reference
    }
    c.seed_after = seed;
    return c;
}

// A grid key as one 64-bit word (the secondary hash is odd, so never 0).
__device__ __forceinline__ uint64_t key_word(const GridKey& k) {
    return ((uint64_t)k.secondary << 32) | k.idx;
}

// Whether this lane is the lowest active lane of its wave holding key k
// (index and secondary hash).  In the reference's lock-step execution the
// lanes of a wave that write one cell have one lane's stores land and one
// lane take its try-lock; lanes whose keys share an index but not a
// secondary hash are different keys, and each keeps its own writer.
// One pass per distinct key of the wave: the lowest lane not yet settled
// broadcasts its key, which settles every lane holding it, and is that
// key's writer.
__device__ __forceinline__ bool first_writer_of_key(const GridKey& k) {
    bool first = false;
    uint64_t todo = __ballot(1);
    const int me = (int)__lane_id();
    while (todo) {
        const int l = __builtin_ctzll(todo);  // wave-uniform
        const uint32_t li = __shfl(k.idx, l);
        const uint32_t ls = __shfl(k.secondary, l);
        todo &= ~__ballot(li == k.idx && ls == k.secondary);
        first |= me == l;
    }
    return first;
}

}  // namespace lmb
