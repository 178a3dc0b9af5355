// api_nrc.cpp - C ABI of the neural radiance cache (creation, weights, the
// host-array inference / gradient / training calls) and the optimiser-step
// driver shared with lmb_trace_nrc.  The cache has no counterpart in the
// reference; its entry points are declared in include/lambent.h.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "lmb_internal.h"
#include "lmb_restir.h"

using namespace lmb;

namespace lmb {

static uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// He-uniform initialisation: U(-a, a), a = sqrt(6 / fan_in); the constant
// feature's row (layer-0 bias) and the unused rows start at zero.
static void init_weights(uint32_t seed, std::vector<float>& w) {
    w.assign(kNrcNumWeights, 0.0f);
    uint64_t st = 0x1234ABCDull ^ ((uint64_t)seed << 32);
    auto uni = [&](float a) {
        const double u = (double)(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
        return (float)((2.0 * u - 1.0) * a);
    };
    const float a0 = std::sqrt(6.0f / (float)(kNrcEncoded + 1));
    for (int in = 0; in < kNrcEncoded; ++in)
        for (int out = 0; out < kNrcWidth; ++out) w[kNrcW0 + in * kNrcWidth + out] = uni(a0);
    const float ah = std::sqrt(6.0f / (float)kNrcWidth);
    for (int i = 0; i < kNrcHidden * kNrcWidth * kNrcWidth; ++i) w[kNrcWH + i] = uni(ah);
    for (int i = 0; i < kNrcWidth * kNrcOut; ++i) w[kNrcW5 + i] = uni(ah);
}

lmb_status nrc_ensure_batch(lmb_nrc n, uint32_t batch) {
    batch = (batch + kNrcChunk - 1) / kNrcChunk * kNrcChunk;
    if (batch == 0) batch = kNrcChunk;
    if (batch > n->batch_cap) {
        const size_t S = batch;
        lmb_status st = n->x0.alloc(S * kNrcIn * 2);
        if (!st) st = n->h.alloc(S * 5 * kNrcWidth * 2);
        if (!st) st = n->dz.alloc(S * 5 * kNrcWidth * 2);
        if (!st) st = n->g.alloc(S * kNrcOut * 2);
        if (!st) st = n->partial.alloc(S / kNrcChunk * kNrcNumWeights * sizeof(float));
        if (st) {
            n->batch_cap = 0;
            return st;
        }
        n->batch_cap = batch;
        n->dev.x0 = n->x0.as<uint16_t>();
        n->dev.h = n->h.as<uint16_t>();
        n->dev.dz = n->dz.as<uint16_t>();
        n->dev.g = n->g.as<uint16_t>();
        n->dev.partial = n->partial.as<float>();
    }
    n->dev.batch = batch;
    return LMB_OK;
}

lmb_status nrc_ensure_queries(lmb_nrc n, uint64_t queries) {
    if (queries > 0xFFFFFFFFull) return fail(LMB_ERR_INVALID_ARGUMENT, "too many cache queries");
    if (queries <= n->dev.query_cap && n->dev.queries) return LMB_OK;
    lmb_status st = n->queries.alloc(queries * sizeof(NrcQuery));
    if (!st) st = n->query_weight.alloc(queries * sizeof(float4));
    if (st) {
        n->dev.query_cap = 0;
        n->dev.queries = nullptr;
        return st;
    }
    n->dev.queries = n->queries.as<NrcQuery>();
    n->dev.query_weight = n->query_weight.as<float4>();
    n->dev.query_cap = (uint32_t)queries;
    return LMB_OK;
}

static AdamParams next_adam(lmb_nrc n) {
    n->steps += 1;
    n->pow1 *= (double)n->opts.beta1;
    n->pow2 *= (double)n->opts.beta2;
    AdamParams a;
    a.lr = n->opts.learning_rate;
    a.beta1 = n->opts.beta1;
    a.beta2 = n->opts.beta2;
    a.eps = n->opts.epsilon;
    a.bias1 = (float)(1.0 - n->pow1);
    a.bias2 = (float)(1.0 - n->pow2);
    return a;
}

lmb_status nrc_run_steps(lmb_nrc n, const NrcDev& d, uint16_t* publish, const NrcQuery* in,
                         const float4* target, const uint32_t* count, uint32_t count_limit,
                         uint32_t steps, bool apply, float* grad_out, hipStream_t stream,
                         KernelTimer* timer) {
    hipError_t e = launch_fill_zero(d.loss, kNrcMaxSteps * sizeof(float), stream);
    if (e != hipSuccess) return fail_hip(e, "loss clear");
    for (uint32_t s = 0; s < steps; ++s) {
        const AdamParams a = apply ? next_adam(n) : AdamParams{};
        NrcDev ds = d;
        if (publish && s + 1 == steps) ds.fwd_pub = publish;
        if (timer) timer->begin(kKernelCacheTrain, stream);
        e = nrc_launch_step(ds, in, target, count, count_limit, steps, s, a, apply, grad_out,
                            nullptr, stream);
        if (timer) timer->end(kKernelCacheTrain, stream);
        if (e != hipSuccess) return fail_hip(e, "cache training step");
    }
    n->last_steps = steps;
    if (apply) n->inf_valid = false;
    return LMB_OK;
}

lmb_status nrc_sync_frags(lmb_nrc n, hipStream_t stream) {
    if (!n->frags_stale) return LMB_OK;
    hipError_t e = nrc_launch_pack(n->dev, stream);
    if (e != hipSuccess) return fail_hip(e, "cache weight pack");
    n->frags_stale = false;
    return LMB_OK;
}

lmb_status nrc_ensure_overlap(lmb_nrc n, uint64_t npix) {
    hipError_t e = hipSuccess;
    if (!n->train_stream) {
        // non-blocking: no implicit ordering against a caller on the null stream
        e = hipStreamCreateWithFlags(&n->train_stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&n->ev_start, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&n->ev_done, hipEventDisableTiming);
        if (e != hipSuccess) return fail_hip(e, "training stream");
    }
    if (!n->ctl_train.ptr) {
        lmb_status st = n->ctl_train.alloc(kNrcCtlWords * 4);
        if (st) return st;
        e = hipMemset(n->ctl_train.ptr, 0, kNrcCtlWords * 4);
        if (e != hipSuccess) return fail_hip(e, "hipMemset");
    }
    if (!n->inf_sets.ptr) {
        // zeroed: the optimiser's publishing writes weight elements only,
        // the padding elements stay zero
        const size_t bytes = 2 * (size_t)kNrcFwdFrags * kFragElems * 2;
        lmb_status st = n->inf_sets.alloc(bytes);
        if (st) return st;
        e = hipMemset(n->inf_sets.ptr, 0, bytes);
        if (e != hipSuccess) return fail_hip(e, "hipMemset");
        n->inf_valid = false;
    }
    const uint64_t grid = std::max<uint64_t>(n->opts.train_paths, 256);
    if (npix <= n->tf_npix && grid <= n->tf_grid) return LMB_OK;
    // the buffers may still be in use by the previous frame's training pass
    e = hipStreamSynchronize(n->train_stream);
    if (e != hipSuccess) return fail_hip(e, "hipStreamSynchronize");
    const uint64_t np = std::max(npix, n->tf_npix);
    lmb_status st = LMB_OK;
    if (np > n->tf_npix) {
        n->tf_npix = 0;
        if (!st) st = n->tf_seed.alloc(np * 4);
        if (!st) st = n->tf_through.alloc(np * 16);
        if (!st) st = n->tf_accum.alloc(np * 16);
        if (!st) st = n->tf_queue0.alloc(np * 32);
        if (!st) st = n->tf_queue1.alloc(np * 32);
        if (!st) st = n->tf_hits.alloc(np * 16);
        if (!st && !n->tf_counters.ptr) st = n->tf_counters.alloc(kNumCounters * 4);
        if (st) return st;
        n->tf_npix = np;
    }
    st = ensure_spill(&n->tf_spill, extend_threads(std::max(grid, n->tf_grid)));
    if (st) return st;
    n->tf_grid = std::max(grid, n->tf_grid);
    return LMB_OK;
}

lmb_status nrc_ensure_wide(lmb_nrc n, uint64_t paths) {
    if (paths <= n->wide_paths) return LMB_OK;
    n->wide_paths = 0;
    lmb_status st = n->w_seed.alloc(paths * 4);
    if (!st) st = n->w_through.alloc(paths * 16);
    if (!st) st = n->w_accum.alloc(paths * 16);
    if (!st) st = n->w_sum.alloc(paths * 16);
    if (!st) st = n->w_queue0.alloc(paths * 32);
    if (!st) st = n->w_queue1.alloc(paths * 32);
    if (!st) st = n->w_hits.alloc(paths * 16);
    if (st) return st;
    n->wide_paths = paths;
    return LMB_OK;
}

}  // namespace lmb

namespace {

lmb_status check(lmb_nrc n) {
    if (!n) return fail(LMB_ERR_INVALID_ARGUMENT, "cache is NULL");
    hipError_t e = hipSetDevice(n->device);
    if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
    return LMB_OK;
}

// Host rows (n x 13) -> device NrcQuery records in n->io_in.
lmb_status stage_inputs(lmb_nrc n, const float* inputs, uint64_t count) {
    std::vector<NrcQuery> q(count);
    for (uint64_t i = 0; i < count; ++i) {
        std::memset(&q[i], 0, sizeof(NrcQuery));
        std::memcpy(q[i].in, inputs + i * kNrcInputs, sizeof(float) * kNrcInputs);
        q[i].pid = (uint32_t)i;
    }
    if (n->io_in.bytes < count * sizeof(NrcQuery)) {
        lmb_status st = n->io_in.alloc(count * sizeof(NrcQuery));
        if (st) return st;
    }
    hipError_t e = hipMemcpy(n->io_in.ptr, q.data(), count * sizeof(NrcQuery), hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy inputs");
    return LMB_OK;
}

lmb_status stage_targets(lmb_nrc n, const float* targets, uint64_t count) {
    std::vector<float> t(count * 4, 0.0f);
    for (uint64_t i = 0; i < count; ++i)
        for (int k = 0; k < 3; ++k) t[i * 4 + k] = targets[i * 3 + k];
    if (n->io_target.bytes < count * 16) {
        lmb_status st = n->io_target.alloc(count * 16);
        if (st) return st;
    }
    hipError_t e = hipMemcpy(n->io_target.ptr, t.data(), count * 16, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy targets");
    return LMB_OK;
}

// Device u32 holding `count`, the record count of a host-array call.
uint32_t* host_count(lmb_nrc n, uint32_t count) {
    uint32_t* p = n->dev.ctl + kNrcCtlHostCount;
    if (hipMemcpy(p, &count, 4, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return p;
}

float read_loss(lmb_nrc n, uint64_t count, uint32_t steps) {
    float sums[kNrcMaxSteps] = {0};
    const uint32_t k = std::min<uint32_t>(steps, kNrcMaxSteps);
    if (hipMemcpy(sums, n->dev.loss, sizeof(float) * k, hipMemcpyDeviceToHost) != hipSuccess)
        return NAN;
    double tot = 0.0;
    for (uint32_t s = 0; s < k; ++s) tot += sums[s];
    return count ? (float)(tot / (3.0 * (double)count)) : 0.0f;
}

lmb_status train_or_grad(lmb_nrc n, const float* inputs, const float* targets, uint64_t count,
                         bool apply, float* grad, float* loss) {
    lmb_status st = check(n);
    if (st) return st;
    if (count == 0 || !inputs || !targets) return fail(LMB_ERR_INVALID_ARGUMENT, "no samples");
    if (count > (1u << 24)) return fail(LMB_ERR_INVALID_ARGUMENT, "at most 2^24 samples per call");
    st = nrc_ensure_batch(n, (uint32_t)count);
    if (!st) st = stage_inputs(n, inputs, count);
    if (!st) st = stage_targets(n, targets, count);
    if (st) return st;
    uint32_t* cnt = host_count(n, (uint32_t)count);
    if (!cnt) return fail(LMB_ERR_DEVICE, "count upload failed");
    float* gdev = nullptr;
    if (grad) {
        if (n->io_grad.bytes < kNrcNumWeights * sizeof(float)) {
            st = n->io_grad.alloc(kNrcNumWeights * sizeof(float));
            if (st) return st;
        }
        gdev = n->io_grad.as<float>();
    }
    st = nrc_run_steps(n, n->dev, nullptr, n->io_in.as<NrcQuery>(), n->io_target.as<float4>(), cnt,
                       (uint32_t)count, 1, apply, gdev, nullptr, nullptr);
    if (st) return st;
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail_hip(e, "cache step");
    if (grad) {
        e = hipMemcpy(grad, gdev, kNrcNumWeights * sizeof(float), hipMemcpyDeviceToHost);
        if (e != hipSuccess) return fail_hip(e, "hipMemcpy gradient");
    }
    if (loss) *loss = read_loss(n, count, 1);
    return LMB_OK;
}

}  // namespace

extern "C" {

void lmb_nrc_opts_default(lmb_nrc_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof *o);
    o->learning_rate = 1e-3f;
    o->beta1 = 0.9f;
    o->beta2 = 0.99f;
    o->epsilon = 1e-8f;
    o->train_paths = 65536;
    o->train_steps = 4;
    o->seed = 1;
    o->ema_decay = 0.0f;
    o->defer_specular = 2u;
    o->overlap_training = 1u;
    o->stratify = 1u;
}

static lmb_status validate(const lmb_nrc_opts& o) {
    if (!(o.learning_rate > 0.0f) || !(o.beta1 >= 0.0f && o.beta1 < 1.0f) ||
        !(o.beta2 >= 0.0f && o.beta2 < 1.0f) || !(o.epsilon >= 0.0f))
        return fail(LMB_ERR_INVALID_ARGUMENT, "invalid Adam parameters");
    if (o.train_steps == 0 || o.train_steps > kNrcMaxSteps)
        return fail(LMB_ERR_INVALID_ARGUMENT, "train_steps must be in 1..256");
    if (o.train_paths > (1u << 24))
        return fail(LMB_ERR_INVALID_ARGUMENT, "train_paths must be <= 2^24");
    if (!(o.ema_decay >= 0.0f && o.ema_decay < 1.0f))
        return fail(LMB_ERR_INVALID_ARGUMENT, "ema_decay must be in [0, 1)");
    if (o.defer_specular > kMaxBounces)
        return fail(LMB_ERR_INVALID_ARGUMENT, "defer_specular must be at most 64");
    if (o.overlap_training > 1u)
        return fail(LMB_ERR_INVALID_ARGUMENT, "overlap_training must be 0 or 1");
    if (o.stratify > 1u) return fail(LMB_ERR_INVALID_ARGUMENT, "stratify must be 0 or 1");
    return LMB_OK;
}

// Record buffers for train_paths plus a margin for the random selection.
static lmb_status ensure_records(lmb_nrc n, uint32_t train_paths) {
    const uint64_t cap = std::max<uint64_t>(kNrcChunk, (uint64_t)train_paths * 5 / 4);
    if (cap <= n->dev.rec_cap && n->dev.rec_in) return LMB_OK;
    lmb_status st = n->rec_in.alloc(cap * sizeof(NrcQuery));
    if (!st) st = n->rec_aux.alloc(cap * 2 * sizeof(float4));
    if (!st) st = n->rec_target.alloc(cap * sizeof(float4));
    if (st) {
        n->dev.rec_cap = 0;
        n->dev.rec_in = nullptr;
        return st;
    }
    n->dev.rec_in = n->rec_in.as<NrcQuery>();
    n->dev.rec_aux = n->rec_aux.as<float4>();
    n->dev.rec_target = n->rec_target.as<float4>();
    n->dev.rec_cap = (uint32_t)cap;
    return LMB_OK;
}

// Inference weights: the EMA copy and its fragments when ema_decay > 0.
static lmb_status ensure_ema(lmb_nrc n) {
    NrcDev& d = n->dev;
    d.ema_decay = n->opts.ema_decay;
    n->inf_valid = false;
    n->frags_stale = false;
    if (n->opts.ema_decay > 0.0f) {
        if (!n->ema.ptr) {
            lmb_status st = n->ema.alloc(kNrcNumWeights * sizeof(float));
            if (!st) st = n->fwd_inf.alloc(kNrcFwdFrags * kFragElems * 2);
            if (st) return st;
            hipError_t e = hipMemcpy(n->ema.ptr, d.master, kNrcNumWeights * sizeof(float),
                                     hipMemcpyDeviceToDevice);
            if (e != hipSuccess) return fail_hip(e, "hipMemcpy weights");
        }
        d.ema = n->ema.as<float>();
        d.fwd_inf = d.fwd_pub = n->fwd_inf.as<uint16_t>();
    } else {
        d.ema = nullptr;
        d.fwd_inf = d.fwd;
        d.fwd_pub = nullptr;
    }
    hipError_t e = nrc_launch_pack(d, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail_hip(e, "cache weight pack");
    return LMB_OK;
}

lmb_status lmb_nrc_create(const lmb_nrc_opts* opts, lmb_nrc* out) {
    if (!out) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    lmb_nrc_opts o;
    lmb_nrc_opts_default(&o);
    if (opts) o = *opts;
    lmb_status vs = validate(o);
    if (vs) return vs;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return fail(LMB_ERR_DEVICE, "no HIP device available");
    }
    lmb_nrc n = new lmb_nrc_s();
    n->opts = o;
    (void)hipGetDevice(&n->device);
    const size_t wbytes = kNrcNumWeights * sizeof(float);
    lmb_status st = n->master.alloc(wbytes);
    if (!st) st = n->adam_m.alloc(wbytes);
    if (!st) st = n->adam_v.alloc(wbytes);
    if (!st) st = n->fwd.alloc(kNrcFwdFrags * kFragElems * 2);
    if (!st) st = n->bwd.alloc(kNrcBwdFrags * kFragElems * 2);
    if (!st) st = n->ctl.alloc(kNrcCtlWords * 4);
    if (!st) st = n->loss.alloc(kNrcMaxSteps * sizeof(float));
    if (!st) {  // the fragment element of every weight, for the Adam kernel's re-pack
        std::vector<int32_t> fpos(kNrcNumWeights, -1), bpos(kNrcNumWeights, -1);
        for (int f = 0; f < kNrcFwdFrags; ++f)
            for (int l = 0; l < 64; ++l)
                for (int e = 0; e < 8; ++e) {
                    const int w = nrc_fwd_index(f, l, e);
                    if (w >= 0) fpos[w] = f * kFragElems + l * 8 + e;
                }
        for (int f = 0; f < kNrcBwdFrags; ++f)
            for (int l = 0; l < 64; ++l)
                for (int e = 0; e < 8; ++e) {
                    const int w = nrc_bwd_index(f, l, e);
                    if (w >= 0) bpos[w] = f * kFragElems + l * 8 + e;
                }
        for (int w = 0; w < kNrcNumWeights; ++w)
            if (fpos[w] < 0) st = fail(LMB_ERR_DEVICE, "weight without a forward fragment element");
        if (!st) st = n->fwd_pos.upload(fpos.data(), fpos.size() * 4);
        if (!st) st = n->bwd_pos.upload(bpos.data(), bpos.size() * 4);
    }
    if (!st) st = ensure_records(n, o.train_paths);
    if (st) {
        delete n;
        return st;
    }
    NrcDev& d = n->dev;
    d.master = n->master.as<float>();
    d.adam_m = n->adam_m.as<float>();
    d.adam_v = n->adam_v.as<float>();
    d.fwd = n->fwd.as<uint16_t>();
    d.bwd = n->bwd.as<uint16_t>();
    d.ctl = n->ctl.as<uint32_t>();
    d.loss = n->loss.as<float>();
    d.fwd_pos = n->fwd_pos.as<int32_t>();
    d.bwd_pos = n->bwd_pos.as<int32_t>();
    d.fwd_inf = d.fwd;
    hipError_t e = hipMemset(n->ctl.ptr, 0, kNrcCtlWords * 4);
    if (e != hipSuccess) {
        delete n;
        return fail_hip(e, "hipMemset");
    }
    std::vector<float> w;
    init_weights(o.seed, w);
    st = lmb_nrc_set_weights(n, w.data());
    if (!st) st = ensure_ema(n);
    if (st) {
        delete n;
        return st;
    }
    *out = n;
    return LMB_OK;
}

lmb_status lmb_nrc_set_opts(lmb_nrc n, const lmb_nrc_opts* opts) {
    lmb_status st = check(n);
    if (st) return st;
    if (!opts) return fail(LMB_ERR_INVALID_ARGUMENT, "opts is NULL");
    st = validate(*opts);
    if (st) return st;
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail_hip(e, "hipDeviceSynchronize");
    st = ensure_records(n, opts->train_paths);
    if (st) return st;
    const uint32_t seed = n->opts.seed;
    const bool beta_changed = opts->beta1 != n->opts.beta1 || opts->beta2 != n->opts.beta2;
    n->opts = *opts;
    n->opts.seed = seed;
    if (beta_changed) {  // the bias corrections restart with the new decays
        n->pow1 = std::pow((double)n->opts.beta1, (double)n->steps);
        n->pow2 = std::pow((double)n->opts.beta2, (double)n->steps);
    }
    return ensure_ema(n);
}

lmb_status lmb_nrc_destroy(lmb_nrc n) {
    if (!n) return LMB_OK;
    (void)hipSetDevice(n->device);
    (void)hipDeviceSynchronize();
    if (n->ev_start) (void)hipEventDestroy(n->ev_start);
    if (n->ev_done) (void)hipEventDestroy(n->ev_done);
    if (n->train_stream) (void)hipStreamDestroy(n->train_stream);
    delete n;
    return LMB_OK;
}

lmb_status lmb_nrc_get_weights(lmb_nrc n, float* w) {
    lmb_status st = check(n);
    if (st) return st;
    if (!w) return fail(LMB_ERR_INVALID_ARGUMENT, "weights is NULL");
    hipError_t e = hipMemcpy(w, n->dev.master, kNrcNumWeights * sizeof(float), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy weights");
    return LMB_OK;
}

lmb_status lmb_nrc_set_weights(lmb_nrc n, const float* w) {
    lmb_status st = check(n);
    if (st) return st;
    if (!w) return fail(LMB_ERR_INVALID_ARGUMENT, "weights is NULL");
    for (uint64_t i = 0; i < kNrcNumWeights; ++i)
        if (!std::isfinite(w[i]))
            return fail(LMB_ERR_INVALID_ARGUMENT, "weight " + std::to_string(i) + " is not finite");
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess)
        e = hipMemcpy(n->dev.master, w, kNrcNumWeights * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(n->dev.adam_m, 0, kNrcNumWeights * sizeof(float));
    if (e == hipSuccess) e = hipMemset(n->dev.adam_v, 0, kNrcNumWeights * sizeof(float));
    if (e == hipSuccess && n->dev.ema)
        e = hipMemcpy(n->dev.ema, w, kNrcNumWeights * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = nrc_launch_pack(n->dev, nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail_hip(e, "cache weight upload");
    n->steps = 0;
    n->pow1 = n->pow2 = 1.0;
    n->inf_valid = false;
    n->frags_stale = false;
    return LMB_OK;
}

lmb_status lmb_nrc_infer(lmb_nrc n, const float* inputs, uint64_t count, float* outputs) {
    lmb_status st = check(n);
    if (st) return st;
    if (count == 0) return LMB_OK;
    if (!inputs || !outputs) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    if (count > 0x7FFFFFFFull) return fail(LMB_ERR_INVALID_ARGUMENT, "too many rows");
    st = stage_inputs(n, inputs, count);
    if (!st) st = nrc_sync_frags(n, nullptr);
    if (st) return st;
    if (n->io_out.bytes < count * 16) {
        st = n->io_out.alloc(count * 16);
        if (st) return st;
    }
    hipError_t e = nrc_launch_infer(n->dev, n->io_in.as<NrcQuery>(), (uint32_t)count,
                                    n->io_out.as<float4>(), nullptr);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail_hip(e, "cache inference");
    std::vector<float> o(count * 4);
    e = hipMemcpy(o.data(), n->io_out.ptr, count * 16, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy outputs");
    for (uint64_t i = 0; i < count; ++i)
        for (int k = 0; k < 3; ++k) outputs[i * 3 + k] = o[i * 4 + k];
    return LMB_OK;
}

lmb_status lmb_nrc_gradient(lmb_nrc n, const float* inputs, const float* targets, uint64_t count,
                            float* grad, float* loss) {
    if (!grad) return fail(LMB_ERR_INVALID_ARGUMENT, "grad is NULL");
    return train_or_grad(n, inputs, targets, count, false, grad, loss);
}

lmb_status lmb_nrc_train(lmb_nrc n, const float* inputs, const float* targets, uint64_t count,
                         float* loss) {
    return train_or_grad(n, inputs, targets, count, true, nullptr, loss);
}

lmb_status lmb_nrc_apply_gradient(lmb_nrc n, const float* grad) {
    lmb_status st = check(n);
    if (st) return st;
    if (!grad) return fail(LMB_ERR_INVALID_ARGUMENT, "grad is NULL");
    if (n->io_grad.bytes < kNrcNumWeights * sizeof(float)) {
        st = n->io_grad.alloc(kNrcNumWeights * sizeof(float));
        if (st) return st;
    }
    hipError_t e = hipMemcpy(n->io_grad.ptr, grad, kNrcNumWeights * sizeof(float),
                             hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy gradient");
    const AdamParams a = next_adam(n);
    e = nrc_launch_step(n->dev, nullptr, nullptr, nullptr, 0, 1, 0, a, true, nullptr,
                        n->io_grad.as<float>(), nullptr);
    n->inf_valid = false;
    n->frags_stale = false;  // the step re-packed every fragment set of n->dev
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail_hip(e, "cache Adam step");
    return LMB_OK;
}

lmb_status lmb_nrc_stats_get(lmb_nrc n, lmb_nrc_stats* s) {
    lmb_status st = check(n);
    if (st) return st;
    if (!s) return fail(LMB_ERR_INVALID_ARGUMENT, "stats is NULL");
    hipError_t e = hipDeviceSynchronize();
    uint32_t ctl[kNrcCtlWords] = {0};
    if (e == hipSuccess) e = hipMemcpy(ctl, n->dev.ctl, sizeof ctl, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy cache counters");
    std::memset(s, 0, sizeof *s);
    s->steps = n->steps;
    uint32_t records = ctl[kNrcCtlRecords];
    if (n->rec_in_train_ctl && e == hipSuccess)
        e = hipMemcpy(&records, (uint32_t*)n->ctl_train.ptr + kNrcCtlRecords, 4, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy cache counters");
    s->records = std::min<uint32_t>(records, n->dev.rec_cap);
    s->queries = ctl[kNrcCtlQueries];
    s->loss = read_loss(n, s->records, n->last_steps);
    return LMB_OK;
}

}  // extern "C"
