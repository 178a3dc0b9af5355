// api_frame.cpp - C ABI: tracer options, per-frame buffers (DataBuffers) and
// the frame passes.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/lambent.h"
#include "lmb_frame_dev.h"
#include "lmb_internal.h"
#include "lmb_guided.h"
#include "lmb_restir.h"

using namespace lmb;

struct lmb_tracer_s {
    lmb_tracer_opts opts;
};

namespace {

// Records a start/stop event pair around each timed launch; events are
// pooled and resolved when the stats are read.
class EventTimer final : public KernelTimer {
  public:
    ~EventTimer() override {
        for (auto& s : pool_) {
            (void)hipEventDestroy(s.start);
            (void)hipEventDestroy(s.stop);
        }
    }
    void begin(int kind, hipStream_t stream) override {
        if (used_ == pool_.size()) {
            Slot s{};
            // Timing only: no system-scope fence (cache writeback and
            // invalidation) when an event completes, which would also slow
            // the launches after it.
            if (hipEventCreateWithFlags(&s.start, hipEventDisableSystemFence) != hipSuccess ||
                hipEventCreateWithFlags(&s.stop, hipEventDisableSystemFence) != hipSuccess) {
                failed_ = true;
                return;
            }
            pool_.push_back(s);
        }
        pool_[used_].kind = kind;
        if (hipEventRecord(pool_[used_].start, stream) != hipSuccess) failed_ = true;
    }
    void end(int kind, hipStream_t stream) override {
        (void)kind;
        if (used_ >= pool_.size()) return;
        if (hipEventRecord(pool_[used_].stop, stream) != hipSuccess) failed_ = true;
        ++used_;
    }
    // Sums the recorded durations per kind and resets.
    bool collect(double ms[kNumKernelKinds], uint64_t launches[kNumKernelKinds]) {
        for (int k = 0; k < kNumKernelKinds; ++k) {
            ms[k] = 0.0;
            launches[k] = 0;
        }
        bool ok = !failed_;
        for (size_t i = 0; i < used_; ++i) {
            float t = 0.f;
            if (hipEventSynchronize(pool_[i].stop) != hipSuccess ||
                hipEventElapsedTime(&t, pool_[i].start, pool_[i].stop) != hipSuccess) {
                ok = false;
                continue;
            }
            ms[pool_[i].kind] += t;
            launches[pool_[i].kind] += 1;
        }
        used_ = 0;
        failed_ = false;
        return ok;
    }

  private:
    struct Slot {
        hipEvent_t start, stop;
        int kind;
    };
    std::vector<Slot> pool_;
    size_t used_ = 0;
    bool failed_ = false;
};

}  // namespace

struct lmb_data_buffers_s {
    std::unique_ptr<EventTimer> timer;
    int device = 0;
    uint32_t width = 0, height = 0, world_size = 0, flags = 0;
    DeviceBuffer out_color, out_normal, out_albedo;
    DeviceBuffer own_seed, pix_sum, ps_seed, ps_through, ps_accum;
    DeviceBuffer queue0, queue1, hits, counters;
    DeviceBuffer hits_primary;  // allocated by the first trace of several samples
    // Plain traces of several samples on a small image trace samples
    // together: per-path state for grp_paths paths (allocated on first use).
    DeviceBuffer grp_seed, grp_through, grp_accum, grp_sum, grp_queue0, grp_queue1, grp_hits;
    SpillBuffer grp_spill;
    uint64_t grp_paths = 0;
    DeviceBuffer gi, gi_old, info, mc, mc_old, wmc, wmc_old;
    SpillBuffer spill;
    uint64_t gi_count = 0, mc_count = 0, wmc_count = 0;
    FrameDev dev{};
    DeviceBuffer guided_pool;  // guided-mode state, allocated on first use
    GuidedDev guided{};
};

namespace {

// Carves the guided-mode arrays out of one allocation (256-byte aligned).
lmb_status ensure_guided(lmb_data_buffers b) {
    const uint64_t n = (uint64_t)b->width * b->height;
    struct Part {
        void** dst;
        uint64_t bytes;
    };
    GuidedDev& g = b->guided;
    Part parts[] = {
        {(void**)&g.tent, n * sizeof(MarkovChain)},
        {(void**)&g.tent_snap, n * sizeof(MarkovChain)},
        {(void**)&g.sum_mc, n * 4},
        {(void**)&g.cand_seed, n * 4},
        {(void**)&g.pix_res, n * sizeof(Reservoir)},
        {(void**)&g.pix_sample, n * sizeof(Reservoir)},
        {(void**)&g.out_rad_sum, n * 16},
        {(void**)&g.pix_last_normal, n * 16},
        {(void**)&g.pix_last_albedo, n * 16},
        {(void**)&g.pix_last_emission, n * 16},
        {(void**)&g.color, n * 16},
        {(void**)&g.scol, n * 16},
        {(void**)&g.orc, n * 16},
        {(void**)&g.sout, n * 16},
        {(void**)&g.aem, n * 16},
        {(void**)&g.aalb, n * 16},
        {(void**)&g.lastn, n * 16},
        {(void**)&g.lastd, n * 16},
        {(void**)&g.s_point, n * 16},
        {(void**)&g.s_normal, n * 16},
        {(void**)&g.s_albedo, n * 16},
        {(void**)&g.s_emission, n * 16},
        {(void**)&g.s_vis_point, n * 16},
        {(void**)&g.s_vis_normal, n * 16},
        {(void**)&g.s_start_dir, n * 16},
        {(void**)&g.gr0, n * 16},
        {(void**)&g.gr1, n * 16},
        {(void**)&g.gr2, n * 16},
        {(void**)&g.gr3, n * 16},
    };
    uint64_t total = 0;
    for (const Part& p : parts) total += (p.bytes + 255) & ~255ull;
    if (b->guided_pool.ptr && b->guided_pool.bytes == total) return LMB_OK;
    lmb_status st = b->guided_pool.alloc(total);
    if (st) return st;
    hipError_t e = hipMemset(b->guided_pool.ptr, 0, total);
    if (e != hipSuccess) return fail_hip(e, "hipMemset");
    char* base = (char*)b->guided_pool.ptr;
    for (const Part& p : parts) {
        *p.dst = base;
        base += (p.bytes + 255) & ~255ull;
    }
    return LMB_OK;
}

}  // namespace

namespace {

lmb_status alloc_zero(DeviceBuffer& b, size_t bytes) {
    lmb_status st = b.alloc(bytes);
    if (st) return st;
    hipError_t e = hipMemset(b.ptr, 0, bytes);
    if (e != hipSuccess) return fail_hip(e, "hipMemset");
    return LMB_OK;
}

void swap_buffers(DeviceBuffer& a, DeviceBuffer& b) {
    std::swap(a.ptr, b.ptr);
    std::swap(a.bytes, b.bytes);
}

// data_buffer/mod.rs:116-148 create_window_required_buffers (+ the output
// images and wavefront queues that the caller owns in the reference).  The
// whole window is allocated before any of it replaces the current one: a
// refused or failed resize (out of memory part way) leaves every buffer and
// the size as they were.
lmb_status alloc_window(lmb_data_buffers b, uint32_t width, uint32_t height) {
    width = width ? width : 1;
    height = height ? height : 1;
    const uint64_t n = (uint64_t)width * height;
    // Ray queues carry the pixel index as a u32 and the kernels split it
    // into (x, y) with 32-bit arithmetic.
    if (n >= (1ull << 31)) return fail(LMB_ERR_INVALID_ARGUMENT, "width * height must be below 2^31");
    const uint64_t gi_count = (b->flags & LMB_BUFFER_SPATIAL_RESAMPLING) ? n : 1;
    const uint64_t mc_count = (b->flags & LMB_BUFFER_MARKOV_CHAIN_SCREEN_SPACE) ? n : 1;
    struct Part {
        DeviceBuffer* cur;
        uint64_t bytes;
    };
    const Part parts[] = {
        {&b->out_color, n * 16}, {&b->out_normal, n * 16}, {&b->out_albedo, n * 16},
        {&b->own_seed, n * 4},   {&b->pix_sum, n * 16},    {&b->ps_seed, n * 4},
        {&b->ps_through, n * 16}, {&b->ps_accum, n * 16},  {&b->queue0, n * 32},
        {&b->queue1, n * 32},    {&b->hits, n * 16},       {&b->counters, kNumCounters * 4},
        {&b->gi, gi_count * sizeof(Reservoir)},
        {&b->gi_old, gi_count * sizeof(Reservoir)},
        {&b->info, n * sizeof(CompressedInfo)},
        {&b->mc, mc_count * sizeof(MarkovChain)},
        {&b->mc_old, mc_count * sizeof(MarkovChain)},
    };
    constexpr size_t kParts = sizeof(parts) / sizeof(parts[0]);
    DeviceBuffer fresh[kParts];  // freed on return: the new window on failure, the old on success
    SpillBuffer spill;
    lmb_status st;
    for (size_t i = 0; i < kParts; ++i)
        if ((st = alloc_zero(fresh[i], parts[i].bytes))) return st;
    if ((st = ensure_spill(&spill, extend_threads(n)))) return st;
    for (size_t i = 0; i < kParts; ++i) swap_buffers(*parts[i].cur, fresh[i]);
    swap_buffers(b->spill.buf, spill.buf);
    std::swap(b->spill.threads, spill.threads);
    std::swap(b->spill.ptr, spill.ptr);
    b->hits_primary.release();
    for (DeviceBuffer* d : {&b->grp_seed, &b->grp_through, &b->grp_accum, &b->grp_sum,
                            &b->grp_queue0, &b->grp_queue1, &b->grp_hits})
        d->release();
    b->grp_paths = 0;
    b->gi_count = gi_count;
    b->mc_count = mc_count;
    b->width = width;
    b->height = height;
    return LMB_OK;
}

void refresh_view(lmb_data_buffers b) {
    FrameDev& f = b->dev;
    f.width = b->width;
    f.height = b->height;
    f.npix = (uint64_t)b->width * b->height;
    f.out_color = b->out_color.as<float4>();
    f.out_normal = b->out_normal.as<float4>();
    f.out_albedo = b->out_albedo.as<float4>();
    f.own_seed = b->own_seed.as<uint32_t>();
    f.pix_sum = b->pix_sum.as<float4>();
    f.ps_seed = b->ps_seed.as<uint32_t>();
    f.ps_through = b->ps_through.as<float4>();
    f.ps_accum = b->ps_accum.as<float4>();
    f.queue[0] = b->queue0.as<float4>();
    f.queue[1] = b->queue1.as<float4>();
    f.hits = b->hits.as<uint4>();
    f.hits_primary = b->hits_primary.as<uint4>();
    f.counters = b->counters.as<uint32_t>();
    f.spill = b->spill.ptr;
    f.gi = b->gi.as<Reservoir>();
    f.gi_old = b->gi_old.as<Reservoir>();
    f.info = b->info.as<CompressedInfo>();
    f.mc = b->mc.as<MarkovChain>();
    f.mc_old = b->mc_old.as<MarkovChain>();
    f.wmc = b->wmc.as<WorldMarkovChain>();
    f.wmc_old = b->wmc_old.as<WorldMarkovChain>();
    f.world_size = (uint32_t)b->wmc_count;
    f.gi_count = b->gi_count;
    f.mc_count = b->mc_count;
    f.flags = b->flags;
}

DeviceBuffer* state_buffer(lmb_data_buffers b, uint32_t which) {
    switch (which) {
        case 0: return &b->gi;
        case 1: return &b->gi_old;
        case 2: return &b->info;
        case 3: return &b->mc;
        case 4: return &b->mc_old;
        case 5: return &b->wmc;
        case 6: return &b->wmc_old;
        default: return nullptr;
    }
}

}  // namespace

extern "C" {

void lmb_tracer_opts_default(lmb_tracer_opts* o) {
    if (!o) return;
    o->quality = LMB_QUALITY_MEDIUM;
    o->samples = 4;      // low_level.rs:101
    o->t_min = 0.0001f;  // low_level.rs:102
    o->t_max = 1000.0f;  // low_level.rs:103
    o->restir = LMB_RESTIR_PER_PIXEL;  // path_tracing/mod.rs:32
    o->bounces = 5;      // general.wesl:50
    o->rr_start = 0xFFFFFFFFu;
    o->guiding = LMB_GUIDING_MARKOV_CHAIN;
}

lmb_status lmb_tracer_create(const lmb_tracer_opts* opts, lmb_tracer* out) {
    if (!opts || !out) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    const bool debug_shader = opts->quality >= LMB_QUALITY_DEBUG_FRONT_FACE &&
                              opts->quality <= LMB_QUALITY_DEBUG_TANGENT;
    if (opts->quality > LMB_QUALITY_SPECTRAL && !debug_shader)
        return fail(LMB_ERR_INVALID_ARGUMENT, "bad quality");
    if (opts->samples == 0)  // NonZeroU32 in the reference
        return fail(LMB_ERR_INVALID_ARGUMENT, "samples must be non-zero");
    if (opts->bounces == 0 || opts->bounces > kMaxBounces)
        return fail(LMB_ERR_INVALID_ARGUMENT,
                    "bounces must be in 1.." + std::to_string(kMaxBounces));
    if (opts->restir > LMB_RESTIR_PER_PIXEL) return fail(LMB_ERR_INVALID_ARGUMENT, "bad restir");
    if (opts->guiding > LMB_GUIDING_MARKOV_CHAIN)
        return fail(LMB_ERR_INVALID_ARGUMENT, "bad guiding mode");
    if (!(opts->t_min >= 0.0f) || !(opts->t_max > opts->t_min))
        return fail(LMB_ERR_INVALID_ARGUMENT, "need 0 <= t_min < t_max");
    lmb_tracer t = new lmb_tracer_s();
    t->opts = *opts;
    *out = t;
    return LMB_OK;
}

lmb_status lmb_tracer_destroy(lmb_tracer tracer) {
    delete tracer;
    return LMB_OK;
}

lmb_status lmb_data_buffers_create(uint32_t width, uint32_t height, uint32_t world_size,
                                   uint32_t buffer_flags, lmb_data_buffers* out) {
    if (!out) return fail(LMB_ERR_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (buffer_flags & ~LMB_BUFFER_ALL) return fail(LMB_ERR_INVALID_ARGUMENT, "bad buffer flags");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
        (void)hipGetLastError();
        return fail(LMB_ERR_DEVICE, "no HIP device available");
    }
    lmb_data_buffers b = new lmb_data_buffers_s();
    (void)hipGetDevice(&b->device);
    b->flags = buffer_flags;
    b->world_size = world_size;
    // data_buffer/mod.rs:162-167: world buffers hold max(world_size, 1)
    // entries when requested, else 1.
    b->wmc_count = (buffer_flags & LMB_BUFFER_MARKOV_CHAIN_WORLD_SPACE)
                       ? (world_size ? world_size : 1)
                       : 1;
    lmb_status st = alloc_zero(b->wmc, b->wmc_count * sizeof(WorldMarkovChain));
    if (!st) st = alloc_zero(b->wmc_old, b->wmc_count * sizeof(WorldMarkovChain));
    if (!st) st = alloc_window(b, width, height);
    if (st) {
        delete b;
        return st;
    }
    refresh_view(b);
    *out = b;
    return LMB_OK;
}

lmb_status lmb_data_buffers_resize(lmb_data_buffers b, uint32_t width, uint32_t height) {
    if (!b) return fail(LMB_ERR_INVALID_ARGUMENT, "buffers is NULL");
    lmb_status st = alloc_window(b, width, height);
    refresh_view(b);
    return st;
}

lmb_status lmb_data_buffers_destroy(lmb_data_buffers b) {
    delete b;
    return LMB_OK;
}

lmb_status lmb_data_buffers_output(lmb_data_buffers b, uint32_t which, void** dev_ptr,
                                   uint64_t* bytes) {
    if (!b || !dev_ptr) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    DeviceBuffer* buf = which == LMB_OUTPUT_COLOR    ? &b->out_color
                        : which == LMB_OUTPUT_NORMAL ? &b->out_normal
                        : which == LMB_OUTPUT_ALBEDO ? &b->out_albedo
                                                     : nullptr;
    if (!buf) return fail(LMB_ERR_INVALID_ARGUMENT, "bad output index");
    *dev_ptr = buf->ptr;
    if (bytes) *bytes = buf->bytes;
    return LMB_OK;
}

lmb_status lmb_data_buffers_read_output(lmb_data_buffers b, uint32_t which, float* host) {
    void* ptr = nullptr;
    uint64_t bytes = 0;
    lmb_status st = lmb_data_buffers_output(b, which, &ptr, &bytes);
    if (st) return st;
    if (!host) return fail(LMB_ERR_INVALID_ARGUMENT, "host is NULL");
    hipError_t e = hipMemcpy(host, ptr, bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy output");
    return LMB_OK;
}

lmb_status lmb_data_buffers_state_size(lmb_data_buffers b, uint32_t which, uint64_t* bytes) {
    if (!b || !bytes) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    DeviceBuffer* buf = state_buffer(b, which);
    if (!buf) return fail(LMB_ERR_INVALID_ARGUMENT, "bad state index");
    *bytes = buf->bytes;
    return LMB_OK;
}

lmb_status lmb_data_buffers_read_state(lmb_data_buffers b, uint32_t which, void* host,
                                       uint64_t bytes) {
    if (!b || !host) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    DeviceBuffer* buf = state_buffer(b, which);
    if (!buf) return fail(LMB_ERR_INVALID_ARGUMENT, "bad state index");
    if (bytes != buf->bytes) return fail(LMB_ERR_INVALID_ARGUMENT, "size mismatch");
    hipError_t e = hipMemcpy(host, buf->ptr, bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy state");
    return LMB_OK;
}

lmb_status lmb_data_buffers_write_state(lmb_data_buffers b, uint32_t which, const void* host,
                                        uint64_t bytes) {
    if (!b || !host) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    DeviceBuffer* buf = state_buffer(b, which);
    if (!buf) return fail(LMB_ERR_INVALID_ARGUMENT, "bad state index");
    if (bytes != buf->bytes) return fail(LMB_ERR_INVALID_ARGUMENT, "size mismatch");
    hipError_t e = hipMemcpy(buf->ptr, host, bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy state");
    return LMB_OK;
}

lmb_status lmb_data_buffers_path_vertices(lmb_data_buffers b, uint64_t* vertices) {
    if (!b || !vertices) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    hipError_t e = hipMemcpy(vertices, b->dev.counters + kCounterVertices, 8,
                             hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy counter");
    return LMB_OK;
}

lmb_status lmb_data_buffers_set_profiling(lmb_data_buffers b, int32_t enable) {
    if (!b) return fail(LMB_ERR_INVALID_ARGUMENT, "buffers is NULL");
    if (enable && !b->timer) {
        b->timer.reset(new EventTimer());
        hipError_t e = hipMemset(b->dev.counters + kCounterVerticesTotal, 0, 8);
        if (e != hipSuccess) return fail_hip(e, "hipMemset");
    } else if (!enable) {
        b->timer.reset();
    }
    return LMB_OK;
}

lmb_status lmb_data_buffers_kernel_stats(lmb_data_buffers b, lmb_kernel_stats* stats) {
    if (!b || !stats) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    if (!b->timer) return fail(LMB_ERR_INVALID_ARGUMENT, "profiling is not enabled");
    double ms[kNumKernelKinds];
    uint64_t n[kNumKernelKinds];
    bool ok = b->timer->collect(ms, n);
    uint64_t verts = 0;
    hipError_t e = hipMemcpy(&verts, b->dev.counters + kCounterVerticesTotal, 8,
                             hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy counter");
    e = hipMemset(b->dev.counters + kCounterVerticesTotal, 0, 8);
    if (e != hipSuccess) return fail_hip(e, "hipMemset");
    if (!ok) return fail(LMB_ERR_DEVICE, "a timing event failed");
    stats->extend_ms = ms[kKernelExtend];
    stats->shade_ms = ms[kKernelShade];
    stats->extend_launches = n[kKernelExtend];
    stats->shade_launches = n[kKernelShade];
    stats->path_vertices = verts;
    stats->cache_infer_ms = ms[kKernelCacheInfer];
    stats->cache_train_ms = ms[kKernelCacheTrain];
    stats->cache_infer_launches = n[kKernelCacheInfer];
    stats->cache_train_steps = n[kKernelCacheTrain];
    return LMB_OK;
}

// Samples the plain tracer traces together on a frame of npix pixels: as
// many as keep the paths of a group at or under a target (2^24 by default,
// 132 B of path state each; the environment variable LMB_SAMPLE_GROUP_PATHS
// overrides it, 0 = one sample at a time).  Measured on the 1M-triangle
// scene at 1080p, 16 spp, 8 bounces: 3.61 ms per sample one at a time, 2.67
// / 2.20 / 2.01 ms in groups of 2 / 4 / 8 (the traversal's per-launch tail
// and launch costs are shared by more rays); a 256 x 256 frame fills the
// GPU only with 16 samples in flight.
static uint32_t sample_group(uint64_t npix, uint32_t samples) {
    uint64_t target = 1ull << 24;
    if (const char* env = std::getenv("LMB_SAMPLE_GROUP_PATHS")) target = std::strtoull(env, nullptr, 10);
    if (npix == 0 || samples < 2) return 1;
    const uint64_t g = std::min<uint64_t>(samples, target / npix);
    return g >= 2 ? (uint32_t)g : 1u;
}

// Per-path state of `paths` grouped sample paths.
static lmb_status ensure_group(lmb_data_buffers b, uint64_t paths) {
    if (b->grp_paths >= paths) return LMB_OK;
    b->grp_paths = 0;
    lmb_status st;
    if ((st = b->grp_seed.alloc(paths * 4)) || (st = b->grp_through.alloc(paths * 16)) ||
        (st = b->grp_accum.alloc(paths * 16)) || (st = b->grp_sum.alloc(paths * 16)) ||
        (st = b->grp_queue0.alloc(paths * 32)) || (st = b->grp_queue1.alloc(paths * 32)) ||
        (st = b->grp_hits.alloc(paths * 16)) ||
        (st = ensure_spill(&b->grp_spill, extend_threads(paths)))) {
        for (DeviceBuffer* d : {&b->grp_seed, &b->grp_through, &b->grp_accum, &b->grp_sum,
                                &b->grp_queue0, &b->grp_queue1, &b->grp_hits})
            d->release();
        return st;
    }
    b->grp_paths = paths;
    return LMB_OK;
}

// The bounce-0 hit buffer that a trace of several samples reuses.
static lmb_status ensure_primary_hits(lmb_data_buffers b, uint32_t samples) {
    const uint64_t bytes = (uint64_t)b->width * b->height * 16;
    if (samples < 2 || b->hits_primary.bytes >= bytes) return LMB_OK;
    lmb_status st = b->hits_primary.alloc(bytes);
    refresh_view(b);
    return st;
}

lmb_status lmb_trace(lmb_tracer tracer, lmb_scene scene, const lmb_camera* cam,
                     lmb_data_buffers bufs, uint32_t seed_offset, void* stream) {
    if (!tracer || !scene || !cam || !bufs) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    const lmb_tracer_opts& o = tracer->opts;
    if (o.guiding != LMB_GUIDING_MARKOV_CHAIN) {
        lmb_status st = ensure_primary_hits(bufs, o.samples);
        if (st) return st;
    }
    TraceParams p;
    p.quality = o.quality;
    p.samples = o.samples;
    p.tmin = o.t_min;
    p.tmax = o.t_max;
    p.restir = o.restir;
    p.bounces = o.bounces;
    p.rr_start = o.rr_start;
    p.guiding = o.guiding;
    p.seed_offset = seed_offset;
    CameraDev cd;
    std::memcpy(&cd, cam, sizeof cd);
    hipError_t e;
    if (o.quality > LMB_QUALITY_SPECTRAL) {
        e = launch_debug(scene->dev, cd, bufs->dev, p, (hipStream_t)stream, bufs->timer.get());
    } else if (o.guiding == LMB_GUIDING_MARKOV_CHAIN) {
        lmb_status st = ensure_guided(bufs);
        if (st) return st;
        e = launch_guided(scene->dev, cd, bufs->dev, bufs->guided, p, (hipStream_t)stream,
                          bufs->timer.get());
    } else {
        uint32_t group = sample_group(bufs->dev.npix, o.samples);
        // Without the memory for a group, the samples go one at a time.
        if (group > 1 && ensure_group(bufs, bufs->dev.npix * group) != LMB_OK) {
            group = 1;
            clear_error();
        }
        if (group > 1) {
            FrameDev fg = bufs->dev;
            fg.ps_seed = bufs->grp_seed.as<uint32_t>();
            fg.ps_through = bufs->grp_through.as<float4>();
            fg.ps_accum = bufs->grp_accum.as<float4>();
            fg.pix_sum = bufs->grp_sum.as<float4>();  // each path's sample colour
            fg.queue[0] = bufs->grp_queue0.as<float4>();
            fg.queue[1] = bufs->grp_queue1.as<float4>();
            fg.hits = bufs->grp_hits.as<uint4>();
            fg.hits_primary = nullptr;
            fg.spill = bufs->grp_spill.ptr;
            fg.own_seed = nullptr;
            fg.out_color = fg.out_normal = fg.out_albedo = nullptr;
            e = launch_trace_grouped(scene->dev, cd, bufs->dev, fg, group, p, (hipStream_t)stream,
                                     bufs->timer.get());
        } else {
            e = launch_trace(scene->dev, cd, bufs->dev, p, (hipStream_t)stream, bufs->timer.get());
        }
    }
    if (e != hipSuccess) return fail_hip(e, "trace launch");
    return LMB_OK;
}

// Samples of an overlapped cached frame traced together (their state is
// 132 bytes per pixel and sample).
constexpr uint32_t kWideSamples = 8;

static lmb_status trace_nrc(lmb_tracer tracer, lmb_scene scene, const lmb_camera* cam,
                            lmb_data_buffers bufs, uint32_t seed_offset, lmb_nrc nrc, void* stream,
                            bool train_only) {
    if (!tracer || !scene || !cam || !bufs || !nrc)
        return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    const lmb_tracer_opts& o = tracer->opts;
    if (o.guiding != LMB_GUIDING_OFF)
        return fail(LMB_ERR_INVALID_ARGUMENT, "the cached tracer needs guiding off");
    if (o.quality > LMB_QUALITY_SPECTRAL)
        return fail(LMB_ERR_INVALID_ARGUMENT, "the cached tracer has no debugging shaders");
    if (!train_only) {
        lmb_status pst = ensure_primary_hits(bufs, o.samples);
        if (pst) return pst;
    }
    TraceParams p;
    p.quality = o.quality;
    p.samples = train_only ? 1u : o.samples;
    p.tmin = o.t_min;
    p.tmax = o.t_max;
    p.restir = o.restir;
    p.bounces = o.bounces;
    p.rr_start = o.rr_start;
    p.guiding = o.guiding;
    p.seed_offset = seed_offset;
    CameraDev cd;
    std::memcpy(&cd, cam, sizeof cd);
    const uint64_t npix = (uint64_t)bufs->width * bufs->height;
    lmb_status st = nrc_ensure_queries(nrc, npix);
    if (st) return st;
    // Steps of ceil(records / steps) samples, records as the current options
    // expect them (the record buffers may be larger from earlier options).
    const uint32_t steps = nrc->opts.train_steps;
    const uint64_t expect = std::max<uint64_t>(kNrcChunk, (uint64_t)nrc->opts.train_paths * 5 / 4);
    const uint64_t rec = std::min<uint64_t>(nrc->dev.rec_cap, expect);
    st = nrc_ensure_batch(nrc, (uint32_t)((rec + steps - 1) / steps));
    if (st) return st;
    NrcTraceParams np{};
    for (int k = 0; k < 3; ++k) {
        const float ext = scene->bounds_hi[k] - scene->bounds_lo[k];
        np.lo[k] = scene->bounds_lo[k];
        np.inv_ext[k] = ext > 0.0f ? 1.0f / ext : 0.0f;
    }
    np.train_expect = (float)nrc->opts.train_paths;  // all in the trace's first sample
    np.defer_specular = nrc->opts.defer_specular;
    np.train_only = train_only ? 1u : 0u;
    np.stratify = !train_only && nrc->opts.stratify ? 1u : 0u;
    np.seq_base = seed_offset * p.samples;  // consecutive offsets continue one sequence
    const hipStream_t hs = (hipStream_t)stream;
    KernelTimer* tm = bufs->timer.get();
    const bool train = nrc->opts.train_paths > 0;
    if (train_only || !nrc->opts.overlap_training) {
        st = nrc_sync_frags(nrc, hs);
        if (st) return st;
        nrc->rec_in_train_ctl = false;
        // The first shading launch zeroes the record count (after counting
        // the previous trace's last sample for the selection probability).
        hipError_t e = launch_trace_nrc(scene->dev, cd, bufs->dev, p, nrc->dev, np, hs, tm);
        if (e != hipSuccess) return fail_hip(e, "cached trace launch");
        if (nrc->opts.train_paths == 0) return LMB_OK;
        return nrc_run_steps(nrc, nrc->dev, nullptr, nrc->dev.rec_in, nrc->dev.rec_target,
                             nrc->dev.ctl + kNrcCtlRecords, nrc->dev.rec_cap, steps, true, nullptr,
                             hs, tm);
    }
    // Overlapped training.  On the caller's stream: every sample ends at its
    // cache vertices.  On the training stream: a training pass (the camera
    // paths of about train_paths random pixels, as in a training frame, with
    // its own control words and path state) and the Adam steps, whose last
    // step publishes the new inference fragments into the set the samples do
    // not read.  The samples are enqueued first, so the GPU starts on them
    // while the host enqueues the training; at the end the caller's stream
    // waits for the training, so the call's work is complete when the
    // caller's stream is.  With train_paths = 0 (a frozen cache) the samples
    // are traced the same way and nothing runs on the training stream.
    st = nrc_ensure_overlap(nrc, npix);
    if (st) return st;
    // Samples traced up to kWideSamples at a time (path id pixel * w + j must
    // fit 31 bits); their state and w * npix queries, allocated before the
    // device views below are taken.
    const uint64_t w_max = std::min<uint64_t>(kWideSamples, 0x7FFFFFFFull / std::max<uint64_t>(npix, 1));
    const uint32_t w = (uint32_t)std::min<uint64_t>(p.samples, w_max);
    const bool wide = w > 1 && bufs->dev.hits_primary;
    if (wide) {
        st = nrc_ensure_wide(nrc, (uint64_t)w * npix);
        if (!st) st = nrc_ensure_queries(nrc, (uint64_t)w * npix);
        if (st) return st;
    }
    const hipStream_t ts = nrc->train_stream;
    uint16_t* sets = nrc->inf_sets.as<uint16_t>();
    uint16_t* cur = sets + (size_t)nrc->inf_cur * kNrcFwdFrags * kFragElems;
    uint16_t* nxt = sets + (size_t)(nrc->inf_cur ^ 1u) * kNrcFwdFrags * kFragElems;
    hipError_t e = hipSuccess;
    if (!nrc->inf_valid) {  // weights changed since the last overlapped frame
        NrcDev dp = nrc->dev;
        dp.fwd_pub = cur;
        e = nrc_launch_pack(dp, hs);
        if (e != hipSuccess) return fail_hip(e, "cache weight pack");
    }
    NrcDev da = nrc->dev;
    da.fwd_inf = cur;
    da.fwd_pub = nullptr;
    NrcDev db = nrc->dev;
    db.ctl = nrc->ctl_train.as<uint32_t>();
    db.fwd_pub = nullptr;
    // The training pass's pixel fraction from the cache vertices per sample
    // counted so far on the caller's stream.
    if (train) {
        e = hipMemcpyAsync(db.ctl + kNrcCtlHitsPrev, da.ctl + kNrcCtlHitsPrev, 4,
                           hipMemcpyDeviceToDevice, hs);
        if (e == hipSuccess) e = hipEventRecord(nrc->ev_start, hs);
        if (e != hipSuccess) return fail_hip(e, "training stream order");
    }
    np.train_expect = 0.0f;  // no training paths among the samples
    if (wide) {
        FrameDev fw = bufs->dev;
        fw.ps_seed = nrc->w_seed.as<uint32_t>();
        fw.ps_through = nrc->w_through.as<float4>();
        fw.ps_accum = nrc->w_accum.as<float4>();
        fw.pix_sum = nrc->w_sum.as<float4>();
        fw.queue[0] = nrc->w_queue0.as<float4>();
        fw.queue[1] = nrc->w_queue1.as<float4>();
        fw.hits = nrc->w_hits.as<uint4>();
        e = launch_trace_nrc_wide(scene->dev, cd, bufs->dev, fw, p, da, np, w, hs, tm);
    } else {
        e = launch_trace_nrc(scene->dev, cd, bufs->dev, p, da, np, hs, tm);
    }
    if (e != hipSuccess) return fail_hip(e, "cached trace launch");
    if (!train) {  // frozen: the samples read the current set, nothing is published
        nrc->inf_valid = true;
        nrc->rec_in_train_ctl = false;
        return LMB_OK;
    }
    e = hipStreamWaitEvent(ts, nrc->ev_start, 0);
    if (e != hipSuccess) return fail_hip(e, "training stream order");
    FrameDev ft = bufs->dev;
    ft.out_color = ft.out_normal = ft.out_albedo = nullptr;  // a training pass writes no image
    ft.pix_sum = nullptr;
    ft.ps_seed = nrc->tf_seed.as<uint32_t>();
    ft.ps_through = nrc->tf_through.as<float4>();
    ft.ps_accum = nrc->tf_accum.as<float4>();
    ft.queue[0] = nrc->tf_queue0.as<float4>();
    ft.queue[1] = nrc->tf_queue1.as<float4>();
    ft.hits = nrc->tf_hits.as<uint4>();
    ft.hits_primary = nullptr;
    ft.counters = nrc->tf_counters.as<uint32_t>();
    ft.spill = nrc->tf_spill.ptr;
    ft.grid_paths = std::max<uint64_t>(nrc->opts.train_paths, 256);
    TraceParams pt = p;
    pt.samples = 1;
    NrcTraceParams npt = np;
    npt.train_expect = (float)nrc->opts.train_paths;
    npt.train_only = 1u;
    npt.stratify = 0u;
    e = launch_trace_nrc(scene->dev, cd, ft, pt, db, npt, ts, tm);
    if (e != hipSuccess) return fail_hip(e, "cached trace training pass");
    st = nrc_run_steps(nrc, db, nxt, db.rec_in, db.rec_target, db.ctl + kNrcCtlRecords, db.rec_cap,
                       steps, true, nullptr, ts, tm);
    if (st) return st;
    e = hipEventRecord(nrc->ev_done, ts);
    if (e == hipSuccess) e = hipStreamWaitEvent(hs, nrc->ev_done, 0);
    if (e != hipSuccess) return fail_hip(e, "training stream order");
    nrc->inf_cur ^= 1u;  // the next frame reads what the steps published
    nrc->inf_valid = true;
    nrc->frags_stale = true;
    nrc->rec_in_train_ctl = true;
    return LMB_OK;
}

lmb_status lmb_trace_nrc(lmb_tracer tracer, lmb_scene scene, const lmb_camera* cam,
                         lmb_data_buffers bufs, uint32_t seed_offset, lmb_nrc nrc, void* stream) {
    return trace_nrc(tracer, scene, cam, bufs, seed_offset, nrc, stream, false);
}

lmb_status lmb_trace_nrc_train(lmb_tracer tracer, lmb_scene scene, const lmb_camera* cam,
                               lmb_data_buffers bufs, uint32_t seed_offset, lmb_nrc nrc,
                               void* stream) {
    if (nrc && nrc->opts.train_paths == 0)
        return fail(LMB_ERR_INVALID_ARGUMENT, "a training frame needs train_paths > 0");
    return trace_nrc(tracer, scene, cam, bufs, seed_offset, nrc, stream, true);
}

static lmb_status need_spatial_buffers(lmb_data_buffers b) {
    if (b->gi_count < (uint64_t)b->width * b->height)
        return fail(LMB_ERR_INVALID_ARGUMENT,
                    "resampling needs per-pixel reservoirs: create the buffers with "
                    "LMB_BUFFER_SPATIAL_RESAMPLING");
    return LMB_OK;
}

lmb_status lmb_temporal_resample(lmb_data_buffers bufs, uint32_t seed_offset, void* stream) {
    if (!bufs) return fail(LMB_ERR_INVALID_ARGUMENT, "buffers is NULL");
    lmb_status st = need_spatial_buffers(bufs);
    if (st) return st;
    hipError_t e = launch_temporal(bufs->dev, seed_offset, (hipStream_t)stream);
    if (e != hipSuccess) return fail_hip(e, "temporal resampling launch");
    return LMB_OK;
}

lmb_status lmb_spatial_resample(lmb_data_buffers bufs, uint32_t is_samples, uint32_t is_space,
                                uint32_t seed_offset, void* stream) {
    if (!bufs) return fail(LMB_ERR_INVALID_ARGUMENT, "buffers is NULL");
    if (is_space < 2 || is_space > 4096)
        return fail(LMB_ERR_INVALID_ARGUMENT, "IS_SPACE must be in 2..4096");
    lmb_status st = need_spatial_buffers(bufs);
    if (st) return st;
    hipError_t e = launch_spatial(bufs->dev, is_samples, is_space, seed_offset, (hipStream_t)stream);
    if (e != hipSuccess) return fail_hip(e, "spatial resampling launch");
    return LMB_OK;
}

// data_buffer/mod.rs:272-349.  "Copy current to old, clear current" is a
// pointer swap plus a clear; the world chains are copied (they persist).
lmb_status lmb_advance_frame(lmb_data_buffers b, uint32_t buffers_to_advance,
                             uint32_t confidence_cap, void* stream) {
    if (!b) return fail(LMB_ERR_INVALID_ARGUMENT, "buffers is NULL");
    if (confidence_cap == 0 || confidence_cap > 0xFFFF)  // NonZeroU16
        return fail(LMB_ERR_INVALID_ARGUMENT, "confidence_cap must be in 1..65535");
    if (buffers_to_advance & ~LMB_BUFFER_ALL)
        return fail(LMB_ERR_INVALID_ARGUMENT, "bad buffer flags");
    hipStream_t s = (hipStream_t)stream;
    hipError_t e = hipSuccess;
    auto swap_buf = [](DeviceBuffer& a, DeviceBuffer& c) {
        std::swap(a.ptr, c.ptr);
        std::swap(a.bytes, c.bytes);
    };
    if (buffers_to_advance & LMB_BUFFER_SPATIAL_RESAMPLING) {
        e = launch_eof_gi(b->gi.as<Reservoir>(), b->gi_count, confidence_cap, s);
        if (e == hipSuccess) {
            swap_buf(b->gi, b->gi_old);
            e = launch_fill_zero(b->gi.ptr, b->gi.bytes, s);
        }
    } else {
        e = launch_fill_zero(b->gi.ptr, b->gi.bytes, s);
        if (e == hipSuccess) e = launch_fill_zero(b->gi_old.ptr, b->gi_old.bytes, s);
    }
    if (e != hipSuccess) return fail_hip(e, "advance_frame (reservoirs)");
    if (buffers_to_advance & LMB_BUFFER_MARKOV_CHAIN_SCREEN_SPACE) {
        swap_buf(b->mc, b->mc_old);
        e = launch_fill_zero(b->mc.ptr, b->mc.bytes, s);
    } else {
        e = launch_fill_zero(b->mc.ptr, b->mc.bytes, s);
        if (e == hipSuccess) e = launch_fill_zero(b->mc_old.ptr, b->mc_old.bytes, s);
    }
    if (e != hipSuccess) return fail_hip(e, "advance_frame (screen chains)");
    if (buffers_to_advance & LMB_BUFFER_MARKOV_CHAIN_WORLD_SPACE) {
        e = launch_eof_world(b->wmc.as<WorldMarkovChain>(), b->wmc_count, s);
        if (e == hipSuccess)
            e = hipMemcpyAsync(b->wmc_old.ptr, b->wmc.ptr, b->wmc.bytes, hipMemcpyDeviceToDevice, s);
    } else {
        e = launch_fill_zero(b->wmc.ptr, b->wmc.bytes, s);
        if (e == hipSuccess) e = launch_fill_zero(b->wmc_old.ptr, b->wmc_old.bytes, s);
    }
    if (e != hipSuccess) return fail_hip(e, "advance_frame (world chains)");
    refresh_view(b);
    return LMB_OK;
}

}  // extern "C"
