// api_batch.cpp - C ABI of the batched tracer: many scenes and resolutions
// rendered by one set of persistent wavefront launches, replayed as a
// captured hipGraph, with a launch-configuration autotuner.  No counterpart in
// the reference, which dispatches one rt_main per image.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "lmb_batch.h"
#include "lmb_internal.h"

using namespace lmb;

namespace {

constexpr int kSeedSlots = 8;

}  // namespace

struct lmb_batch_s {
    int device = 0;
    lmb_tracer_opts opts{};
    lmb_batch_config cfg{};
    std::vector<lmb_scene> scenes;  // not owned
    std::vector<BatchItemDev> items;
    DeviceBuffer scene_table, item_table, seed_word;
    DeviceBuffer out_color, out_normal, out_albedo, own_seed, pix_sum, ps_seed, ps_through,
        ps_accum, queue0, queue1, hits, counters;
    DeviceBuffer hits_primary;  // per pixel: camera-ray hits reused by later samples
    // Samples traced together (path id pixel * group + j) and their per-path
    // state; group 1: one sample at a time.
    uint32_t group = 1;
    DeviceBuffer grp_seed, grp_through, grp_accum, grp_sum, grp_queue0, grp_queue1, grp_hits;
    uint64_t queue_cap = 0;
    SpillBuffer spill;
    FrameDev dev{};
    BatchDev bdev{};
    BatchLaunch launch{};
    uint64_t npix = 0;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    uint32_t* pinned = nullptr;  // seed slots
    hipEvent_t slot_ev[kSeedSlots] = {};
    int next_slot = 0;

    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~lmb_batch_s() {
        (void)hipSetDevice(device);
        (void)hipDeviceSynchronize();
        drop_graph();
        for (hipEvent_t& e : slot_ev)
            if (e) (void)hipEventDestroy(e);
        if (pinned) (void)hipHostFree(pinned);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
    }
};

namespace {

lmb_status validate_config(const lmb_batch_config& c) {
    if (c.blocks_per_cu < 1 || c.blocks_per_cu > 64)
        return fail(LMB_ERR_INVALID_ARGUMENT, "blocks_per_cu must be in 1..64");
    if (c.refill < 1 || c.refill > 64) return fail(LMB_ERR_INVALID_ARGUMENT, "refill must be in 1..64");
    if (c.tile_w < 1 || c.tile_h < 1 || (uint64_t)c.tile_w * c.tile_h > 4096)
        return fail(LMB_ERR_INVALID_ARGUMENT, "tile must be at least 1 x 1 and at most 4096 pixels");
    return LMB_OK;
}

// Tiles, passes, queue capacity and persistent grid for `c`; drops the graph.
lmb_status apply_config(lmb_batch b, const lmb_batch_config& c) {
    lmb_status st = validate_config(c);
    if (st) return st;
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail_hip(e, "hipDeviceSynchronize");
    b->drop_graph();
    uint64_t tiles = 0;
    for (BatchItemDev& it : b->items) {
        it.tiles_x = (it.width + c.tile_w - 1) / c.tile_w;
        it.tiles_y = (it.height + c.tile_h - 1) / c.tile_h;
        it.tile_base = tiles;
        tiles += (uint64_t)it.tiles_x * it.tiles_y;
    }
    const uint64_t tile_px = (uint64_t)c.tile_w * c.tile_h;
    uint64_t pass_tiles = c.pass_pixels ? std::max<uint64_t>(1, c.pass_pixels / tile_px) : tiles;
    pass_tiles = std::min(pass_tiles, tiles);
    const uint64_t cap = pass_tiles * tile_px;
    if (cap > 0xFFFFFFFFull) return fail(LMB_ERR_INVALID_ARGUMENT, "pass too large");
    if (cap > b->queue_cap) {
        st = b->queue0.alloc(cap * 32);
        if (!st) st = b->queue1.alloc(cap * 32);
        if (!st) st = b->hits.alloc(cap * 16);
        if (st) {
            b->queue_cap = 0;
            return st;
        }
        b->queue_cap = cap;
    }
    const uint64_t blocks = (uint64_t)device_cu_count() * c.blocks_per_cu;
    st = ensure_spill(&b->spill, blocks * 256);
    if (st) return st;
    st = b->item_table.upload(b->items.data(), b->items.size() * sizeof(BatchItemDev));
    if (st) return st;
    b->cfg = c;
    b->launch.blocks = (uint32_t)blocks;
    b->launch.refill = c.refill;
    b->launch.pass_tiles = pass_tiles;
    BatchDev& bd = b->bdev;
    bd.items = b->item_table.as<BatchItemDev>();
    bd.tile_w = c.tile_w;
    bd.tile_h = c.tile_h;
    bd.tiles_total = tiles;
    FrameDev& f = b->dev;
    f.queue[0] = b->queue0.as<float4>();
    f.queue[1] = b->queue1.as<float4>();
    f.hits = b->hits.as<uint4>();
    f.spill = b->spill.ptr;
    return LMB_OK;
}

// Samples a batch of npix pixels traces together: as many as keep a group
// at or under 2^27 paths (132 B of path state each, 17.7 GB; the environment
// variable LMB_BATCH_GROUP_PATHS overrides it, 0 = one sample at a time).
uint32_t batch_group(uint64_t npix, uint32_t samples) {
    uint64_t target = 1ull << 27;
    if (const char* env = std::getenv("LMB_BATCH_GROUP_PATHS")) target = std::strtoull(env, nullptr, 10);
    if (npix == 0 || samples < 2) return 1;
    const uint64_t g = std::min<uint64_t>(samples, target / npix);
    return g >= 2 ? (uint32_t)g : 1u;
}

// Per-path buffers of the sample groups; without the memory for them the
// batch traces one sample at a time.
void alloc_groups(lmb_batch b) {
    b->group = batch_group(b->npix, b->opts.samples);
    if (b->group < 2) return;
    const uint64_t paths = b->npix * b->group;
    if (b->grp_seed.alloc(paths * 4) || b->grp_through.alloc(paths * 16) ||
        b->grp_accum.alloc(paths * 16) || b->grp_sum.alloc(paths * 16) ||
        b->grp_queue0.alloc(paths * 32) || b->grp_queue1.alloc(paths * 32) ||
        b->grp_hits.alloc(paths * 16)) {
        for (DeviceBuffer* d : {&b->grp_seed, &b->grp_through, &b->grp_accum, &b->grp_sum,
                                &b->grp_queue0, &b->grp_queue1, &b->grp_hits})
            d->release();
        b->group = 1;
    }
}

// The frame view of the sample groups (nullptr: one sample at a time).
const FrameDev* group_view(lmb_batch b, FrameDev* g) {
    if (b->group < 2) return nullptr;
    *g = b->dev;
    g->group = b->group;
    g->grid_paths = b->npix * b->group;
    g->ps_seed = b->grp_seed.as<uint32_t>();
    g->ps_through = b->grp_through.as<float4>();
    g->ps_accum = b->grp_accum.as<float4>();
    g->pix_sum = b->grp_sum.as<float4>();
    g->queue[0] = b->grp_queue0.as<float4>();
    g->queue[1] = b->grp_queue1.as<float4>();
    g->hits = b->grp_hits.as<uint4>();
    return g;
}

TraceParams params_of(const lmb_tracer_opts& o) {
    TraceParams p;
    p.quality = o.quality;
    p.samples = o.samples;
    p.tmin = o.t_min;
    p.tmax = o.t_max;
    p.restir = o.restir;
    p.bounces = o.bounces;
    p.rr_start = o.rr_start;
    p.guiding = o.guiding;
    p.seed_offset = 0;  // the batch reads its seed from bdev.seed
    return p;
}

lmb_status ensure_graph(lmb_batch b) {
    if (b->exec) return LMB_OK;
    hipError_t e = hipSuccess;
    if (!b->cap_stream) e = hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking);
    if (e != hipSuccess) return fail_hip(e, "hipStreamCreate");
    e = hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return fail_hip(e, "hipStreamBeginCapture");
    FrameDev gv;
    hipError_t le = launch_batch(b->bdev, b->dev, group_view(b, &gv), params_of(b->opts), b->launch,
                                 b->cap_stream);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(b->cap_stream, &g);
    if (le != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return fail_hip(le, "batch launch capture");
    }
    if (e != hipSuccess) return fail_hip(e, "hipStreamEndCapture");
    b->graph = g;
    e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        b->exec = nullptr;
        return fail_hip(e, "hipGraphInstantiate");
    }
    return LMB_OK;
}

lmb_status run_frame(lmb_batch b, uint32_t seed_offset, hipStream_t stream) {
    // the seed goes through a pinned slot; a slot is reused only after the
    // copy that last read it has executed
    const int k = b->next_slot;
    b->next_slot = (k + 1) % kSeedSlots;
    hipError_t e = hipEventSynchronize(b->slot_ev[k]);
    if (e != hipSuccess) return fail_hip(e, "hipEventSynchronize");
    b->pinned[k] = seed_offset;
    e = hipMemcpyAsync(b->seed_word.ptr, b->pinned + k, 4, hipMemcpyHostToDevice, stream);
    if (e == hipSuccess) e = hipEventRecord(b->slot_ev[k], stream);
    if (e != hipSuccess) return fail_hip(e, "seed upload");
    if (b->cfg.use_graph) {
        lmb_status st = ensure_graph(b);
        if (st) return st;
        e = hipGraphLaunch(b->exec, stream);
        if (e != hipSuccess) return fail_hip(e, "hipGraphLaunch");
        return LMB_OK;
    }
    FrameDev gv;
    e = launch_batch(b->bdev, b->dev, group_view(b, &gv), params_of(b->opts), b->launch, stream);
    if (e != hipSuccess) return fail_hip(e, "batch launch");
    return LMB_OK;
}

lmb_status check(lmb_batch b) {
    if (!b) return fail(LMB_ERR_INVALID_ARGUMENT, "batch is NULL");
    hipError_t e = hipSetDevice(b->device);
    if (e != hipSuccess) return fail_hip(e, "hipSetDevice");
    return LMB_OK;
}

}  // namespace

extern "C" {

void lmb_batch_config_default(lmb_batch_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof *c);
    c->blocks_per_cu = 8;
    c->refill = 16;
    c->tile_w = 8;
    c->tile_h = 8;
    c->pass_pixels = 0;
    c->use_graph = 1;
}

lmb_status lmb_batch_create(const lmb_batch_item* items, uint32_t n, const lmb_tracer_opts* opts,
                            const lmb_batch_config* cfg, lmb_batch* out) {
    if (!out || !items || !opts) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    *out = nullptr;
    if (n == 0) return fail(LMB_ERR_INVALID_ARGUMENT, "a batch needs at least one item");
    if (opts->guiding != LMB_GUIDING_OFF)
        return fail(LMB_ERR_INVALID_ARGUMENT, "the batched tracer needs guiding off");
    if (opts->samples == 0 || opts->bounces == 0 || opts->bounces > kMaxBounces ||
        opts->quality > LMB_QUALITY_SPECTRAL)
        return fail(LMB_ERR_INVALID_ARGUMENT, "bad tracer options");
    lmb_batch_config c;
    lmb_batch_config_default(&c);
    if (cfg) c = *cfg;
    lmb_status st = validate_config(c);
    if (st) return st;
    uint64_t npix = 0;
    for (uint32_t i = 0; i < n; ++i) {
        if (!items[i].scene) return fail(LMB_ERR_INVALID_ARGUMENT, "item " + std::to_string(i) + " has no scene");
        if (items[i].scene->dev.alpha_test)
            return fail(LMB_ERR_INVALID_ARGUMENT,
                        "item " + std::to_string(i) + ": batches take opaque scenes only");
        if (items[i].width == 0 || items[i].height == 0)
            return fail(LMB_ERR_INVALID_ARGUMENT, "item " + std::to_string(i) + " has an empty image");
        npix += (uint64_t)items[i].width * items[i].height;
    }
    if (npix >= 0xFFFFFFFFull) return fail(LMB_ERR_INVALID_ARGUMENT, "more than 2^32 pixels");
    lmb_batch b = new lmb_batch_s();
    b->opts = *opts;
    (void)hipGetDevice(&b->device);
    b->npix = npix;
    std::vector<SceneDev> table(n);
    uint64_t base = 0;
    for (uint32_t i = 0; i < n; ++i) {
        b->scenes.push_back(items[i].scene);
        table[i] = items[i].scene->dev;
        BatchItemDev it{};
        std::memcpy(&it.cam, &items[i].camera, sizeof it.cam);
        it.width = items[i].width;
        it.height = items[i].height;
        it.pix_base = base;
        base += (uint64_t)it.width * it.height;
        b->items.push_back(it);
    }
    st = b->scene_table.upload(table.data(), n * sizeof(SceneDev));
    if (!st) st = b->seed_word.alloc(256);
    if (!st) st = b->out_color.alloc(npix * 16);
    if (!st) st = b->out_normal.alloc(npix * 16);
    if (!st) st = b->out_albedo.alloc(npix * 16);
    if (!st) st = b->own_seed.alloc(npix * 4);
    if (!st) st = b->pix_sum.alloc(npix * 16);
    if (!st) st = b->ps_seed.alloc(npix * 4);
    if (!st) st = b->ps_through.alloc(npix * 16);
    if (!st) st = b->ps_accum.alloc(npix * 16);
    if (!st) st = b->counters.alloc(kNumCounters * 4);
    if (!st && opts->samples > 1) st = b->hits_primary.alloc(npix * 16);
    if (st) {
        delete b;
        return st;
    }
    hipError_t e = hipHostMalloc((void**)&b->pinned, kSeedSlots * 4, hipHostMallocDefault);
    for (int k = 0; e == hipSuccess && k < kSeedSlots; ++k) e = hipEventCreate(&b->slot_ev[k]);
    if (e == hipSuccess) e = hipMemset(b->counters.ptr, 0, kNumCounters * 4);
    if (e != hipSuccess) {
        delete b;
        return fail_hip(e, "batch set-up");
    }
    FrameDev& f = b->dev;
    f.width = 0;
    f.height = 0;
    f.npix = npix;
    f.out_color = b->out_color.as<float4>();
    f.out_normal = b->out_normal.as<float4>();
    f.out_albedo = b->out_albedo.as<float4>();
    f.own_seed = b->own_seed.as<uint32_t>();
    f.pix_sum = b->pix_sum.as<float4>();
    f.ps_seed = b->ps_seed.as<uint32_t>();
    f.ps_through = b->ps_through.as<float4>();
    f.ps_accum = b->ps_accum.as<float4>();
    f.counters = b->counters.as<uint32_t>();
    f.hits_primary = b->hits_primary.as<uint4>();  // indexed by pixel here
    alloc_groups(b);
    BatchDev& bd = b->bdev;
    bd.scenes = b->scene_table.as<SceneDev>();
    bd.num_items = n;
    bd.seed = b->seed_word.as<uint32_t>();
    st = apply_config(b, c);
    if (st) {
        delete b;
        return st;
    }
    *out = b;
    return LMB_OK;
}

lmb_status lmb_batch_destroy(lmb_batch b) {
    delete b;
    return LMB_OK;
}

lmb_status lmb_batch_set_config(lmb_batch b, const lmb_batch_config* c) {
    lmb_status st = check(b);
    if (st) return st;
    if (!c) return fail(LMB_ERR_INVALID_ARGUMENT, "config is NULL");
    return apply_config(b, *c);
}

lmb_status lmb_batch_get_config(lmb_batch b, lmb_batch_config* c) {
    if (!b || !c) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    *c = b->cfg;
    return LMB_OK;
}

lmb_status lmb_batch_set_camera(lmb_batch b, uint32_t i, const lmb_camera* cam) {
    lmb_status st = check(b);
    if (st) return st;
    if (!cam || i >= b->items.size()) return fail(LMB_ERR_INVALID_ARGUMENT, "bad item or camera");
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail_hip(e, "hipDeviceSynchronize");
    std::memcpy(&b->items[i].cam, cam, sizeof(CameraDev));
    e = hipMemcpy(b->item_table.as<BatchItemDev>() + i, &b->items[i], sizeof(BatchItemDev),
                  hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy camera");
    return LMB_OK;
}

lmb_status lmb_batch_trace(lmb_batch b, uint32_t seed_offset, void* stream) {
    lmb_status st = check(b);
    if (st) return st;
    return run_frame(b, seed_offset, (hipStream_t)stream);
}

lmb_status lmb_batch_read_output(lmb_batch b, uint32_t i, float* host) {
    lmb_status st = check(b);
    if (st) return st;
    if (!host || i >= b->items.size()) return fail(LMB_ERR_INVALID_ARGUMENT, "bad item or buffer");
    const BatchItemDev& it = b->items[i];
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess)
        e = hipMemcpy(host, b->out_color.as<float4>() + it.pix_base,
                      (size_t)it.width * it.height * 16, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy output");
    return LMB_OK;
}

lmb_status lmb_batch_path_vertices(lmb_batch b, uint64_t* vertices) {
    lmb_status st = check(b);
    if (st) return st;
    if (!vertices) return fail(LMB_ERR_INVALID_ARGUMENT, "NULL argument");
    hipError_t e = hipDeviceSynchronize();
    if (e == hipSuccess)
        e = hipMemcpy(vertices, b->dev.counters + kCounterVertices, 8, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return fail_hip(e, "hipMemcpy counter");
    return LMB_OK;
}

uint32_t lmb_batch_default_sweep(lmb_batch_config* out, uint32_t cap) {
    // tile shape x grid size x refill threshold x pass size
    const uint32_t tiles[][2] = {{8, 8}, {16, 16}, {64, 1}};
    const uint32_t bpc[] = {4, 8, 16};
    const uint32_t refill[] = {8, 16, 32};
    const uint32_t pass[] = {0, 1u << 22};
    uint32_t n = 0;
    lmb_batch_config c;
    lmb_batch_config_default(&c);
    for (auto& t : tiles)
        for (uint32_t k : bpc)
            for (uint32_t r : refill)
                for (uint32_t pp : pass) {
                    c.tile_w = t[0];
                    c.tile_h = t[1];
                    c.blocks_per_cu = k;
                    c.refill = r;
                    c.pass_pixels = pp;
                    if (out && n < cap) out[n] = c;
                    ++n;
                }
    return n;
}

lmb_status lmb_batch_autotune(lmb_batch b, const lmb_batch_config* candidates, uint32_t n,
                              uint32_t frames, lmb_batch_config* best, double* ms) {
    lmb_status st = check(b);
    if (st) return st;
    if (!candidates || n == 0) return fail(LMB_ERR_INVALID_ARGUMENT, "no candidate configurations");
    for (uint32_t i = 0; i < n; ++i) {
        st = validate_config(candidates[i]);
        if (st) return st;
    }
    std::vector<lmb_batch_config> cands(candidates, candidates + n);
    if (frames == 0) frames = 1;
    const lmb_batch_config keep = b->cfg;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    hipError_t e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    if (e != hipSuccess) return fail_hip(e, "hipEventCreate");
    double best_ms = 1e300;
    size_t best_i = 0;
    for (size_t i = 0; i < cands.size() && !st; ++i) {
        st = apply_config(b, cands[i]);
        if (!st) st = run_frame(b, 0, nullptr);  // warm-up (and graph build)
        if (!st && hipDeviceSynchronize() != hipSuccess) st = fail(LMB_ERR_DEVICE, "autotune frame");
        if (st) break;
        (void)hipEventRecord(t0, nullptr);
        for (uint32_t f = 0; f < frames && !st; ++f) st = run_frame(b, f + 1, nullptr);
        (void)hipEventRecord(t1, nullptr);
        (void)hipEventSynchronize(t1);
        float el = 0.0f;
        (void)hipEventElapsedTime(&el, t0, t1);
        const double per = (double)el / frames;
        if (ms) ms[i] = per;
        if (per < best_ms) {
            best_ms = per;
            best_i = i;
        }
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    if (st) {
        (void)apply_config(b, keep);
        return st;
    }
    st = apply_config(b, cands[best_i]);
    if (!st && best) *best = cands[best_i];
    return st;
}

}  // extern "C"
