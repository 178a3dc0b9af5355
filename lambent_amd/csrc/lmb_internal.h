// lmb_internal.h - host-side objects behind the opaque C-ABI handles.
#pragma once

#include <hip/hip_runtime_api.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/lambent.h"
#include "bvh_build.h"
#include "lmb_frame_dev.h"
#include "lmb_layout.h"
#include "lmb_nrc.h"
#include "lmb_scene_dev.h"

namespace lmb {

lmb_status fail(lmb_status code, const std::string& msg);
void clear_error();  // forgets the last error message (a failure handled internally)
lmb_status fail_hip(hipError_t e, const char* what);

// Owning device allocation.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    lmb_status alloc(size_t n);  // frees the previous allocation first
    lmb_status upload(const void* host, size_t n);
    void release();
    template <class T>
    T* as() const {
        return static_cast<T*>(ptr);
    }
};

// Per-thread traversal-stack overflow area of the persistent ray kernels.
struct SpillBuffer {
    DeviceBuffer buf;
    uint64_t threads = 0;
    uint2* ptr = nullptr;
};
lmb_status ensure_spill(SpillBuffer* sb, uint64_t threads);

struct HostImage {
    uint32_t width = 0, height = 0;
    std::vector<uint32_t> texels;  // RGBA8, r in the low byte
};

}  // namespace lmb

struct lmb_textures_s {
    std::vector<lmb::HostImage> diffuse, emission, attributes;
    lmb::HostImage recolour;
    bool has_recolour = false;
    lmb::HostImage background[6];
    bool has_background = false;
};

namespace lmb {

// One instance of an instanced scene with triangles in the BVH: an
// instance-pure subtree whose root's node position the top tree decides.
struct InstancePart {
    uint32_t instance = 0;             // index in the creation desc
    uint32_t first_tri = 0, num_tris = 0;  // leaf-order triangle range
    uint32_t height = 0;               // node height of its root (0: no internal children)
    uint32_t levels = 0;               // node levels of the subtree
    uint32_t num_nodes = 0;
    double sah = 0.0;                  // node areas relative to its root, at creation
    float obj_lo[3] = {0.f, 0.f, 0.f}, obj_hi[3] = {0.f, 0.f, 0.f};  // object-space box
};

// What moving the instances of an instanced scene needs (api_instances.cpp).
struct InstanceState {
    uint32_t num_instances = 0;
    std::vector<uint32_t> geometry;    // per desc instance
    std::vector<int32_t> part_of;      // per desc instance: part index or -1
    std::vector<uint8_t> hidden_at_creation;  // per desc instance: triangles but mask 0
    std::vector<InstancePart> parts;
    std::vector<uint32_t> level_off;   // static refit entries of height h: [off[h], off[h + 1])
    uint32_t region = 0;               // node positions of the top tree and the part roots
    DeviceBuffer entries;              // static RefitEntry, by height
    DeviceBuffer top;                  // RefitEntry of the top tree, by height
    DeviceBuffer root_pos;             // u32 per part
    DeviceBuffer first;                // u32 per part + 1: leaf-order triangle ranges
    DeviceBuffer xf;                   // PartTransform per part
    DeviceBuffer box;                  // 2 float4 per node position
    DeviceBuffer error;                // u32
};

}  // namespace lmb

struct lmb_scene_s {
    int device = 0;
    lmb::SceneDev dev{};
    lmb::DeviceBuffer nodes, tris, prim_material, materials, texels, tex;
    lmb::DeviceBuffer obj_tris;  // instanced scenes: object-space triangles (or empty)
    std::unique_ptr<lmb::InstanceState> inst;  // instanced scenes only
    lmb::DeviceBuffer fetch;  // ray counter of the query kernels
    lmb::SpillBuffer spill;
    lmb_bvh_stats stats{};
    uint64_t num_prims = 0;
    float bounds_lo[3] = {0.f, 0.f, 0.f};  // box of the vertices
    float bounds_hi[3] = {0.f, 0.f, 0.f};
};

struct lmb_nrc_s {
    int device = 0;
    lmb_nrc_opts opts{};
    lmb::NrcDev dev{};
    lmb::DeviceBuffer master, adam_m, adam_v, fwd, bwd, ema, fwd_inf;
    lmb::DeviceBuffer fwd_pos, bwd_pos;  // fragment element of each weight
    lmb::DeviceBuffer x0, h, dz, g, partial;  // training scratch for batch_cap samples
    uint32_t batch_cap = 0;
    lmb::DeviceBuffer ctl, loss;
    lmb::DeviceBuffer queries, query_weight;  // tracer queries (query_cap)
    lmb::DeviceBuffer rec_in, rec_aux, rec_target;
    lmb::DeviceBuffer io_in, io_target, io_out, io_grad;  // staging of the host-array calls
    uint64_t steps = 0;                  // Adam steps taken (t)
    double pow1 = 1.0, pow2 = 1.0;       // beta1^t, beta2^t
    uint32_t last_steps = 0;             // optimiser steps of the last call
    // Overlapped training (opts.overlap_training): the training pass and the
    // optimiser steps of a cached frame run on train_stream beside the
    // frame's samples, with their own control words and tracing state; the
    // samples read inference fragment set inf_cur while the steps publish
    // the next set into the other one.
    hipStream_t train_stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_done = nullptr;
    lmb::DeviceBuffer ctl_train, inf_sets;  // inf_sets: 2 x kNrcFwdFrags fragments
    uint32_t inf_cur = 0;
    bool inf_valid = false;    // set inf_cur holds the current inference weights
    bool frags_stale = false;  // dev.fwd_pub (EMA fragments) behind the weights
    bool rec_in_train_ctl = false;  // the last trace's records are counted in ctl_train
    uint64_t tf_npix = 0, tf_grid = 0;
    lmb::DeviceBuffer tf_seed, tf_through, tf_accum, tf_queue0, tf_queue1, tf_hits, tf_counters;
    lmb::SpillBuffer tf_spill;
    // Wide tracing of an overlapped frame's samples: state for wide_paths
    // path ids (samples traced together x pixels).
    uint64_t wide_paths = 0;
    lmb::DeviceBuffer w_seed, w_through, w_accum, w_sum, w_queue0, w_queue1, w_hits;
};

namespace lmb {
lmb_status scene_upload(lmb_scene s, const lmb_scene_desc* desc, uint64_t ntri,
                        const BvhBuildResult& bvh);
void scene_free(lmb_scene s);

// Neural cache helpers (api_nrc.cpp).
lmb_status nrc_ensure_batch(lmb_nrc n, uint32_t batch);
lmb_status nrc_ensure_queries(lmb_nrc n, uint64_t queries);
// `steps` optimiser steps over the min(*count, count_limit) records of
// in / target (*count is a device word), record i * steps + s in step s.
// Zeroes the per-step loss sums first.
// d is the cache's device view to launch with (n->dev, or a copy with other
// control words); publish (optional) receives the inference fragments of the
// weights after the last step.
lmb_status nrc_run_steps(lmb_nrc n, const NrcDev& d, uint16_t* publish, const NrcQuery* in,
                         const float4* target, const uint32_t* count, uint32_t count_limit,
                         uint32_t steps, bool apply, float* grad_out, hipStream_t stream,
                         KernelTimer* timer);
// Brings the synchronous inference fragments (dev.fwd_inf) up to date after
// overlapped training, on `stream`.
lmb_status nrc_sync_frags(lmb_nrc n, hipStream_t stream);
// The overlapped form of a cached frame's training: buffers for a frame of
// npix pixels, the stream and events (created on first use).
lmb_status nrc_ensure_overlap(lmb_nrc n, uint64_t npix);
// State of the wide tracing for `paths` path ids.
lmb_status nrc_ensure_wide(lmb_nrc n, uint64_t paths);
}  // namespace lmb
