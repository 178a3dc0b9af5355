// lmb_nrc.h - neural radiance cache: the network's shape and weight layout
// (shared by host and device), the device view of a cache, and the launch
// entry points of its kernels.
//
// Network: 13 inputs in [0, 1] (position in the scene box, direction and
// normal mapped from [-1, 1], roughness, albedo) -> frequency encoding
// sin / cos (2^k pi x), k = 0..11 (312 features) plus a constant 1 (the first
// layer's bias), zero-padded to 320 -> 64 -> 64 -> 64 -> 64 -> 64 -> 3, ReLU
// after every hidden layer, linear output.  Weights are bf16 on the matrix
// cores with fp32 accumulation; the optimiser keeps fp32 master weights.
//
// Matrix-core formulation.  Every layer computes Y^T = W^T X^T with
// v_mfma_f32_32x32x16_bf16: rows = output features, columns = 32 samples,
// K = input features.  The accumulator layout of that instruction (lane l,
// register r holds row 8(r/4) + 4(l/32) + r%4, column l%32) is turned into
// the next layer's B operand (lane l holds column l%32 and K = 8(l/32)..+7)
// without any data movement by permuting the K order of the next layer's
// weights (nrc_perm_hidden): activations never leave registers.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#include "lmb_frame_dev.h"

namespace lmb {

constexpr int kNrcInputs = 13;
constexpr int kNrcFreqs = 12;
constexpr int kNrcEncoded = kNrcInputs * kNrcFreqs * 2;  // 312; feature 312 = 1
constexpr int kNrcIn = 320;
constexpr int kNrcWidth = 64;
constexpr int kNrcHidden = 4;  // 64 -> 64 layers
constexpr int kNrcOut = 3;

// fp32 master weights, each layer [in][out] row-major.
constexpr int kNrcW0 = 0;                                    // 320 x 64
constexpr int kNrcWH = kNrcIn * kNrcWidth;                   // 4 x (64 x 64)
constexpr int kNrcW5 = kNrcWH + kNrcHidden * kNrcWidth * kNrcWidth;  // 64 x 3
constexpr int kNrcNumWeights = kNrcW5 + kNrcWidth * kNrcOut;  // 37056

// One MFMA A-operand fragment: 64 lanes x 8 bf16 = 1 KiB, lane-major.
constexpr int kFragElems = 512;
// Forward fragments: layer 0 (2 row tiles x 20 K steps), hidden layers
// (4 x 2 x 4), output (1 x 4).  Backward fragments (W instead of W^T, for
// the hidden and output layers): 4 x 2 x 4 + 2 x 1.
constexpr int kNrcFwdFrags = 40 + kNrcHidden * 8 + 4;  // 76
constexpr int kNrcBwdFrags = kNrcHidden * 8 + 2;       // 34
constexpr int kNrcFwdFragHidden = 40;
constexpr int kNrcFwdFragOut = 72;
constexpr int kNrcBwdFragOut = 32;

// Most optimiser steps of one call (one loss sum per step).
constexpr uint32_t kNrcMaxSteps = 256;

// Samples of one training column chunk of the weight-gradient kernel.
constexpr int kNrcChunk = 256;
// Output tiles (32 x 32) of the weight-gradient kernel: layer 0 10 x 2,
// hidden 4 x (2 x 2), output 2 x 1.
constexpr int kNrcGradTiles = 20 + kNrcHidden * 4 + 2;  // 38

// Feature index of K position (j, h, e) of a layer whose input comes from
// the accumulator of the previous 64-wide layer: K step j (16 features),
// lane half h, element e.
__host__ __device__ inline int nrc_perm_hidden(int j, int h, int e) {
    const int r = (j & 1) * 8 + e;
    return 32 * (j >> 1) + (r >> 2) * 8 + h * 4 + (r & 3);
}

// Master-weight index of element e of lane `lane` of forward fragment
// `frag`, or -1 for padding.
__host__ __device__ inline int nrc_fwd_index(int frag, int lane, int e) {
    const int h = lane >> 5, row = lane & 31;
    if (frag < kNrcFwdFragHidden) {
        const int mt = frag / 20, j = frag % 20;
        const int in = 16 * j + 8 * h + e, out = mt * 32 + row;
        return kNrcW0 + in * kNrcWidth + out;
    }
    if (frag < kNrcFwdFragOut) {
        const int f = frag - kNrcFwdFragHidden;
        const int layer = f / 8, mt = (f % 8) / 4, j = f % 4;
        const int in = nrc_perm_hidden(j, h, e), out = mt * 32 + row;
        return kNrcWH + layer * kNrcWidth * kNrcWidth + in * kNrcWidth + out;
    }
    const int j = frag - kNrcFwdFragOut;
    const int in = nrc_perm_hidden(j, h, e), out = row;
    return out < kNrcOut ? kNrcW5 + in * kNrcOut + out : -1;
}

// Master-weight index of an element of backward fragment `frag` (rows =
// input features, K = output features in accumulator order), or -1.
__host__ __device__ inline int nrc_bwd_index(int frag, int lane, int e) {
    const int h = lane >> 5, row = lane & 31;
    if (frag < kNrcBwdFragOut) {
        const int layer = frag / 8, mt = (frag % 8) / 4, j = frag % 4;
        const int in = mt * 32 + row, out = nrc_perm_hidden(j, h, e);
        return kNrcWH + layer * kNrcWidth * kNrcWidth + in * kNrcWidth + out;
    }
    const int mt = frag - kNrcBwdFragOut;
    const int in = mt * 32 + row, out = nrc_perm_hidden(0, h, e);
    return out < kNrcOut ? kNrcW5 + in * kNrcOut + out : -1;
}

// One cache query / training record: the 13 network inputs, the path id,
// and two spare words.
struct NrcQuery {
    float in[kNrcInputs];
    uint32_t pid;
    float spare[2];
};
static_assert(sizeof(NrcQuery) == 64, "NrcQuery is 64 bytes");
static_assert(offsetof(NrcQuery, pid) == 4 * kNrcInputs, "pid follows the inputs");

// Control words of a cache (device u32), each on a 256-byte line of its own:
// every wave of the tracer's shading does an atomic on the first two, and
// atomics on one line queue behind each other in L2.
constexpr uint32_t kNrcCtlStride = 64;
constexpr uint32_t kNrcCtlQueries = 0 * kNrcCtlStride;      // queries of the current sample
constexpr uint32_t kNrcCtlRecords = 1 * kNrcCtlStride;      // records reserved in the current trace
                                                            // (may pass rec_cap; those are dropped)
constexpr uint32_t kNrcCtlRecordsBase = 2 * kNrcCtlStride;  // min(records, rec_cap) at the sample's start
constexpr uint32_t kNrcCtlHitsPrev = 3 * kNrcCtlStride;     // cache vertices of the last sample
                                                            // traced (kept across traces)
constexpr uint32_t kNrcCtlHostCount = 4 * kNrcCtlStride;    // record count of the host-array calls
constexpr uint32_t kNrcCtlFrac = 5 * kNrcCtlStride;         // training frame: its pixel fraction (f32)
constexpr uint32_t kNrcCtlWords = 6 * kNrcCtlStride;

struct NrcDev {
    float* master;  // kNrcNumWeights fp32
    float* adam_m;
    float* adam_v;
    uint16_t* fwd;  // kNrcFwdFrags fragments (bf16 bits)
    uint16_t* bwd;  // kNrcBwdFrags fragments
    // Inference weights: an exponential moving average of the masters and its
    // forward fragments, or (ema == nullptr) the masters and `fwd` themselves.
    float* ema;
    uint16_t* fwd_inf;  // read by the inference kernels
    uint16_t* fwd_pub;  // written by k_nrc_pack from the inference weights (null: not written)
    // Fragment element of each master weight (nrc_fwd_index / nrc_bwd_index
    // inverted; bwd_pos -1 for layer 0): the Adam kernel re-packs the
    // weights it updates.
    const int32_t* fwd_pos;
    const int32_t* bwd_pos;
    float ema_decay;
    // Training scratch, feature-major bf16 with row stride `batch`:
    // x0 [320][batch], h [5][64][batch], dz [5][64][batch], g [3][batch].
    uint16_t* x0;
    uint16_t* h;
    uint16_t* dz;
    uint16_t* g;
    float* partial;  // [batch / kNrcChunk][kNrcNumWeights]
    uint32_t batch;  // samples per optimiser step (multiple of kNrcChunk)
    uint32_t* ctl;
    float* loss;  // per optimiser step (kNrcMaxSteps): sum of (y - t)^2 / (y^2 + 0.01)
    // queries of the tracer (one per path that reaches the cache vertex)
    NrcQuery* queries;
    float4* query_weight;  // path throughput at the cache vertex
    uint32_t query_cap;
    // training records of the tracer
    NrcQuery* rec_in;
    float4* rec_aux;  // 2 per record: (T.xyz, S.x), (S.y, S.z, -, -)
    float4* rec_target;
    uint32_t rec_cap;
};

struct AdamParams {
    float lr, beta1, beta2, eps;
    float bias1, bias2;  // 1 - beta1^t, 1 - beta2^t
};

// Re-pack the bf16 fragments from the master weights.
hipError_t nrc_launch_pack(const NrcDev& d, hipStream_t stream);
// out[i] = network(q[i].in) for i < n (n <= 2^31); out is float4 per query.
hipError_t nrc_launch_infer(const NrcDev& d, const NrcQuery* q, uint32_t n, float4* out,
                            hipStream_t stream);
// Tracer form: for the d.ctl[kNrcCtlQueries] queries of d.queries,
// pix_sum[pid] += query_weight * max(network, 0).
hipError_t nrc_launch_infer_accumulate(const NrcDev& d, float4* pix_sum, hipStream_t stream);
// The cached tracer: per sample generate, extend / shade (paths end at the
// second vertex with a query or become training records), then inference.
struct NrcTraceParams {
    float lo[3];            // scene box, position inputs = (p - lo) * inv_ext
    float inv_ext[3];
    float train_expect;     // expected training records per sample
    uint32_t sample;
    uint32_t defer_specular;  // most vertices a specular chain moves the cache vertex
    uint32_t train_only;      // a training frame: random pixels, no image
    uint32_t wide;            // > 1: samples traced up to this many at a time (path id
                              // pixel * wide + j for sample `sample` + j)
    uint32_t wide_count;      // samples of the current wide group (<= wide)
    uint32_t stratify;        // first scatter from the pixel's scrambled Sobol point
                              // seq_base + sample (+ j in a wide group)
    uint32_t seq_base;
};
hipError_t launch_trace_nrc(const SceneDev& s, const CameraDev& cam, const FrameDev& f,
                            const TraceParams& p, const NrcDev& d, NrcTraceParams np,
                            hipStream_t stream, KernelTimer* timer);
// The samples of a trace w at a time, without training paths: f_img is the
// image's frame (outputs, camera rays and their hits), fw the wide state
// (ps_*, pix_sum, queues and hits for w * npix path ids; counters and spill
// may be f_img's).  d.queries must hold w * npix queries.
hipError_t launch_trace_nrc_wide(const SceneDev& s, const CameraDev& cam, const FrameDev& f_img,
                                 const FrameDev& fw, const TraceParams& p, const NrcDev& d,
                                 NrcTraceParams np, uint32_t w, hipStream_t stream,
                                 KernelTimer* timer);

// One optimiser step on batch `step` of `steps`: records i * steps + step,
// i < d.batch, of the min(*count, count_limit) records `in` / `target`
// (*count is a device word).  grad_out
// (optional) receives the summed gradient; apply = false skips the update.
// grad_in (optional) replaces the computed gradient (no forward / backward).
hipError_t nrc_launch_step(const NrcDev& d, const NrcQuery* in, const float4* target,
                           const uint32_t* count, uint32_t count_limit, uint32_t steps,
                           uint32_t step,
                           const AdamParams& adam, bool apply, float* grad_out,
                           const float* grad_in, hipStream_t stream);

}  // namespace lmb
