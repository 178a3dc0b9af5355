// nrc.hip - neural radiance cache network on the gfx950 matrix cores.
//
// Kernels (all bf16 operands, fp32 accumulation, v_mfma_f32_32x32x16_bf16):
//   k_nrc_pack     fp32 master weights -> bf16 A-operand fragments (forward
//                  W^T and backward W, K order permuted as lmb_nrc.h states)
//   k_nrc_infer    fused encode + 6-layer MLP; the 76 KiB of forward
//                  fragments sit in LDS, activations stay in registers; one
//                  wave evaluates 64 samples per pass (two column tiles)
//   k_nrc_fwd_bwd  training forward + relative-L2 loss + backward for 32
//                  samples per wave; stores the layer inputs and the
//                  pre-activation gradients feature-major for
//   k_nrc_dw       dW = X . dZ^T per 32 x 32 weight tile and 256-sample chunk
//                  (K = samples), partial sums per chunk
//   k_nrc_adam     sums the chunk partials, applies Adam to the masters and
//                  writes the updated weights' bf16 fragment elements
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>

#include "lmb_nrc.h"
#include "lmb_scene_dev.h"

namespace lmb {

typedef __bf16 bf16x8 __attribute__((__vector_size__(8 * sizeof(__bf16))));
typedef float f32x8 __attribute__((__vector_size__(8 * sizeof(float))));
typedef float f32x16 __attribute__((__vector_size__(16 * sizeof(float))));

constexpr int kNrcBlock = 256;  // 4 waves

template <class Fn, int... I>
__device__ __forceinline__ void static_for_impl(Fn&& fn, std::integer_sequence<int, I...>) {
    (fn(std::integral_constant<int, I>{}), ...);
}
template <int N, class Fn>
__device__ __forceinline__ void static_for(Fn&& fn) {
    static_for_impl(fn, std::make_integer_sequence<int, N>{});
}

__device__ __forceinline__ f32x16 mfma(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x8 to_bf16(f32x8 v) { return __builtin_convertvector(v, bf16x8); }
__device__ __forceinline__ uint16_t bits(__bf16 x) { return __builtin_bit_cast(uint16_t, x); }

// Copies n 16-byte fragment elements from global memory into LDS with a
// BLOCK-thread block, issuing each thread's loads back to back in batches of
// kFillBatch before their LDS writes (a load-wait-write loop would pay one
// L2 round trip per element).
constexpr int kFillBatch = 8;
template <int BLOCK>
__device__ __forceinline__ void fill_lds(bf16x8* __restrict__ lds, const bf16x8* __restrict__ src,
                                         int n) {
    for (int base = 0; base < n; base += BLOCK * kFillBatch) {
        bf16x8 v[kFillBatch];
#pragma unroll
        for (int k = 0; k < kFillBatch; ++k) {
            const int i = base + k * BLOCK + (int)threadIdx.x;
            if (i < n) v[k] = src[i];
        }
#pragma unroll
        for (int k = 0; k < kFillBatch; ++k) {
            const int i = base + k * BLOCK + (int)threadIdx.x;
            if (i < n) lds[i] = v[k];
        }
    }
}

// fill_lds over two sources into one LDS range: elements [0, n0) from s0,
// [n0, n0 + n1) from s1, kFillBatch2 loads in flight per thread.
constexpr int kFillBatch2 = 14;
template <int BLOCK>
__device__ __forceinline__ void fill_lds2(bf16x8* __restrict__ lds, const bf16x8* __restrict__ s0,
                                          int n0, const bf16x8* __restrict__ s1, int n1) {
    const int n = n0 + n1;
    for (int base = 0; base < n; base += BLOCK * kFillBatch2) {
        bf16x8 v[kFillBatch2];
#pragma unroll
        for (int k = 0; k < kFillBatch2; ++k) {
            const int i = base + k * BLOCK + (int)threadIdx.x;
            if (i < n) v[k] = i < n0 ? s0[i] : s1[i - n0];
        }
#pragma unroll
        for (int k = 0; k < kFillBatch2; ++k) {
            const int i = base + k * BLOCK + (int)threadIdx.x;
            if (i < n) lds[i] = v[k];
        }
    }
}

struct FragPtr {
    const bf16x8* p;
    __device__ __forceinline__ bf16x8 operator()(int frag, int lane) const {
        return p[frag * 64 + lane];
    }
};

// K step j of layer 0 for one sample: features 16 j + 8 h .. +7, i.e. block
// b = 2 j + h = 3 s + m: scalar s, frequencies 4 m .. 4 m + 3 as (sin, cos)
// pairs; block 39 is the constant feature 312 = 1 followed by zeros.
// sin / cos (2^k pi x) = sin / cos (2 pi fract(2^(k-1) x)): the argument
// reduction is exact; one hardware sin / cos (unit = revolutions) per block,
// then angle doubling.
template <int J>
__device__ __forceinline__ bf16x8 encode_kstep(const float (&x)[kNrcInputs], int h) {
    constexpr int b0 = 2 * J, b1 = 2 * J + 1;
    constexpr int s0 = b0 / 3, s1 = b1 / 3 < kNrcInputs ? b1 / 3 : kNrcInputs - 1;
    constexpr float sc0 = (b0 % 3 == 0) ? 0.5f : (b0 % 3 == 1) ? 8.0f : 128.0f;
    constexpr float sc1 = (b1 % 3 == 0) ? 0.5f : (b1 % 3 == 1) ? 8.0f : 128.0f;
    const float xv = h ? x[s1] : x[s0];
    float rev = xv * (h ? sc1 : sc0);
    rev = rev - floorf(rev);
    float sn = __builtin_amdgcn_sinf(rev), cs = __builtin_amdgcn_cosf(rev);
    f32x8 v;
    v[0] = sn;
    v[1] = cs;
#pragma unroll
    for (int m = 1; m < 4; ++m) {
        const float s2 = 2.0f * sn * cs;
        const float c2 = fmaf(-2.0f * sn, sn, 1.0f);
        sn = s2;
        cs = c2;
        v[2 * m] = sn;
        v[2 * m + 1] = cs;
    }
    if (b1 / 3 >= kNrcInputs && h) {
        v = f32x8{1.0f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    }
    return to_bf16(v);
}

// Accumulator tiles (two row tiles of 32 features) -> ReLU -> B operand.
template <int NT>
__device__ __forceinline__ void relu_to_b(const f32x16 (&c)[2][NT], bf16x8 (&hb)[4][NT]) {
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            f32x8 v;
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = fmaxf(c[j >> 1][t][(j & 1) * 8 + e], 0.0f);
            hb[j][t] = to_bf16(v);
        }
}

// Layer 0 (320 -> 64); `emit(j, t, b)` sees every encoded K step.
template <int NT, class F, class Emit>
__device__ __forceinline__ void layer0(const F& fr, const float (&x)[NT][kNrcInputs], int lane,
                                       f32x16 (&c)[2][NT], Emit&& emit) {
    const int h = lane >> 5;
#pragma unroll
    for (int t = 0; t < NT; ++t) c[0][t] = c[1][t] = f32x16{};
    static_for<20>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        const bf16x8 a0 = fr(j, lane), a1 = fr(20 + j, lane);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const bf16x8 b = encode_kstep<j>(x[t], h);
            emit(j, t, b);
            c[0][t] = mfma(a0, b, c[0][t]);
            c[1][t] = mfma(a1, b, c[1][t]);
        }
        // one K step's fragment reads at a time: hoisting all 40 of them
        // ahead of the MFMAs needs more than 256 registers (spills)
        __builtin_amdgcn_sched_barrier(0);
    });
}

// 64 -> 64 layer from B-form input; fragments fbase .. fbase + 7.
template <int NT, class F>
__device__ __forceinline__ void layer64(const F& fr, int fbase, const bf16x8 (&in)[4][NT], int lane,
                                        f32x16 (&c)[2][NT]) {
#pragma unroll
    for (int t = 0; t < NT; ++t) c[0][t] = c[1][t] = f32x16{};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const bf16x8 a0 = fr(fbase + j, lane), a1 = fr(fbase + 4 + j, lane);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            c[0][t] = mfma(a0, in[j][t], c[0][t]);
            c[1][t] = mfma(a1, in[j][t], c[1][t]);
        }
    }
}

// 64 -> 3 output layer: rows 0..2 of the result sit in registers 0..2 of
// lanes 0..31 (lane = sample).
template <int NT, class F>
__device__ __forceinline__ void layer_out(const F& fr, const bf16x8 (&in)[4][NT], int lane,
                                          f32x16 (&y)[NT]) {
#pragma unroll
    for (int t = 0; t < NT; ++t) y[t] = f32x16{};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const bf16x8 a = fr(kNrcFwdFragOut + j, lane);
#pragma unroll
        for (int t = 0; t < NT; ++t) y[t] = mfma(a, in[j][t], y[t]);
    }
}

__device__ __forceinline__ void load_inputs(const NrcQuery* __restrict__ q, uint64_t i, bool valid,
                                            float (&x)[kNrcInputs]) {
    if (valid) {
        const float4* p = reinterpret_cast<const float4*>(q + i);
        const float4 a = p[0], b = p[1], c = p[2], d = p[3];
        x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w;
        x[4] = b.x; x[5] = b.y; x[6] = b.z; x[7] = b.w;
        x[8] = c.x; x[9] = c.y; x[10] = c.z; x[11] = c.w;
        x[12] = d.x;
    } else {
#pragma unroll
        for (int k = 0; k < kNrcInputs; ++k) x[k] = 0.0f;
    }
}

__global__ __launch_bounds__(kNrcBlock) void k_nrc_pack(NrcDev d) {
    const int i = blockIdx.x * kNrcBlock + threadIdx.x;
    if (i >= (kNrcFwdFrags + kNrcBwdFrags) * kFragElems) return;
    const int frag = i / kFragElems, lane = (i % kFragElems) / 8, e = i % 8;
    if (frag < kNrcFwdFrags) {
        const int w = nrc_fwd_index(frag, lane, e);
        d.fwd[i] = w >= 0 ? bits((__bf16)d.master[w]) : (uint16_t)0;
        if (d.fwd_pub)
            d.fwd_pub[i] = w >= 0 ? bits((__bf16)(d.ema ? d.ema[w] : d.master[w])) : (uint16_t)0;
    } else {
        const int w = nrc_bwd_index(frag - kNrcFwdFrags, lane, e);
        d.bwd[i - kNrcFwdFrags * kFragElems] = w >= 0 ? bits((__bf16)d.master[w]) : (uint16_t)0;
    }
}

// Inference.  kAccumulate: pix_sum[pid] += weight * max(y, 0) for the
// queries counted in *count; otherwise out[i] = y for i < count.
// Inference blocks: 8 waves share one LDS copy of the forward fragments,
// one block per CU (2 waves per SIMD; 4 waves at two blocks per CU, or
// 4-wave blocks, measured slower).
constexpr int kNrcInferWaves = 8;
constexpr int kNrcInferBlock = 64 * kNrcInferWaves;
constexpr int kNrcInferBlocksPerCu = 1;

template <bool kAccumulate>
__global__ __launch_bounds__(kNrcInferBlock) void k_nrc_infer(NrcDev d, const NrcQuery* __restrict__ q,
                                                         const float4* __restrict__ weight,
                                                         const uint32_t* __restrict__ count_ptr,
                                                         uint32_t count_fixed,
                                                         float4* __restrict__ dst) {
    __shared__ bf16x8 lds[kNrcFwdFrags * 64];
    const bf16x8* src = reinterpret_cast<const bf16x8*>(d.fwd_inf);
    fill_lds<kNrcInferBlock>(lds, src, kNrcFwdFrags * 64);
    __syncthreads();
    const uint32_t count = count_ptr ? *count_ptr : count_fixed;
    const int lane = threadIdx.x & 63, n = lane & 31, h = lane >> 5;
    const uint32_t tiles = (count + 63u) / 64u;
    const FragPtr fr{lds};
    for (uint32_t tile = blockIdx.x * (uint32_t)kNrcInferWaves + (threadIdx.x >> 6); tile < tiles;
         tile += gridDim.x * (uint32_t)kNrcInferWaves) {
        float x[2][kNrcInputs];
        uint32_t idx[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            idx[t] = tile * 64u + t * 32u + n;
            load_inputs(q, idx[t], idx[t] < count, x[t]);
        }
        f32x16 c[2][2];
        bf16x8 hb[4][2];
        layer0<2>(fr, x, lane, c, [](int, int, bf16x8) {});
        relu_to_b<2>(c, hb);
#pragma unroll 1
        for (int l = 0; l < kNrcHidden; ++l) {
            layer64<2>(fr, kNrcFwdFragHidden + 8 * l, hb, lane, c);
            relu_to_b<2>(c, hb);
        }
        f32x16 y[2];
        layer_out<2>(fr, hb, lane, y);
        if (h == 0) {
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                if (idx[t] >= count) continue;
                if (kAccumulate) {
                    const uint32_t pid = q[idx[t]].pid;
                    const float4 w = weight[idx[t]];
                    const float4 ps = dst[pid];
                    dst[pid] = make_float4(ps.x + w.x * fmaxf(y[t][0], 0.0f),
                                           ps.y + w.y * fmaxf(y[t][1], 0.0f),
                                           ps.z + w.z * fmaxf(y[t][2], 0.0f), ps.w);
                } else {
                    dst[idx[t]] = make_float4(y[t][0], y[t][1], y[t][2], 0.0f);
                }
            }
        }
    }
}

// Stores a B-form K step (lane = sample column i, 8 features) feature-major.
__device__ __forceinline__ void store_kstep(uint16_t* __restrict__ base, uint32_t stride,
                                            uint32_t col, bf16x8 v, int j, int h, bool hidden) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const int row = hidden ? nrc_perm_hidden(j, h, e) : 16 * j + 8 * h + e;
        base[(size_t)row * stride + col] = bits(v[e]);
    }
}

// dz = bf16(d) where the ReLU output h is positive, else 0 (B form).
__device__ __forceinline__ void relu_grad(const f32x16 (&dacc)[2][1], const bf16x8 (&hb)[4][1],
                                          bf16x8 (&dz)[4][1]) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        f32x8 v;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float hv = (float)hb[j][0][e];
            v[e] = hv > 0.0f ? dacc[j >> 1][0][(j & 1) * 8 + e] : 0.0f;
        }
        dz[j][0] = to_bf16(v);
    }
}

// Training forward + loss + backward for 32 samples per wave.  Sample i of
// this batch is record i * steps + step.
__global__ __launch_bounds__(kNrcBlock) void k_nrc_fwd_bwd(NrcDev d, const NrcQuery* __restrict__ in,
                                                           const float4* __restrict__ target,
                                                           const uint32_t* __restrict__ count_ptr,
                                                           uint32_t count_limit, uint32_t steps,
                                                           uint32_t step) {
    __shared__ bf16x8 lds[(kNrcFwdFrags + kNrcBwdFrags) * 64];
    fill_lds2<kNrcBlock>(lds, reinterpret_cast<const bf16x8*>(d.fwd), kNrcFwdFrags * 64,
                         reinterpret_cast<const bf16x8*>(d.bwd), kNrcBwdFrags * 64);
    __syncthreads();
    const FragPtr fr{lds}, br{lds + kNrcFwdFrags * 64};
    const uint32_t S = d.batch;
    const uint32_t count = min(*count_ptr, count_limit);
    const uint32_t nvalid = count > step ? min(S, (count - step + steps - 1u) / steps) : 0u;
    const float gscale = nvalid ? 2.0f / (3.0f * (float)nvalid) : 0.0f;
    const int lane = threadIdx.x & 63, n = lane & 31, h = lane >> 5;
    const uint32_t base = (blockIdx.x * 4u + (threadIdx.x >> 6)) * 32u;
    if (base >= S) return;
    const uint32_t i = base + n;
    const bool valid = i < nvalid;
    const uint64_t rec = (uint64_t)i * steps + step;
    float x[1][kNrcInputs];
    load_inputs(in, rec, valid, x[0]);

    f32x16 c[2][1];
    bf16x8 hb[5][4][1];
    layer0<1>(fr, x, lane, c,
              [&](int j, int, bf16x8 b) { store_kstep(d.x0, S, i, b, j, h, false); });
    relu_to_b<1>(c, hb[0]);
    static_for<kNrcHidden>([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        layer64<1>(fr, kNrcFwdFragHidden + 8 * l, hb[l], lane, c);
        relu_to_b<1>(c, hb[l + 1]);
    });
    static_for<5>([&](auto lc) {
        constexpr int l = decltype(lc)::value;
        uint16_t* hp = d.h + (size_t)l * kNrcWidth * S;
#pragma unroll
        for (int j = 0; j < 4; ++j) store_kstep(hp, S, i, hb[l][j][0], j, h, true);
    });
    f32x16 y[1];
    layer_out<1>(fr, hb[4], lane, y);

    // relative L2: sum (y - t)^2 / (sg(y)^2 + 0.01); dL/dy = 2 (y - t) / (y^2 + 0.01)
    float gy[3] = {0.f, 0.f, 0.f};
    float lsum = 0.0f;
    if (h == 0 && valid) {
        const float4 t4 = target[rec];
        const float tv[3] = {t4.x, t4.y, t4.z};
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const float dd = y[0][k] - tv[k];
            const float den = y[0][k] * y[0][k] + 0.01f;
            lsum += dd * dd / den;
            gy[k] = gscale * dd / den;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_xor(lsum, off);
    if (lane == 0) atomicAdd(d.loss + step, lsum);
    bf16x8 gb[4][1];
    gb[0][0] = to_bf16(f32x8{gy[0], gy[1], gy[2], 0.f, 0.f, 0.f, 0.f, 0.f});
    if (h == 0) {
#pragma unroll
        for (int k = 0; k < 3; ++k) d.g[(size_t)k * S + i] = bits(gb[0][0][k]);
    }
    // backward through the output layer: dh5 = W5 . g (K = 16, one step)
    f32x16 dacc[2][1];
    dacc[0][0] = mfma(br(kNrcBwdFragOut, lane), gb[0][0], f32x16{});
    dacc[1][0] = mfma(br(kNrcBwdFragOut + 1, lane), gb[0][0], f32x16{});
    bf16x8 dz[4][1];
    relu_grad(dacc, hb[4], dz);
    static_for<kNrcHidden + 1>([&](auto kc) {
        constexpr int k = decltype(kc)::value;  // dz holds dZ_{5-k}
        constexpr int l = 4 - k;                // its index among h / dz arrays
        uint16_t* zp = d.dz + (size_t)l * kNrcWidth * S;
#pragma unroll
        for (int j = 0; j < 4; ++j) store_kstep(zp, S, i, dz[j][0], j, h, true);
        if constexpr (l > 0) {
            // dh_l = W_l . dz_{l+1} with the backward fragments of hidden layer l-1
            layer64<1>(br, 8 * (l - 1), dz, lane, dacc);
            relu_grad(dacc, hb[l - 1], dz);
        }
    });
}

// dW tile (32 input rows x 32 output columns) of one layer over one chunk of
// kNrcChunk samples: K = samples, both operands loaded straight from the
// feature-major scratch (16 contiguous bytes per lane).
// Weight-gradient tile `tile` (32 x 32; layer 0: 10 x 2 tiles, hidden
// layers 2 x 2 each, output 2 x 1): its layer and row / column tile.
__device__ __forceinline__ void grad_tile(uint32_t tile, int& layer, int& rt, int& ct) {
    if (tile < 20) {
        layer = 0;
        rt = tile >> 1;
        ct = tile & 1;
    } else if (tile < 36) {
        layer = 1 + (tile - 20) / 4;
        rt = ((tile - 20) & 3) >> 1;
        ct = (tile - 20) & 1;
    } else {
        layer = 5;
        rt = tile - 36;
        ct = 0;
    }
}

// Master-weight index of accumulator register r of lane `lane` of a tile's
// product, or -1 (output layer columns past the third).
__device__ __forceinline__ int grad_tile_weight(int layer, int rt, int ct, int r, int lane) {
    const int h = lane >> 5;
    const int inr = rt * 32 + (r >> 2) * 8 + h * 4 + (r & 3);
    const int out = ct * 32 + (lane & 31);
    if (layer == 0) return kNrcW0 + inr * kNrcWidth + out;
    if (layer < 5) return kNrcWH + (layer - 1) * kNrcWidth * kNrcWidth + inr * kNrcWidth + out;
    return out < kNrcOut ? kNrcW5 + inr * kNrcOut + out : -1;
}

// One wave's partial dW of a tile over sample chunk `chunk` (K = samples,
// 16-byte operand loads straight from the feature-major layer inputs and
// pre-activation gradients).
__device__ __forceinline__ f32x16 grad_tile_chunk(const NrcDev& d, int layer, int rt, int ct,
                                                  uint32_t chunk, int lane) {
    const uint32_t S = d.batch;
    const int h = lane >> 5;
    const uint16_t* A = layer == 0 ? d.x0 : d.h + (size_t)(layer - 1) * kNrcWidth * S;
    const uint16_t* B = layer == 5 ? d.g : d.dz + (size_t)layer * kNrcWidth * S;
    const int arow = rt * 32 + (lane & 31);
    const int bcol = ct * 32 + (lane & 31);
    const bool bvalid = layer != 5 || bcol < kNrcOut;
    const size_t k0 = (size_t)chunk * kNrcChunk + 8 * h;
    const bf16x8* ap = reinterpret_cast<const bf16x8*>(A + (size_t)arow * S + k0);
    const bf16x8* bp = reinterpret_cast<const bf16x8*>(B + (size_t)(bvalid ? bcol : 0) * S + k0);
    f32x16 acc = f32x16{};
#pragma unroll 4
    for (int k = 0; k < kNrcChunk / 16; ++k) {
        const bf16x8 a = ap[2 * k];
        const bf16x8 b = bvalid ? bp[2 * k] : bf16x8{};
        acc = mfma(a, b, acc);
    }
    return acc;
}

__global__ __launch_bounds__(kNrcBlock) void k_nrc_dw(NrcDev d) {
    const uint32_t wave = (blockIdx.x * kNrcBlock + threadIdx.x) >> 6;
    const uint32_t chunks = d.batch / kNrcChunk;
    if (wave >= chunks * kNrcGradTiles) return;
    const uint32_t tile = wave % kNrcGradTiles, chunk = wave / kNrcGradTiles;
    int layer, rt, ct;
    grad_tile(tile, layer, rt, ct);
    const int lane = threadIdx.x & 63;
    const f32x16 acc = grad_tile_chunk(d, layer, rt, ct, chunk, lane);
    float* dst = d.partial + (size_t)chunk * kNrcNumWeights;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int w = grad_tile_weight(layer, rt, ct, r, lane);
        if (w >= 0) dst[w] = acc[r];
    }
}

// Adam on master weight w with gradient g; writes the weight's bf16
// fragment elements (k_nrc_pack's, one weight at a time).
__device__ __forceinline__ void adam_update(const NrcDev& d, const AdamParams& p, int w, float g) {
    const float m = p.beta1 * d.adam_m[w] + (1.0f - p.beta1) * g;
    const float v = p.beta2 * d.adam_v[w] + (1.0f - p.beta2) * (g * g);
    d.adam_m[w] = m;
    d.adam_v[w] = v;
    const float mhat = m / p.bias1;
    const float vhat = v / p.bias2;
    const float wn = d.master[w] - p.lr * mhat / (sqrtf(vhat) + p.eps);
    d.master[w] = wn;
    float we = wn;
    if (d.ema) {
        we = d.ema_decay * d.ema[w] + (1.0f - d.ema_decay) * wn;
        d.ema[w] = we;
    }
    const int32_t fp = d.fwd_pos[w], bp = d.bwd_pos[w];
    d.fwd[fp] = bits((__bf16)wn);
    if (bp >= 0) d.bwd[bp] = bits((__bf16)wn);
    if (d.fwd_pub) d.fwd_pub[fp] = bits((__bf16)we);
}

__global__ __launch_bounds__(kNrcBlock) void k_nrc_adam(NrcDev d, AdamParams p,
                                                        const uint32_t* __restrict__ count_ptr,
                                                        uint32_t step, int apply,
                                                        float* __restrict__ grad_out,
                                                        const float* __restrict__ grad_in) {
    const int w = blockIdx.x * kNrcBlock + threadIdx.x;
    if (w >= kNrcNumWeights) return;
    float g;
    if (grad_in) {
        g = grad_in[w];
    } else {
        if (*count_ptr <= step) return;  // empty batch: no step
        const uint32_t chunks = d.batch / kNrcChunk;
        g = 0.0f;
        for (uint32_t c = 0; c < chunks; ++c) g += d.partial[(size_t)c * kNrcNumWeights + w];
    }
    if (grad_out) grad_out[w] = g;
    if (apply) adam_update(d, p, w, g);
}

// Publishes the inference fragments of a step whose batch was empty (no
// update): the fragments of the unchanged weights.
__global__ __launch_bounds__(kNrcBlock) void k_nrc_publish(NrcDev d, const uint32_t* __restrict__ count_ptr,
                                                           uint32_t step) {
    const int w = blockIdx.x * kNrcBlock + threadIdx.x;
    if (w >= kNrcNumWeights || !d.fwd_pub || *count_ptr > step) return;
    d.fwd_pub[d.fwd_pos[w]] = bits((__bf16)(d.ema ? d.ema[w] : d.master[w]));
}

hipError_t nrc_launch_pack(const NrcDev& d, hipStream_t stream) {
    const int total = (kNrcFwdFrags + kNrcBwdFrags) * kFragElems;
    hipLaunchKernelGGL(k_nrc_pack, dim3((total + kNrcBlock - 1) / kNrcBlock), dim3(kNrcBlock), 0,
                       stream, d);
    return hipGetLastError();
}

static uint32_t infer_blocks(uint64_t n) {
    const uint64_t want = (n + kNrcInferWaves * 64 - 1) / (kNrcInferWaves * 64);
    const uint64_t cap = (uint64_t)device_cu_count() * kNrcInferBlocksPerCu;
    return (uint32_t)std::max<uint64_t>(1, std::min(want, cap));
}

hipError_t nrc_launch_infer(const NrcDev& d, const NrcQuery* q, uint32_t n, float4* out,
                            hipStream_t stream) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_nrc_infer<false>, dim3(infer_blocks(n)), dim3(kNrcInferBlock), 0, stream, d, q,
                       (const float4*)nullptr, (const uint32_t*)nullptr, n, out);
    return hipGetLastError();
}

hipError_t nrc_launch_infer_accumulate(const NrcDev& d, float4* pix_sum, hipStream_t stream) {
    hipLaunchKernelGGL(k_nrc_infer<true>, dim3(infer_blocks(d.query_cap)), dim3(kNrcInferBlock), 0,
                       stream, d, (const NrcQuery*)d.queries, (const float4*)d.query_weight,
                       (const uint32_t*)(d.ctl + kNrcCtlQueries), 0u, pix_sum);
    return hipGetLastError();
}

hipError_t nrc_launch_step(const NrcDev& d, const NrcQuery* in, const float4* target,
                           const uint32_t* count, uint32_t count_limit, uint32_t steps,
                           uint32_t step,
                           const AdamParams& adam, bool apply, float* grad_out,
                           const float* grad_in, hipStream_t stream) {
    if (!grad_in) {
        const uint32_t waves = d.batch / 32;
        hipLaunchKernelGGL(k_nrc_fwd_bwd, dim3((waves + 3) / 4), dim3(kNrcBlock), 0, stream, d, in,
                           target, count, count_limit, steps, step);
        const uint32_t gw = (d.batch / kNrcChunk) * kNrcGradTiles;
        hipLaunchKernelGGL(k_nrc_dw, dim3((gw + 3) / 4), dim3(kNrcBlock), 0, stream, d);
    }
    const dim3 wgrid((kNrcNumWeights + kNrcBlock - 1) / kNrcBlock);
    hipLaunchKernelGGL(k_nrc_adam, wgrid, dim3(kNrcBlock), 0, stream, d, adam, count, step,
                       apply ? 1 : 0, grad_out, grad_in);
    // k_nrc_adam re-packs the fragments of the weights it updates; a step
    // that found no records still publishes the inference fragments
    if (apply && d.fwd_pub && !grad_in)
        hipLaunchKernelGGL(k_nrc_publish, wgrid, dim3(kNrcBlock), 0, stream, d, count, step);
    return hipGetLastError();
}

}  // namespace lmb
