// dev_compact.h - queue compaction with one atomic per block and counter.
//
// Shading kernels append survivors to the next bounce's queue (and the cached
// tracer its queries and training records) through device-wide counters.
// Atomics on one address are serialized where the device keeps them coherent;
// with one atomic per wave, a 2M-path launch spends a large part of its time
// queued on them.  Reserving per block of BLOCK threads cuts their number by
// BLOCK / 64, for up to three counters with the same two barriers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lmb {

template <int BLOCK, int N>
struct ReserveLds {
    uint32_t w[N][BLOCK / 64 + 1];  // per wave: its offset in the block; [W]: the block's base
};

// Every thread with pred[k] gets a distinct index idx[k] from *counter[k]
// (k < N), in thread order within the block.  Every thread of the block must
// call it the same number of times (block-uniform trip counts).
template <int BLOCK, int N>
__device__ __forceinline__ void block_reserve(const bool (&pred)[N], uint32_t* const (&counter)[N],
                                              ReserveLds<BLOCK, N>& lds, uint32_t (&idx)[N]) {
    static_assert(BLOCK % 64 == 0 && N >= 1 && N <= 3, "block_reserve shape");
    constexpr int W = BLOCK / 64;
    const uint32_t lane = __lane_id(), wave = threadIdx.x / 64u;
    uint64_t mask[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        mask[k] = __ballot(pred[k]);
        if (lane == 0) lds.w[k][wave] = (uint32_t)__popcll(mask[k]);
    }
    __syncthreads();
    if (threadIdx.x < (uint32_t)N) {  // one thread per counter, all in wave 0
        const uint32_t k = threadIdx.x;
        uint32_t* const c = k == 0 ? counter[0] : (k == 1 ? counter[N > 1 ? 1 : 0] : counter[N - 1]);
        uint32_t total = 0;
        for (int v = 0; v < W; ++v) {
            const uint32_t n = lds.w[k][v];
            lds.w[k][v] = total;
            total += n;
        }
        lds.w[k][W] = total ? atomicAdd(c, total) : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; ++k)
        idx[k] = lds.w[k][W] + lds.w[k][wave] + (uint32_t)__popcll(mask[k] & ((1ull << lane) - 1ull));
    __syncthreads();  // the scratch is reused by the next call
}

// Trip count of a grid-stride loop over `count` items that is the same for
// every thread of a block of BLOCK threads.
template <int BLOCK>
__device__ __forceinline__ uint64_t block_uniform_count(uint32_t count) {
    return ((uint64_t)count + BLOCK - 1) / BLOCK * BLOCK;
}

}  // namespace lmb
