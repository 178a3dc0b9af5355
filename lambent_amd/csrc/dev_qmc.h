// dev_qmc.h - stratified sample points for the first scatter of the cached
// tracer: the first three dimensions of the Sobol sequence (Sobol 1967), in
// natural index order, under Owen's nested uniform scrambling in the
// hash-based form of Burley ("Practical Hash-based Owen Scrambling", JCGT
// 9(4), 2020), with the Laine-Karras-style permutation hash of N. Vegdahl
// ("Building a Better LK Hash", 2021).
//
// Dimension 0 is the base-2 radical inverse of the index.  Dimensions 1 and
// 2 use the primitive polynomials x + 1 and x^2 + x + 1 with initial
// direction numbers m = 1 and m = 1, 3; their direction numbers
// v_k = m_k 2^(32-k) follow the polynomials' recurrences written on v:
// v_k = v_(k-1) ^ (v_(k-1) >> 1) and v_k = v_(k-1) ^ v_(k-2) ^ (v_(k-2) >> 2).
// Dimensions 0 and 1 form a (0, 2)-sequence: the first 2^m points put one
// point in every elementary box of area 2^-m.  Owen scrambling keeps that
// and makes every single point uniform on [0, 1)^3, so an estimator that
// uses one point per sample stays unbiased.
//
// Host and device code: the tests compile this header on the CPU.
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#define LMB_QMC_HD __host__ __device__ __forceinline__
#else
#define LMB_QMC_HD inline
#endif

namespace lmb {

LMB_QMC_HD uint32_t qmc_reverse_bits(uint32_t x) {
    x = ((x >> 1) & 0x55555555u) | ((x & 0x55555555u) << 1);
    x = ((x >> 2) & 0x33333333u) | ((x & 0x33333333u) << 2);
    x = ((x >> 4) & 0x0F0F0F0Fu) | ((x & 0x0F0F0F0Fu) << 4);
    x = ((x >> 8) & 0x00FF00FFu) | ((x & 0x00FF00FFu) << 8);
    return (x >> 16) | (x << 16);
}

// Sobol dimension 1 (x + 1, m_1 = 1).
LMB_QMC_HD uint32_t sobol_dim1(uint32_t n) {
    uint32_t x = 0;
    for (uint32_t v = 1u << 31; n; n >>= 1, v ^= v >> 1)
        if (n & 1u) x ^= v;
    return x;
}

// Sobol dimension 2 (x^2 + x + 1, m_1 = 1, m_2 = 3).
LMB_QMC_HD uint32_t sobol_dim2(uint32_t n) {
    uint32_t x = 0, cur = 1u << 31, nxt = 3u << 30;
    for (; n; n >>= 1) {
        if (n & 1u) x ^= cur;
        const uint32_t v = nxt ^ cur ^ (cur >> 2);
        cur = nxt;
        nxt = v;
    }
    return x;
}

// A permutation of 32-bit words in which every bit depends only on itself
// and the bits below it (each step: xor with an even multiple, add, odd
// multiply), so that applied to bit-reversed values it is a nested uniform
// scramble: the constants are Vegdahl's.
LMB_QMC_HD uint32_t lk_permute(uint32_t x, uint32_t seed) {
    x ^= x * 0x3d20adeau;
    x += seed;
    x *= (seed >> 16) | 1u;
    x ^= x * 0x05526c56u;
    x ^= x * 0x53a22864u;
    return x;
}

LMB_QMC_HD uint32_t owen_scramble(uint32_t x, uint32_t seed) {
    return qmc_reverse_bits(lk_permute(qmc_reverse_bits(x), seed));
}

// C. Wellons' "lowbias32" integer hash, for the scrambling seeds.
LMB_QMC_HD uint32_t qmc_hash(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352du;
    x ^= x >> 15;
    x *= 0x846ca68bu;
    x ^= x >> 16;
    return x;
}

// 24-bit float in [0, 1).
LMB_QMC_HD float qmc_unit(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }

// Point n of the scrambled sequence of `stream` (a pixel): three
// independently scrambled dimensions.
LMB_QMC_HD void qmc_point3(uint32_t n, uint32_t stream, float out[3]) {
    const uint32_t base = qmc_hash(stream ^ 0xA511E9B3u);
    out[0] = qmc_unit(owen_scramble(qmc_reverse_bits(n), qmc_hash(base)));
    out[1] = qmc_unit(owen_scramble(sobol_dim1(n), qmc_hash(base + 0x9E3779B9u)));
    out[2] = qmc_unit(owen_scramble(sobol_dim2(n), qmc_hash(base + 0x3C6EF372u)));
}

}  // namespace lmb
