// qmc_host.cpp - the cached tracer's stratified sampler (dev_qmc.h) built
// for the CPU, so that tests can check its points.  Test infrastructure:
// only tests/ load it.
#include <stdint.h>

#include "dev_qmc.h"

extern "C" {

// Unscrambled Sobol value (32-bit fraction) of index n in dimension 0, 1 or 2.
uint32_t qmc_sobol(uint32_t dim, uint32_t n) {
    if (dim == 0) return lmb::qmc_reverse_bits(n);
    return dim == 1 ? lmb::sobol_dim1(n) : lmb::sobol_dim2(n);
}

uint32_t qmc_owen(uint32_t x, uint32_t seed) { return lmb::owen_scramble(x, seed); }

// Points n0 .. n0 + count - 1 of stream `stream` (count x 3 floats).
void qmc_points(uint32_t n0, uint32_t count, uint32_t stream, float* out) {
    for (uint32_t i = 0; i < count; ++i) lmb::qmc_point3(n0 + i, stream, out + 3 * (uint64_t)i);
}
}
