// lmb_restir.h - launch entry points of the ReSTIR and end-of-frame kernels.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "lmb_frame_dev.h"
#include "lmb_layout.h"

namespace lmb {

hipError_t launch_temporal(const FrameDev& f, uint32_t seed_offset, hipStream_t stream);
hipError_t launch_spatial(const FrameDev& f, uint32_t is_samples, uint32_t is_space,
                          uint32_t seed_offset, hipStream_t stream);
hipError_t launch_eof_gi(Reservoir* res, uint64_t n, uint32_t cap, hipStream_t stream);
hipError_t launch_eof_world(WorldMarkovChain* w, uint64_t n, hipStream_t stream);
// Zeroes `bytes` (a multiple of 4) with a kernel.
hipError_t launch_fill_zero(void* p, uint64_t bytes, hipStream_t stream);

}  // namespace lmb
