// lmb_guided.h - per-pixel / per-path state of the guided (reference) mode
// and its launch entry point.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "lmb_frame_dev.h"
#include "lmb_layout.h"
#include "lmb_scene_dev.h"

namespace lmb {

struct GuidedDev {
    // per pixel, across samples
    MarkovChain* tent;       // tentative chain during the sample
    MarkovChain* tent_snap;  // tentative chain at sample start (after the shuffle)
    float* sum_mc;           // sum of candidate weights of the last resampling
    uint32_t* cand_seed;     // seed the screen-space candidate draws started from
    float4* out_rad_sum;     // rt_main's out_radiance accumulator
    Reservoir* pix_res;      // rt_main's reservoir R
    Reservoir* pix_sample;   // rt_main's per-pixel Sample (RESTIR per pixel)
    float4* pix_last_normal;    // xyz, w = ty bits
    float4* pix_last_albedo;
    float4* pix_last_emission;
    // per path (rt_sample locals)
    float4* color;  // xyz throughput, w picked IOR parameter
    float4* scol;   // xyz sample colour, w cached_pdf
    float4* orc;    // xyz out_radiance_colour, w cached_brdf
    float4* sout;   // xyz sample out_radiance, w flags
    float4* aem;    // accum_emission
    float4* aalb;   // accum_albedo
    float4* lastn;  // last_normal
    float4* lastd;  // last_dir
    // per path, sample fields of SampleReturn
    float4* s_point;       // xyz, w = ty bits
    float4* s_normal;      // xyz, w = pdf
    float4* s_albedo;
    float4* s_emission;
    float4* s_vis_point;
    float4* s_vis_normal;
    float4* s_start_dir;
    // per queue slot: hit record handed from shade to guide
    float4* gr0;  // geometric normal, roughness.x
    float4* gr1;  // tangent, refractive index
    float4* gr2;  // chosen normal, meta bits (type | reflect << 8 | done << 16)
    float4* gr3;  // hit colour
};

// Number of float4 / 4-byte arrays above, for allocation.
constexpr int kGuidedFloat4PerPixel = 24;

hipError_t launch_guided(const SceneDev& s, const CameraDev& cam, const FrameDev& f,
                         const GuidedDev& g, const TraceParams& p, hipStream_t stream,
                         KernelTimer* timer);

}  // namespace lmb
