// lmb_batch.h - device view of a batch of scenes rendered by one set of
// persistent wavefront launches, and its launcher.
#pragma once

#include <hip/hip_runtime_api.h>
#include <stdint.h>

#include "lmb_frame_dev.h"
#include "lmb_layout.h"
#include "lmb_scene_dev.h"

namespace lmb {

// One image of the batch.  Its pixels are pid = pix_base + y * width + x of
// the batch's per-pixel arrays; its tiles are the global tiles
// [tile_base, tile_base + tiles_x * tiles_y).
struct BatchItemDev {
    CameraDev cam;
    uint32_t width, height;
    uint64_t pix_base;
    uint64_t tile_base;
    uint32_t tiles_x, tiles_y;
};

struct BatchDev {
    const SceneDev* scenes;  // per item
    const BatchItemDev* items;
    uint32_t num_items;
    uint32_t tile_w, tile_h;
    uint64_t tiles_total;
    const uint32_t* seed;  // device word: seed_offset of the frame
};

struct BatchLaunch {
    uint32_t blocks;      // persistent grid of every stage
    uint32_t refill;      // traversal refill threshold (idle lanes)
    uint64_t pass_tiles;  // tiles per wavefront pass
};

// One frame, then colour = max(pix_sum / samples, 0) for every pixel.
// Capturable.  With `fg` (per-path buffers for fg->group samples of every
// pixel: path id pixel * group + j) and one pass over all tiles, the camera
// rays are generated and traversed once and the samples are traced in
// groups of fg->group, as the single-image tracer does; otherwise, for each
// pass of pass_tiles tiles, every sample's wavefronts one sample at a time.
// Both give the same images.
hipError_t launch_batch(const BatchDev& b, const FrameDev& f, const FrameDev* fg,
                        const TraceParams& p, const BatchLaunch& l, hipStream_t stream);

}  // namespace lmb
