#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "darkslide_gpu.h"

namespace ds {

struct DemosaicArgs {
    const float* cfa_f32;     // normalised CFA (device), or null
    const uint16_t* raw_u16;  // raw DN (device), or null
    size_t stride;            // elements per input row
    size_t w, h;
    uint8_t pattern[4];
    float black;
    float white[3];
    float mul[3];
    int algo;     // DS_ALGO_RCD / DS_ALGO_MHC
    float* rgb;   // interleaved out (device)
    float* clip;  // clip plane out (device) or null
};

// variant: tile shape, 0 .. demosaic_variants(algo) - 1 (same results)
int demosaic_variants(int algo);
hipError_t launch_demosaic(const DemosaicArgs& a, hipStream_t stream, int variant = 0);

// cfa_clip_levels (demosaic.rs:419-430) of a's CFA (rgb, clip and algo are
// not used) into out3 (device, 3 floats); partial: CLIP_PARTS * 3 floats.
constexpr int CLIP_PARTS = 256;
hipError_t launch_cfa_clip_levels(const DemosaicArgs& a, float* partial, float* out3, hipStream_t stream);

}  // namespace ds
