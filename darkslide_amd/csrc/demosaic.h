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
    float black_site[4];  // per CFA site (index (y & 1) * 2 + (x & 1)) when per_site is set, else black everywhere
    int per_site;
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

// superpixel (demosaic.rs:44-87) of a's CFA into a.rgb (w/2 x h/2 x 3) and
// a.clip (w/2 x h/2).
hipError_t launch_superpixel(const DemosaicArgs& a, hipStream_t stream);

// The normalised CFA of a's input (decode.rs:300-303) into out (w x h).
hipError_t launch_cfa_normalize(const DemosaicArgs& a, float* out, hipStream_t stream);

}  // namespace ds
