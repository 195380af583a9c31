#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace ds {

// Workspace of launch_reconstruct_highlights: a 64-float header (the
// clipped-anywhere flag) and 9 planes.
size_t highlights_ws_floats(size_t w, size_t h);

// reconstruct_highlights (demosaic.rs:435-502) in place on interleaved RGB.
// mul: host values; raw_clip: 3 device floats (cfa_clip_levels' output, so
// the pipeline never waits on the host).  Every stage reads a device flag
// set when some clip value is positive and does nothing otherwise, which is
// the reference's early return without a host round trip.
hipError_t launch_reconstruct_highlights(float* rgb, const float* clip, size_t w, size_t h, const float mul[3], const float* raw_clip,
                                         float* ws, hipStream_t st);

}  // namespace ds
