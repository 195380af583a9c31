#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "nlm.h"

namespace ds {

// Workspace (floats) of launch_nlm_reference_order: three planes per chunk.
size_t nlm_ref_workspace_floats(size_t w, size_t h, long search, size_t patch, int threads);

// nlm(p, h, search, patch) (enhance.rs:34-72) in the reference's summation
// order for a rayon pool of `threads` threads: bit for bit the CPU path's
// result.  search >= 0 (search < 0 leaves out untouched).
hipError_t launch_nlm_reference_order(const float* in, float* out, size_t w, size_t h, NlmH hp, long search, size_t patch, int threads,
                                      float* ws, hipStream_t st);

}  // namespace ds
