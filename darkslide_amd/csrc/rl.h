#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace ds {

// Taps of gaussian_blur_exact (image.rs:342-345) into k[0..2r]; returns r.
int gaussian_taps(float sigma, float* k);

// Workspace of raw_details in floats.
size_t rl_workspace_floats(size_t w, size_t h);

// raw_details(rgb, sigma, iters) -> rgb_out (enhance.rs:96-119).
// variant: tile shape of the iteration kernel, 0 .. RL_VARIANTS - 1 (same results)
constexpr int RL_VARIANTS = 3;
hipError_t launch_raw_details(const float* rgb, size_t w, size_t h, float sigma, size_t iters, float* ws, float* rgb_out, hipStream_t st,
                              int variant = 0);

}  // namespace ds
