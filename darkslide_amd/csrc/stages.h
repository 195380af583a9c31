#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "common.h"

namespace ds {

void box_radii_for_gauss(float sigma, size_t radii[3]);
// gaussian_blur(in, sigma) (image.rs:327); ws: 2 planes.
hipError_t launch_gaussian_blur(const float* in, float* ws, float* out, size_t w, size_t h, float sigma, hipStream_t st);
// guided_filter(I, p, r, eps) (image.rs:385); ws: 8 planes.
// launch_guided_coeffs stops before the last step (image.rs:404): the
// box-blurred a and b are left in ws planes 4 and 5.
hipError_t launch_guided_coeffs(const float* I, const float* P, size_t w, size_t h, size_t r, float eps, float* ws, hipStream_t st);
hipError_t launch_guided_filter(const float* I, const float* P, size_t w, size_t h, size_t r, float eps, float* ws, float* out,
                                hipStream_t st);
hipError_t launch_oklab_fwd_interleaved(const float* rgb, size_t n, const OklabM& ok, float* lab, hipStream_t st);
hipError_t launch_oklab_inv(const float* lab, size_t n, const OklabM& ok, float* rgb, hipStream_t st);
hipError_t launch_luminance(const float* rgb, size_t n, float* out, hipStream_t st);
// The order-free hash (libm_hash_term) of ds_cbrtf (fn 0) or ds_expf (fn 1)
// over all 2^32 inputs, added into *acc (zeroed here).
hipError_t launch_libm_hash(int fn, unsigned long long* acc, hipStream_t st);

}  // namespace ds
