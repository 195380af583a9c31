#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "common.h"

namespace ds {

hipError_t launch_oklab_fwd(const float* rgb, size_t n, const OklabM& ok, float* L, float* A, float* B, hipStream_t st);

// Samples estimate_noise reads from a w x h plane.
size_t noise_sample_count(size_t w, size_t h);
// Scratch of launch_estimate_noise in u32: a first-level histogram and the samples.
size_t noise_scratch_u32(size_t w, size_t h);
// The largest scratch any plane needs: the histograms and control words
// (NOISE_HDR_U32) plus twice the sample count ceil(n / max(n / 200000, 1)),
// which peaks at 399,999 (n = 399,999, step 1; above 400,000 pixels the
// step grows and the count stays below 300,000).
constexpr size_t NOISE_HDR_U32 = 4096 + 1024 + 64;
constexpr size_t NOISE_MAX_SAMPLES = 399999;
constexpr size_t NOISE_SCRATCH_MAX_U32 = NOISE_HDR_U32 + 2 * NOISE_MAX_SAMPLES;
// Number of samples estimate_noise keeps (off the 1-pixel border) for a
// w x h plane sampled every step pixels.
size_t noise_valid_samples(size_t w, size_t h, size_t step);
// estimate_noise(plane) into *sigma_out (device).  scratch: noise_scratch_u32 u32.
hipError_t launch_estimate_noise(const float* plane, size_t w, size_t h, uint32_t* scratch, float* sigma_out, hipStream_t st);

// The two guided filters of denoise (enhance.rs:86-91) and the inverse Oklab:
// I = L + (LF - L) m; a2 = gf(I, A), b2 = gf(I, B) with eps = (2 max(sigma,1e-4))^2;
// rgb = max(Oklab^-1(I, a2, b2), 0).  ws: 12 planes of w*h floats.  LF is
// overwritten (a spare plane once the guide is formed).  guide_ready: the
// first four planes of ws already hold I, I A, I B and I I (formed by
// launch_nlm_guide), and LF is only a spare plane.
hipError_t launch_guided_pair(const float* L, float* LF, const float* A, const float* B, float m, size_t r, const float* sigma_dev,
                              const OklabM& ok, float* ws, float* rgb, size_t w, size_t h, hipStream_t st, bool guide_ready = false);

}  // namespace ds
