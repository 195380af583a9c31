#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "common.h"

namespace ds {

// NoiseReduction (params.rs:319-326) and Sharpening (params.rs:304-309), in
// the reference's 0..100 units.
struct NrParams {
    float luminance, luminance_detail, luminance_contrast, color, color_detail, color_smoothness;
};
struct SharpParams {
    float amount, radius, detail, masking;
};

// Workspace floats of launch_noise_reduction / launch_sharpen.
size_t nr_ws_floats(size_t w, size_t h);
size_t sharpen_ws_floats(size_t w, size_t h);

// detail::noise_reduction (pipeline/detail.rs:13-92) in place on
// interleaved ProPhoto.  field: the local noise plane (w*h) or null;
// field_max: its maximum (the fold at detail.rs:34), read by the caller.
// The early returns and every radius / eps are decided on the host from
// the same f32 expressions as the reference.
hipError_t launch_noise_reduction(float* rgb, size_t w, size_t h, const NrParams& nr, const float* field, float field_max, float scale,
                                  const OklabM& ok, float* ws, hipStream_t st);

// detail::sharpen (pipeline/detail.rs:96-139) in place.  field: the local
// sharpness plane or null.
hipError_t launch_sharpen(float* rgb, size_t w, size_t h, const SharpParams& sh, const float* field, float scale, float* ws,
                          hipStream_t st);

// The export post-steps after render_full (export.rs:244-247), in place on
// interleaved ProPhoto.  ws: export_detail_ws_floats(w, h) floats.
size_t export_detail_ws_floats(size_t w, size_t h);
// enhance.rs:122-131 super_res_refine
hipError_t launch_super_res_refine(float* rgb, size_t w, size_t h, float* ws, hipStream_t st);
// export.rs:199-223 output_sharpen; mode SharpenFor (export.rs:63-68),
// amount SharpenAmount (export.rs:71-75), numbered in declaration order.
enum { OUT_SHARPEN_NONE = 0, OUT_SHARPEN_SCREEN = 1, OUT_SHARPEN_MATTE = 2, OUT_SHARPEN_GLOSSY = 3 };
enum { OUT_AMOUNT_LOW = 0, OUT_AMOUNT_STANDARD = 1, OUT_AMOUNT_HIGH = 2 };
hipError_t launch_output_sharpen(float* rgb, size_t w, size_t h, int mode, int amount, float* ws, hipStream_t st);

// max(0, max of field) (the fold at detail.rs:34) into out (1 device float);
// partial: FIELD_MAX_PARTS floats.
constexpr int FIELD_MAX_PARTS = 256;
hipError_t launch_field_max(const float* field, size_t n, float* partial, float* out, hipStream_t st);

// Plane::resize (image.rs:238-262) and fast_guided_filter (image.rs:408-431),
// also used on their own by the C ABI.  fgf ws: fgf_ws_floats(w, h, sub).
hipError_t launch_plane_resize(const float* in, size_t w, size_t h, size_t w2, size_t h2, float* out, hipStream_t st);
size_t fgf_ws_floats(size_t w, size_t h, size_t sub);
hipError_t launch_fast_guided_filter(const float* I, const float* P, size_t w, size_t h, size_t r, float eps, size_t sub, float* ws,
                                     float* out, hipStream_t st);
// The chroma step of noise_reduction (detail.rs:79-82): A and B, in place,
// each mixed by k towards fast_guided_filter(I, plane, r, eps, sub).  The
// two filters share the guide, so its resized copy and the means of it and
// of its square (bit-identical in the two calls) are formed once.
// ws: fgf_pair_ws_floats(w, h, sub).
size_t fgf_pair_ws_floats(size_t w, size_t h, size_t sub);
hipError_t launch_fgf_pair_mix(const float* I, float* A, float* B, size_t w, size_t h, size_t r, float eps, size_t sub, float k, float* ws,
                               hipStream_t st);

}  // namespace ds
