// Per-pixel steps of the guided filter pair of denoise (image.rs:384-405,
// enhance.rs:86-91), shared by the pointwise kernels and the fused vertical
// box passes so that both compute the same operations in the same order.
#pragma once

#include "common.h"

namespace ds {

// eps = (2 max(sigma, 1e-4))^2 (enhance.rs:82, :88)
__device__ __forceinline__ float gf_eps(const float* sigma_dev) {
#pragma clang fp contract(off)
    const float s2 = fmaxr(*sigma_dev, 1e-4f) * 2.0f;
    return s2 * s2;
}

// image.rs:395-400 for the a and b planes from the six means of I, a, b,
// I a, I b, I I: a_x = cov_x / (var + eps), b_x = mean_x - a_x mean_I.
__device__ __forceinline__ void gf_coef_px(float mi, float mpa, float mpb, float mipa, float mipb, float mii, float eps, float& aa, float& ba,
                                           float& ab, float& bb) {
#pragma clang fp contract(off)
    const float var = mii - mi * mi;
    const float cov_a = mipa - mi * mpa;
    aa = cov_a / (var + eps);
    ba = mpa - aa * mi;
    const float cov_b = mipb - mi * mpb;
    ab = cov_b / (var + eps);
    bb = mpb - ab * mi;
}

// image.rs:404 mean(a) I + mean(b) for both planes, then Oklab -> ProPhoto
// and max(0) (enhance.rs:91) into rgb[0..2].
__device__ __forceinline__ void gf_final_px(float I, float ma_a, float mb_a, float ma_b, float mb_b, const OklabM& ok, float* rgb) {
#pragma clang fp contract(off)
    float lab[3] = {I, ma_a * I + mb_a, ma_b * I + mb_b};
    float l[3], cube[3], p[3];
    mul_mv(ok.m2_inv, lab, l);
    cube[0] = l[0] * l[0] * l[0];
    cube[1] = l[1] * l[1] * l[1];
    cube[2] = l[2] * l[2] * l[2];
    mul_mv(ok.from_lms, cube, p);
    rgb[0] = fmaxr(p[0], 0.0f);
    rgb[1] = fmaxr(p[1], 0.0f);
    rgb[2] = fmaxr(p[2], 0.0f);
}

}  // namespace ds
