#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace ds {

// Parameter layout (FP32 master copy, one flat array):
//   [0, 320)        input layer   W_in[co 32][tap 9], b_in[32]
//   [320, 55808)    6 middle layers, each W[co 32][tap 9][ci 32] then b[32]
//   [55808, 56097)  output layer  W_out[tap 9][ci 32], b_out
constexpr size_t CNN_OFF_IN = 0;
constexpr size_t CNN_OFF_MID = 320;
constexpr size_t CNN_MID_STRIDE = 32 * 9 * 32 + 32;
constexpr size_t CNN_OFF_OUT = CNN_OFF_MID + 6 * CNN_MID_STRIDE;
constexpr size_t CNN_NPARAM = CNN_OFF_OUT + 9 * 32 + 1;
// bf16 copies of the weights that the MFMA convolutions use, one array:
//   [0, 6*9216)        middle layers, forward layout [co][tap][ci]
//   [6*9216, 12*9216)  middle layers, data-gradient layout [ci][8 - tap][co]
//   CNN_WPK_IN         input layer as a 32 x 32 A operand [co][k], k = tap (k >= 9: 0)
//   CNN_WPK_OUT        output layer [tap][ci]
constexpr size_t CNN_WMID = 32 * 9 * 32;
constexpr size_t CNN_WPK_BWD = 6 * CNN_WMID;
constexpr size_t CNN_WPK_IN = 12 * CNN_WMID;
constexpr size_t CNN_WPK_OUT = CNN_WPK_IN + 32 * 32;
constexpr size_t CNN_WPK = CNN_WPK_OUT + 9 * 32;

size_t cnn_param_count();
hipError_t cnn_pack(const float* params, uint16_t* wpk, hipStream_t st);
hipError_t cnn_forward(const float* params, const uint16_t* wpk, const float* x, int n_img, int H, int W, uint16_t* const* acts, float* out,
                       hipStream_t st);
// Inference forward in two fused streaming passes (activation a4 in HBM,
// cnn_fused_act_bytes); equals cnn_forward bit for bit.  With lum set, out
// receives lum + max(*sigma_dev, 1e-4) (y - x), the denoised luminance,
// instead of y.
size_t cnn_fused_act_bytes(int n_img, int H, int W);
hipError_t cnn_forward_fused(const float* params, const uint16_t* wpk, const float* x, int n_img, int H, int W, uint16_t* act, float* out,
                             hipStream_t st, const float* lum = nullptr, const float* sigma_dev = nullptr);
size_t cnn_wgrad_blocks(int n_img, int H, int W);
size_t cnn_scratch_floats(int n_img, int H, int W);
hipError_t cnn_backward(const float* params, const uint16_t* wpk, const float* x, const float* tgt, int n_img, int H, int W, int margin,
                        uint16_t* const* acts, const float* out, uint16_t* const* gacts, float* scratch, float* grad, float* loss_part_out,
                        hipStream_t st);
hipError_t cnn_adam(float* params, const float* grad, float* m, float* v, int step, float lr, hipStream_t st);
hipError_t cnn_input(const float* L, int w, int h, const float* sigma_dev, float* x, hipStream_t st);
hipError_t cnn_target(const float* x, const float* L, const float* LF, size_t n, const float* sigma_dev, float* t, hipStream_t st);
hipError_t cnn_crops(const float* a, const float* b, int w, int h, int c, int n, uint64_t seed, float* ca, float* cb, hipStream_t st);

}  // namespace ds
