// Learned luminance denoiser: a residual CNN that stands in for the NLM
// stage of denoise.  It has no counterpart in the reference; its training
// targets are the NLM outputs of this library.
//
// Network (input x, 1 channel; 32 channels inside; 8 conv3x3 layers):
//     a1 = relu(conv_in(x))            1 -> 32   (FP32 VALU)
//     a{k+1} = relu(conv_k(a_k))       32 -> 32, k = 1..6   (BF16 MFMA)
//     out = x + conv_out(a7)           32 -> 1   (FP32 VALU, residual epilogue)
// with zero padding.  Activations are NHWC bf16 (32 channels = 64 bytes per
// pixel); every accumulation is FP32.
//
// The 32 -> 32 convolutions are implicit GEMMs on v_mfma_f32_16x16x32_bf16:
// D[co][px] = sum over (tap, ci) of W[co][tap][ci] * X[px + tap][ci], with the
// weights (A operand, 16 x 32 per MFMA) held in registers for the whole
// block and the input tile plus its 1-pixel halo (B operand) in LDS.  The
// data-gradient pass is the same kernel with the weights flipped and
// transposed; the weight gradient is a GEMM over pixels through an im2col
// tile in LDS.
//
// MFMA 16x16x32 operand layout (lane l of the wave, g = l / 16):
//   A: lane holds A[m = l % 16][k = 8g .. 8g + 7]
//   B: lane holds B[k = 8g .. 8g + 7][n = l % 16]
//   D: lane holds D[m = 4g + i][n = l % 16], i = 0..3

#include <type_traits>

#include "cnn.h"
#include "common.h"

namespace ds {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int C = 32;       // channels
constexpr int CP = 40;      // LDS pixel stride in bf16 (80 bytes: 16-byte aligned, spreads banks)
constexpr int TW = 64;      // output tile width
constexpr int TH = 4;       // output tile rows (one per wave)
constexpr int NT = 256;

__device__ __forceinline__ uint16_t f2bf(float f) {
    __bf16 b = (__bf16)f;  // round to nearest even
    return __builtin_bit_cast(uint16_t, b);
}
__device__ __forceinline__ float bf2f(uint16_t u) { return __uint_as_float((uint32_t)u << 16); }

// ---------------------------------------------------------------- input / output layers (MFMA)

// Input layer: a1[co][px] = relu(b_in[co] + sum_tap W_in[co][tap] x[px + tap]),
// one v_mfma_f32_16x16x32_bf16 per 16 x 16 tile with K = 32 (taps 0..8, the
// rest zero); x is rounded to bf16 as the MFMA's B operand.  Tile 64 x 4, one
// row per wave; x tile with a 1-pixel halo in LDS (zero outside the image).
__global__ __launch_bounds__(256) void k_conv_in(const float* __restrict__ x, const uint16_t* __restrict__ win, const float* __restrict__ b_in,
                                                 uint16_t* __restrict__ act, int H, int Wd) {
    constexpr int TW = 64, TH = 4;
    __shared__ float s_x[(TH + 2) * (TW + 2)];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, g = lane >> 4, ln = lane & 15;
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
    const size_t ioff = (size_t)blockIdx.z * H * Wd;
    for (int i = tid; i < (TH + 2) * (TW + 2); i += 256) {
        const int gy = y0 + i / (TW + 2) - 1, gx = x0 + i % (TW + 2) - 1;
        s_x[i] = (gy >= 0 && gy < H && gx >= 0 && gx < Wd) ? x[ioff + (size_t)gy * Wd + gx] : 0.0f;
    }
    bf16x8 wa[2];
#pragma unroll
    for (int mt = 0; mt < 2; mt++) wa[mt] = *reinterpret_cast<const bf16x8*>(win + (mt * 16 + ln) * 32 + 8 * g);
    __syncthreads();
    const int oy = y0 + wave;
#pragma unroll
    for (int nt = 0; nt < TW / 16; nt++) {
        bf16x8 bv;
#pragma unroll
        for (int e = 0; e < 8; e++) {
            const int t = 8 * g + e;
            const float v = t < 9 ? s_x[(wave + t / 3) * (TW + 2) + nt * 16 + ln + t % 3] : 0.0f;
            bv[e] = (__bf16)v;
        }
        const int ox = x0 + nt * 16 + ln;
#pragma unroll
        for (int mt = 0; mt < 2; mt++) {
            const f32x4 acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[mt], bv, f32x4{0, 0, 0, 0}, 0, 0, 0);
            if (oy < H && ox < Wd) {
                const int co = mt * 16 + 4 * g;
                uint2 o;
                o.x = f2bf(fmaxf(acc[0] + b_in[co], 0.0f)) | ((uint32_t)f2bf(fmaxf(acc[1] + b_in[co + 1], 0.0f)) << 16);
                o.y = f2bf(fmaxf(acc[2] + b_in[co + 2], 0.0f)) | ((uint32_t)f2bf(fmaxf(acc[3] + b_in[co + 3], 0.0f)) << 16);
                *reinterpret_cast<uint2*>(act + (ioff + (size_t)oy * Wd + ox) * C + co) = o;
            }
        }
    }
}

__device__ __forceinline__ void load_px32(const uint16_t* p, float* v) {
    const uint4* s = reinterpret_cast<const uint4*>(p);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uint4 u = s[k];
        uint32_t w4[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
        for (int j = 0; j < 4; j++) {
            v[8 * k + 2 * j] = bf2f((uint16_t)(w4[j] & 0xffff));
            v[8 * k + 2 * j + 1] = bf2f((uint16_t)(w4[j] >> 16));
        }
    }
}

// Output layer: y[px] = x[px] + (((E_0[px - 1] + E_1[px]) + E_2[px + 1]) + b_out),
// where E_kx[q] = sum over (ky, ci) of W_out[ky][kx][ci] a7[q + (ky - 1) rows][ci].
// Three MFMAs (one per kernel row ky, K = 32 channels, a chain) with
// W_out[ky][kx] in row kx of the A operand give E_0..E_2 of 16 pixels at once;
// the fused inference kernel sums in the same order.  Tile and a7 staging as
// k_conv_mid.
constexpr int CP_ = 40;
__global__ __launch_bounds__(256) void k_conv_out(const uint16_t* __restrict__ act, const float* __restrict__ x, const uint16_t* __restrict__ wout,
                                                  const float* __restrict__ b_out, float* __restrict__ out, int H, int Wd) {
    constexpr int TW = 64, TH = 4, NE = TW + 2;  // E columns x0 - 1 .. x0 + TW
    constexpr int NET = (NE + 15) / 16;           // E tiles of 16 columns
    __shared__ __attribute__((aligned(16))) uint16_t s_in[(TH + 2) * (TW + 2) * CP_];
    __shared__ float s_e[TH][3][NET * 16];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, g = lane >> 4, ln = lane & 15;
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
    const size_t ioff = (size_t)blockIdx.z * H * Wd;
    {
        constexpr int N16 = (TH + 2) * (TW + 2) * (C / 8);
        constexpr int K = (N16 + 255) / 256;
        uint4 r[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int i = tid + k * 256;
            const int pix = i >> 2, part = i & 3;
            const int gy = y0 + pix / (TW + 2) - 1, gx = x0 + pix % (TW + 2) - 1;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (i < N16 && gy >= 0 && gy < H && gx >= 0 && gx < Wd) v = reinterpret_cast<const uint4*>(act + (ioff + (size_t)gy * Wd + gx) * C)[part];
            r[k] = v;
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int i = tid + k * 256;
            if (i < N16) reinterpret_cast<uint4*>(s_in + (i >> 2) * CP_)[i & 3] = r[k];
        }
    }
    bf16x8 wa[3];
#pragma unroll
    for (int ky = 0; ky < 3; ky++) {
        bf16x8 z = {};
        if (ln < 3) z = *reinterpret_cast<const bf16x8*>(wout + (ky * 3 + ln) * C + 8 * g);
        wa[ky] = z;
    }
    const float bo = b_out[0];
    __syncthreads();
#pragma unroll
    for (int nt = 0; nt < NET; nt++) {
        const int j = nt * 16 + ln;  // E column x0 - 1 + j
        f32x4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int ky = 0; ky < 3; ky++) {
            bf16x8 bv = {};
            if (j < NE) bv = *reinterpret_cast<const bf16x8*>(s_in + ((wave + ky) * (TW + 2) + j) * CP_ + 8 * g);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[ky], bv, acc, 0, 0, 0);
        }
        if (g == 0) {
#pragma unroll
            for (int kx = 0; kx < 3; kx++) s_e[wave][kx][j] = acc[kx];
        }
    }
    __syncthreads();
    const int ox = x0 + lane, oy = y0 + wave;
    if (oy < H && ox < Wd) {
        const size_t p = ioff + (size_t)oy * Wd + ox;
        out[p] = x[p] + (((s_e[wave][0][lane] + s_e[wave][1][lane + 1]) + s_e[wave][2][lane + 2]) + bo);
    }
}

// ---------------------------------------------------------------- 32 -> 32 conv on MFMA

// MODE 0: forward, epilogue relu(acc + bias).
// MODE 1: data gradient, epilogue acc * (mask > 0) (mask = the layer input's activation).
template <int MODE>
__global__ __launch_bounds__(NT) void k_conv_mid(const uint16_t* __restrict__ in, const uint16_t* __restrict__ wgt,
                                                 const float* __restrict__ bias, const uint16_t* __restrict__ mask,
                                                 uint16_t* __restrict__ out, int H, int Wd) {
    __shared__ __attribute__((aligned(16))) uint16_t s_in[(TH + 2) * (TW + 2) * CP];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int x0 = blockIdx.x * TW, y0 = blockIdx.y * TH;
    const size_t img_off = (size_t)blockIdx.z * H * Wd;
    // stage the input tile + 1-pixel halo, 16 bytes per load, zero outside the image
    {
        constexpr int N16 = (TH + 2) * (TW + 2) * (C / 8);
        constexpr int K = (N16 + NT - 1) / NT;
        uint4 r[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int i = tid + k * NT;
            const int pix = i >> 2, part = i & 3;
            const int ty = pix / (TW + 2), tx = pix % (TW + 2);
            const int gy = y0 + ty - 1, gx = x0 + tx - 1;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (i < N16 && gy >= 0 && gy < H && gx >= 0 && gx < Wd)
                v = reinterpret_cast<const uint4*>(in + (img_off + (size_t)gy * Wd + gx) * C)[part];
            r[k] = v;
        }
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int i = tid + k * NT;
            if (i < N16) reinterpret_cast<uint4*>(s_in + (i >> 2) * CP)[i & 3] = r[k];
        }
    }
    // weights as A operands: rows co = mt*16 + lane%16, k = tap*32 + 8*(lane/16) .. +7
    const int g = lane >> 4, ln = lane & 15;
    bf16x8 wa[2][9];
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int t = 0; t < 9; t++)
            wa[mt][t] = *reinterpret_cast<const bf16x8*>(wgt + ((mt * 16 + ln) * 9 + t) * C + 8 * g);
    float bco[2][4];
#pragma unroll
    for (int mt = 0; mt < 2; mt++)
#pragma unroll
        for (int i = 0; i < 4; i++) bco[mt][i] = MODE == 0 ? bias[mt * 16 + 4 * g + i] : 0.0f;
    __syncthreads();

    const int oy = y0 + wave;
#pragma unroll
    for (int nt = 0; nt < TW / 16; nt++) {
        f32x4 acc[2] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
#pragma unroll
        for (int t = 0; t < 9; t++) {
            const int ky = t / 3, kx = t % 3;
            const bf16x8 bv = *reinterpret_cast<const bf16x8*>(s_in + ((wave + ky) * (TW + 2) + nt * 16 + ln + kx) * CP + 8 * g);
            acc[0] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[0][t], bv, acc[0], 0, 0, 0);
            acc[1] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[1][t], bv, acc[1], 0, 0, 0);
        }
        const int ox = x0 + nt * 16 + ln;
        if (oy < H && ox < Wd) {
            const size_t p = img_off + (size_t)oy * Wd + ox;
#pragma unroll
            for (int mt = 0; mt < 2; mt++) {
                float v[4];
                if (MODE == 0) {
#pragma unroll
                    for (int i = 0; i < 4; i++) v[i] = fmaxf(acc[mt][i] + bco[mt][i], 0.0f);
                } else {
                    const uint2 mk = *reinterpret_cast<const uint2*>(mask + p * C + mt * 16 + 4 * g);
                    const uint16_t m4[4] = {(uint16_t)(mk.x & 0xffff), (uint16_t)(mk.x >> 16), (uint16_t)(mk.y & 0xffff), (uint16_t)(mk.y >> 16)};
#pragma unroll
                    for (int i = 0; i < 4; i++) v[i] = bf2f(m4[i]) > 0.0f ? acc[mt][i] : 0.0f;
                }
                uint2 o;
                o.x = f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16);
                o.y = f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16);
                *reinterpret_cast<uint2*>(out + p * C + mt * 16 + 4 * g) = o;
            }
        }
    }
}

// ---------------------------------------------------------------- weight gradient of a 32 -> 32 layer (MFMA)

// dW[co][tap][ci] = sum_p dY[p][co] * X[p + tap][ci];  db[co] = sum_p dY[p][co].
// Each block walks `segs` 64-pixel row segments (image row order), builds for
// each the im2col tile Xc[n = tap*32 + ci][64 px] and dY^T[co][64 px] in LDS,
// and accumulates D[n][co] (18 x 2 MFMA tiles, split over the 4 waves).
// Partial sums per block go to `part` (C*9*C + C floats per block).
constexpr int WG_PX = 64;
constexpr int WG_XS = WG_PX + 8;  // LDS row stride of the [n][px] tiles (bf16)
__global__ __launch_bounds__(NT) void k_wgrad_mid(const uint16_t* __restrict__ X, const uint16_t* __restrict__ dY, int H, int Wd, int n_img,
                                                  int segs, float* __restrict__ part) {
    __shared__ __attribute__((aligned(16))) uint16_t s_xc[9 * C * WG_XS];
    __shared__ __attribute__((aligned(16))) uint16_t s_dy[C * WG_XS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, g = lane >> 4, ln = lane & 15;
    const int segs_per_row = (Wd + WG_PX - 1) / WG_PX;
    const long total_segs = (long)segs_per_row * H * n_img;
    // wave w owns n-tiles {w, w+4, w+8, ...} of the 18, both co tiles
    f32x4 acc[5][2];
#pragma unroll
    for (int a = 0; a < 5; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0, 0, 0, 0};
    float dbacc = 0.0f;
    for (int s = 0; s < segs; s++) {
        const long seg = (long)blockIdx.x * segs + s;
        if (seg >= total_segs) break;
        const int img = (int)(seg / ((long)segs_per_row * H));
        const int rem = (int)(seg % ((long)segs_per_row * H));
        const int y = rem / segs_per_row, x0 = (rem % segs_per_row) * WG_PX;
        const size_t ioff = (size_t)img * H * Wd;
        __syncthreads();
        // dY^T tile: s_dy[co][px]; and the bias gradient
        for (int i = tid; i < WG_PX * C; i += NT) {
            const int px = i / C, co = i % C;
            const int gx = x0 + px;
            uint16_t v = 0;
            if (gx < Wd) v = dY[(ioff + (size_t)y * Wd + gx) * C + co];
            s_dy[co * WG_XS + px] = v;
        }
        // im2col: s_xc[tap*32 + ci][px] = X[y + ky - 1][x0 + px + kx - 1][ci]
        for (int i = tid; i < 9 * WG_PX * C; i += NT) {
            const int t = i / (WG_PX * C), r = i % (WG_PX * C);
            const int px = r / C, ci = r % C;
            const int gy = y + t / 3 - 1, gx = x0 + px + t % 3 - 1;
            uint16_t v = 0;
            if (gy >= 0 && gy < H && gx >= 0 && gx < Wd && x0 + px < Wd) v = X[(ioff + (size_t)gy * Wd + gx) * C + ci];
            s_xc[(t * C + ci) * WG_XS + px] = v;
        }
        __syncthreads();
        if (tid < C) {
            for (int px = 0; px < WG_PX; px++) dbacc += bf2f(s_dy[tid * WG_XS + px]);
        }
        // D[n][co] += sum_px Xc[n][px] * dY[px][co]; K = 64 px = 2 MFMA steps
#pragma unroll
        for (int a = 0; a < 5; a++) {
            const int nt = wave + 4 * a;
            if (nt >= 18) break;
#pragma unroll
            for (int kk = 0; kk < 2; kk++) {
                const bf16x8 av = *reinterpret_cast<const bf16x8*>(s_xc + (nt * 16 + ln) * WG_XS + kk * 32 + 8 * g);
#pragma unroll
                for (int b = 0; b < 2; b++) {
                    const bf16x8 bv = *reinterpret_cast<const bf16x8*>(s_dy + (b * 16 + ln) * WG_XS + kk * 32 + 8 * g);
                    acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[a][b], 0, 0, 0);
                }
            }
        }
    }
    // write partials: D[n][co] at lane: n = nt*16 + 4g + i, co = b*16 + ln -> dW[co][tap][ci] with n = tap*32 + ci
    float* pb = part + (size_t)blockIdx.x * (C * 9 * C + C);
#pragma unroll
    for (int a = 0; a < 5; a++) {
        const int nt = wave + 4 * a;
        if (nt >= 18) break;
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int n = nt * 16 + 4 * g + i, co = b * 16 + ln;
                pb[co * 9 * C + n] = acc[a][b][i];
            }
    }
    if (tid < C) pb[C * 9 * C + tid] = dbacc;
}

// Sum block partials: out[j] = sum_b part[b][j].  With split > 0, entries
// j >= split go to out2[j - split] instead.
__global__ void k_reduce_parts(const float* __restrict__ part, int nblk, int n, float* __restrict__ out, int split, float* __restrict__ out2) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    float s = 0.0f;
    for (int b = 0; b < nblk; b++) s += part[(size_t)b * n + j];
    if (split > 0 && j >= split)
        out2[j - split] = s;
    else
        out[j] = s;
}

// ---------------------------------------------------------------- loss and the VALU layers' gradients

// dout = 2 (out - tgt) / n_valid inside the crop margin, 0 outside; loss partials.
__global__ void k_mse(const float* __restrict__ out, const float* __restrict__ tgt, int H, int Wd, int n_img, int margin, float inv_n,
                      float* __restrict__ dout, float* __restrict__ loss_part) {
    __shared__ float red[256];
    const size_t npx = (size_t)H * Wd;
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.0f;
    if (p < npx * n_img) {
        const size_t q = p % npx;
        const int y = (int)(q / Wd), x = (int)(q % Wd);
        const bool in = y >= margin && y < H - margin && x >= margin && x < Wd - margin;
        const float e = out[p] - tgt[p];
        dout[p] = in ? 2.0f * e * inv_n : 0.0f;
        l = in ? e * e : 0.0f;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// g7[p][ci] = (a7[p][ci] > 0) * sum_tap W[tap][ci] * dout[p - tap]
__global__ __launch_bounds__(256) void k_out_dgrad(const float* __restrict__ dout, const uint16_t* __restrict__ W, const uint16_t* __restrict__ a7,
                                                   uint16_t* __restrict__ g7, int H, int Wd, int n_img) {
    __shared__ float sw[9 * C];
    for (int i = threadIdx.x; i < 9 * C; i += blockDim.x) sw[i] = bf2f(W[i]);  // the forward's bf16 weights
    __syncthreads();
    const size_t npx = (size_t)H * Wd;
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= npx * n_img) return;
    const size_t img = p / npx, q = p % npx;
    const int y = (int)(q / Wd), x = (int)(q % Wd);
    float d[9];
#pragma unroll
    for (int t = 0; t < 9; t++) {
        // out[p'] used a7[p' + off(t)]; so a7[p] feeds out[p - off(t)]
        const int yy = y - (t / 3 - 1), xx = x - (t % 3 - 1);
        d[t] = (yy >= 0 && yy < H && xx >= 0 && xx < Wd) ? dout[img * npx + (size_t)yy * Wd + xx] : 0.0f;
    }
    float a[C];
    load_px32(a7 + p * C, a);
    uint16_t o[C];
#pragma unroll
    for (int ci = 0; ci < C; ci++) {
        float s = 0.0f;
#pragma unroll
        for (int t = 0; t < 9; t++) s += sw[t * C + ci] * d[t];
        o[ci] = f2bf(a[ci] > 0.0f ? s : 0.0f);
    }
    uint4* dst = reinterpret_cast<uint4*>(g7 + p * C);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        uint4 u;
        u.x = o[8 * k] | ((uint32_t)o[8 * k + 1] << 16);
        u.y = o[8 * k + 2] | ((uint32_t)o[8 * k + 3] << 16);
        u.z = o[8 * k + 4] | ((uint32_t)o[8 * k + 5] << 16);
        u.w = o[8 * k + 6] | ((uint32_t)o[8 * k + 7] << 16);
        dst[k] = u;
    }
}

// Weight gradients of the output layer (dW[tap][ci] = sum_p dout[p] a7[p + tap][ci],
// db = sum dout) and of the input layer (dW[co][tap] = sum_p g1[p][co] x[p + tap],
// db[co] = sum g1[p][co]).  One thread per (block, parameter) sums a pixel range.
__global__ void k_io_wgrad(const float* __restrict__ dout, const uint16_t* __restrict__ a7, const uint16_t* __restrict__ g1,
                           const float* __restrict__ x, int H, int Wd, int n_img, int px_per_blk, float* __restrict__ part) {
    // parameters of this pass: out layer 9*C + 1, in layer C*9 + C
    constexpr int NOUT = 9 * C + 1, NIN = C * 9 + C, NP = NOUT + NIN;
    const int j = threadIdx.x + blockDim.x * blockIdx.y;  // parameter index
    if (j >= NP) return;
    const size_t npx = (size_t)H * Wd, tot = npx * n_img;
    const size_t p0 = (size_t)blockIdx.x * px_per_blk, p1 = p0 + px_per_blk < tot ? p0 + px_per_blk : tot;
    float s = 0.0f;
    if (j < NOUT) {
        if (j == 9 * C) {
            for (size_t p = p0; p < p1; p++) s += dout[p];
        } else {
            const int t = j / C, ci = j % C, dy = t / 3 - 1, dx = t % 3 - 1;
            for (size_t p = p0; p < p1; p++) {
                const size_t img = p / npx, q = p % npx;
                const int y = (int)(q / Wd) + dy, xx = (int)(q % Wd) + dx;
                if (y < 0 || y >= H || xx < 0 || xx >= Wd) continue;
                s += dout[p] * bf2f(a7[(img * npx + (size_t)y * Wd + xx) * C + ci]);
            }
        }
    } else {
        const int jj = j - NOUT;
        if (jj >= C * 9) {
            const int co = jj - C * 9;
            for (size_t p = p0; p < p1; p++) s += bf2f(g1[p * C + co]);
        } else {
            const int co = jj / 9, t = jj % 9, dy = t / 3 - 1, dx = t % 3 - 1;
            for (size_t p = p0; p < p1; p++) {
                const size_t img = p / npx, q = p % npx;
                const int y = (int)(q / Wd) + dy, xx = (int)(q % Wd) + dx;
                if (y < 0 || y >= H || xx < 0 || xx >= Wd) continue;
                s += bf2f(g1[p * C + co]) * bf2f(f2bf(x[img * npx + (size_t)y * Wd + xx]));  // the forward's bf16 x
            }
        }
    }
    part[(size_t)blockIdx.x * NP + j] = s;
}

// ---------------------------------------------------------------- Adam and weight packing

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ gr, float* __restrict__ m, float* __restrict__ v, int n, float lr,
                       float b1, float b2, float eps, float bc1, float bc2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = gr[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// bf16 copies of the input and output layers: W_in as a [co][k] A operand
// (k = tap, zero for k >= 9), W_out as [tap][ci].
__global__ void k_pack_io(const float* __restrict__ params, uint16_t* __restrict__ win, uint16_t* __restrict__ wout) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < C * 32) {
        const int co = i / 32, k = i % 32;
        win[i] = k < 9 ? f2bf(params[CNN_OFF_IN + co * 9 + k]) : (uint16_t)0;
    } else if (i < C * 32 + 9 * C) {
        const int j = i - C * 32;
        wout[j] = f2bf(params[CNN_OFF_OUT + j]);
    }
}

// bf16 copies of the six middle layers: forward layout [co][tap][ci] and the
// data-gradient layout W'[ci][tap'][co] = W[co][8 - tap'][ci].
__global__ void k_pack_mid(const float* __restrict__ wmid, uint16_t* __restrict__ fwd, uint16_t* __restrict__ bwd) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    constexpr int PER = C * 9 * C;
    if (i >= 6 * PER) return;
    const int l = i / PER, r = i % PER;
    const int co = r / (9 * C), t = (r / C) % 9, ci = r % C;
    const uint16_t b = f2bf(wmid[(size_t)l * (PER + C) + r]);
    fwd[(size_t)l * PER + r] = b;
    bwd[(size_t)l * PER + ((ci * 9) + (8 - t)) * C + co] = b;
}

// ---------------------------------------------------------------- input normalisation and crops

// x = clamp((L - mean5x5(L)) / s, -64, 64) with s = max(sigma, 1e-4); the
// 5x5 mean uses clamped coordinates.
// x = clamp((L - mean5x5(L)) / s, +-64) with clamped edges; the 5x5 window
// is summed row by row, left to right, from a 64x16 tile plus its 2-pixel
// halo staged in LDS.
constexpr int CI_TW = 64, CI_TH = 16;
__global__ __launch_bounds__(256) void k_cnn_input(const float* __restrict__ L, int w, int h, const float* __restrict__ sigma_dev,
                                                   float* __restrict__ x) {
    constexpr int SW = CI_TW + 4, SH = CI_TH + 4;
    __shared__ float t[SH][SW + 1];
    const int x0 = blockIdx.x * CI_TW, y0 = blockIdx.y * CI_TH;
    for (int i = threadIdx.x; i < SW * SH; i += 256) {
        const int u = i % SW, v = i / SW;
        t[v][u] = L[(size_t)clampi(y0 + v - 2, 0, h - 1) * w + clampi(x0 + u - 2, 0, w - 1)];
    }
    __syncthreads();
    const float sg = fmaxf(*sigma_dev, 1e-4f);
    const int u = threadIdx.x % CI_TW;
    for (int v = threadIdx.x / CI_TW; v < CI_TH; v += 256 / CI_TW) {
        const int px = x0 + u, py = y0 + v;
        if (px >= w || py >= h) continue;
        float s = 0.0f;
#pragma unroll
        for (int dy = 0; dy < 5; dy++)
#pragma unroll
            for (int dx = 0; dx < 5; dx++) s += t[v + dy][u + dx];
        x[(size_t)py * w + px] = fminf(fmaxf((t[v + 2][u + 2] - s * (1.0f / 25.0f)) / sg, -64.0f), 64.0f);
    }
}

// target of training: x + (LF - L) / s
__global__ void k_cnn_target(const float* __restrict__ x, const float* __restrict__ L, const float* __restrict__ LF, size_t n,
                             const float* __restrict__ sigma_dev, float* __restrict__ t) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    t[i] = x[i] + (LF[i] - L[i]) / fmaxf(*sigma_dev, 1e-4f);
}

// lf = L + s * (out - x)

// cut n_crops crops of c x c from two w x h planes at seeded positions
__global__ void k_crops(const float* __restrict__ a, const float* __restrict__ b, int w, int h, int c, int n_crops, uint64_t seed,
                        float* __restrict__ ca, float* __restrict__ cb) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t per = (size_t)c * c;
    if (i >= per * n_crops) return;
    const int k = (int)(i / per), r = (int)(i % per);
    uint64_t z = seed * 0x9E3779B97F4A7C15ull + (uint64_t)k * 0xBF58476D1CE4E5B9ull;
    z ^= z >> 31;
    z *= 0x94D049BB133111EBull;
    z ^= z >> 29;
    const int ox = (int)((z & 0xffffffffull) % (uint64_t)(w - c + 1)), oy = (int)((z >> 32) % (uint64_t)(h - c + 1));
    const size_t src = (size_t)(oy + r / c) * w + ox + r % c;
    ca[i] = a[src];
    cb[i] = b[src];
}

// ---------------------------------------------------------------- fused inference forward

// The forward in two streaming passes of three MFMA layers each:
//   pass A: x -> a1 (FP32) -> a2 -> a3 -> a4, a4 written to HBM (bf16 NHWC);
//   pass B: a4 -> a5 -> a6 -> a7 -> y (FP32 output layer and residual).
// A block owns a strip of FS_OUT output columns and FS_SH output rows and
// walks down it one row per step, keeping the last three rows of each
// activation it needs in LDS rings.  Rows are FS_W = FS_OUT + 2 FS_HALO
// pixels wide (6 MFMA column tiles); a pass reaches 4 pixels, so the halo
// columns' values go stale one column per layer from the strip edge inwards
// and never reach the stored columns.  Activations outside the image are
// stored as 0 (the zero padding of every layer).  The next input row is
// fetched into registers one step ahead.  The arithmetic is that of
// k_conv_in / k_conv_mid<0> / k_conv_out operation by operation, so the
// result equals the layer-by-layer forward bit for bit.
constexpr int FS_HALO = 4;
constexpr int FS_W = 96;                    // 6 MFMA column tiles
constexpr int FS_OUT = FS_W - 2 * FS_HALO;  // 88
constexpr int FS_RW = FS_W + 2;             // ring row: 1 zero pixel each side
constexpr int FS_SH = 192;
constexpr int FS_NT = 256;
constexpr size_t FS_ROW_BYTES = (size_t)FS_RW * C * sizeof(uint16_t);
// Pass A's input-layer operands for one a1 row (one slot): per computed
// column the 32-element K vector of k_conv_in (taps 0..8 as bf16, then
// zeros) in four 16-byte chunks, chunk-major (FS_IM_CHUNK bytes per chunk,
// the column's ring entry inside it).  Chunks 2 and 3 are zeros.  Three
// slots, one per a1 row mod 3.
constexpr size_t FS_IM_CHUNK = (size_t)FS_W * 16;
constexpr size_t FS_IM_SLOT = 4 * FS_IM_CHUNK;
constexpr size_t FS_IM_BYTES = 3 * FS_IM_SLOT;
// pass A: rings a1..a3, input-layer operands, x ring; pass B: rings a4..a7
constexpr size_t FS_LDS_A = 3 * 3 * FS_ROW_BYTES + FS_IM_BYTES + 3 * FS_RW * sizeof(float);
constexpr size_t FS_LDS_B = 4 * 3 * FS_ROW_BYTES;

// Ring rows are stored class-major: computed column j (0 <= j < FS_W) sits
// at entry (j % FS_NC) * 16 + j / FS_NC, so the MFMA column tile "class s"
// (lane ln <-> column s + FS_NC * ln) is 16 consecutive entries.  Tap kx of
// class t reads columns (t + kx - 1) + FS_NC * ln, which is class t + kx - 1
// for 1 <= t + kx <= FS_NC; class -1 is class FS_NC - 1 one lane down and
// class FS_NC is class 0 one lane up, with column -1 and column FS_W (zero
// padding, or in the input ring of pass B the neighbouring image columns) at
// entries FS_W and FS_W + 1.  Every B fragment is therefore loaded once and
// feeds the three horizontal taps of its neighbours, and each MFMA still sees
// exactly the operands of the layer-by-layer kernels.
//
// A pixel's four 16-byte channel chunks are XOR-swizzled by (entry >> 1) & 3,
// so that the eight lanes of a ds_read_b128 group, which read the same chunk
// of eight consecutive entries, hit eight different bank groups (64-byte
// entries would otherwise put every other entry on the same banks).
constexpr int FS_NC = FS_W / 16;   // column classes (6)
constexpr int FS_CPW = FS_NC / 2;  // classes per wave (a wave owns one half of the row)
constexpr int FS_WIN = FS_CPW + 2; // classes a wave reads: its own and one either side
static_assert(FS_NC % 2 == 0, "two waves per channel half split the classes");
__host__ __device__ constexpr int fs_entry(int col) { return col < 0 ? FS_W : col >= FS_W ? FS_W + 1 : (col % FS_NC) * 16 + col / FS_NC; }
__device__ __forceinline__ int fs_chunk(int e, int c) { return e * 4 + (c ^ ((e >> 1) & 3)); }  // uint4 index in a ring row

// Byte offsets, within a ring row, of the operands a lane touches.  Wave
// half h owns classes FS_CPW h .. FS_CPW h + FS_CPW - 1; its B fragments are
// those of the classes FS_CPW h - 1 + i, i < FS_WIN.  The epilogue writes
// class FS_CPW h + q at st + 1024 q (16 entries of 64 bytes further per class;
// the swizzle depends on the entry modulo 8 only).
struct FsLane {
    int b[FS_WIN];
    int st;
};
__device__ __forceinline__ FsLane fs_lane(int lane, int mt, int half) {
    const int g = lane >> 4, ln = lane & 15;
    FsLane o;
#pragma unroll
    for (int i = 0; i < FS_WIN; i++) o.b[i] = fs_chunk(fs_entry(FS_CPW * half - 1 + i + FS_NC * ln), g) * 16;
    o.st = fs_chunk(FS_CPW * half * 16 + ln, 2 * mt + (g >> 1)) * 16 + 8 * (g & 1);
    return o;
}
template <class T>
__device__ __forceinline__ T lds_at(const void* base, int byte_off) {
    return *reinterpret_cast<const T*>(reinterpret_cast<const uint8_t*>(base) + byte_off);
}
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
// relu(a + ba), relu(b + bb) rounded to bf16 (nearest even) and packed low/high
__device__ __forceinline__ uint32_t fs_pack(float a, float ba, float b, float bb) {
    const f32x2 v = {fmaxf(a + ba, 0.0f), fmaxf(b + bb, 0.0f)};
    return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, bf16x2));
}

// Kernel row ky of one MFMA layer row: taps 3 ky .. 3 ky + 2 of the wave's
// classes from source row `src`, accumulated into acc (ky = 0, 1, 2 in turn
// gives the tap order of k_conv_mid).
__device__ __forceinline__ void fs_mma_row(f32x4 (&acc)[FS_CPW], const bf16x8 (&wa)[9], int ky, const uint16_t* src, const FsLane& lo) {
    bf16x8 b[FS_WIN];
#pragma unroll
    for (int i = 0; i < FS_WIN; i++) b[i] = lds_at<bf16x8>(src, lo.b[i]);
#pragma unroll
    for (int kx = 0; kx < 3; kx++)
#pragma unroll
        for (int q = 0; q < FS_CPW; q++) acc[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wa[3 * ky + kx], b[q + kx], acc[q], 0, 0, 0);
}

// Epilogue of one MFMA layer row: relu(acc + bias) as bf16, packed per class
// (channels 4g .. 4g + 3 of the wave's M tile), zero outside the image
// (row_in false, or col_in[q] false for the class's column).
__device__ __forceinline__ void fs_pack_row(const f32x4 (&acc)[FS_CPW], const float (&bias)[4], bool row_in, const bool (&col_in)[FS_CPW],
                                            uint2 (&o)[FS_CPW]) {
#pragma unroll
    for (int q = 0; q < FS_CPW; q++) {
        const uint32_t keep = row_in && col_in[q] ? 0xffffffffu : 0u;
        o[q].x = fs_pack(acc[q][0], bias[0], acc[q][1], bias[1]) & keep;
        o[q].y = fs_pack(acc[q][2], bias[2], acc[q][3], bias[3]) & keep;
    }
}
// the wave's classes of a packed row into a ring row
__device__ __forceinline__ void fs_put_row(const uint2 (&o)[FS_CPW], uint16_t* dst, const FsLane& lo) {
#pragma unroll
    for (int q = 0; q < FS_CPW; q++) *reinterpret_cast<uint2*>(reinterpret_cast<uint8_t*>(dst) + lo.st + q * 1024) = o[q];
}

// PASS 0 (A): x -> a4 (act_out, n*H*W*32 bf16).  PASS 1 (B): a4 (act_in) -> y.
//
// A step at row r has two phases.  Phase 1 reads only rows published before
// the step and does all of the step's MFMAs: layer l of the pass computes
// row r - 1 - 2l (A: a1 row r first, from its operands, then a2 row r - 2,
// a3 row r - 4, a4 row r - 6; B: a5 row r - 1, a6 row r - 3, a7 row r - 5
// and output row r - 7), so each layer's newest source row was written in
// the step before.  After a barrier, phase 2 writes the new rows into the
// ring slots of the rows phase 1 read last (row - 3), commits the next input
// row, fetches the one after and sends this step's a4 / output row to HBM;
// a second barrier publishes them.  A layer whose row no stored row depends
// on is skipped.
// The row loop is unrolled by three (y0 and the first row are multiples of
// three), so every ring slot ((row) mod 3) is a compile-time constant and
// the LDS addresses are a lane's fixed offsets plus immediates.
// B with lum != nullptr: out receives lum + max(*sigma_dev, 1e-4) (y - x), the
// denoised luminance, in place of y.
template <int PASS>
__global__ __launch_bounds__(FS_NT, 2) void k_cnn_fused(const float* __restrict__ x, const uint16_t* __restrict__ act_in,
                                                        const float* __restrict__ params, const uint16_t* __restrict__ wpk,
                                                        uint16_t* __restrict__ act_out, float* __restrict__ out, int H, int Wd,
                                                        const float* __restrict__ lum, const float* __restrict__ sigma_dev) {
    extern __shared__ __attribute__((aligned(16))) uint8_t fs_lds[];
    constexpr int NR = PASS == 0 ? 3 : 4;  // LDS rings
    uint16_t* ring = reinterpret_cast<uint16_t*>(fs_lds);
    uint8_t* im = fs_lds + NR * 3 * FS_ROW_BYTES;            // A: input-layer operands, 3 slots (a1 row mod 3)
    float* xr = reinterpret_cast<float*>(im + FS_IM_BYTES);  // A: x ring [3][FS_RW]
    const int tid = threadIdx.x, lane = tid & 63, g = lane >> 4, ln = lane & 15;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int x0 = blockIdx.x * FS_OUT, y0 = blockIdx.y * FS_SH;
    static_assert(FS_SH % 3 == 0, "ring slots follow the row index mod 3");
    const size_t ioff = (size_t)blockIdx.z * H * Wd;
    const int gx0 = x0 - FS_HALO;  // global column of computed column 0
    // ring k, slot s (= row mod 3)
    auto RING = [&](int k, int slot) { return ring + (size_t)(k * 3 + slot) * FS_RW * C; };

    constexpr int ZERO_BYTES = NR * 3 * FS_ROW_BYTES + (PASS == 0 ? FS_IM_BYTES : 0);  // rings and (A) the operand chunks
    for (int i = tid; i < ZERO_BYTES / 16; i += FS_NT) reinterpret_cast<uint4*>(fs_lds)[i] = make_uint4(0, 0, 0, 0);
    // input / output layer operands (as k_conv_in / k_conv_out: W_out[ky][kx] in row kx of the A operand)
    bf16x8 win = {}, wout[3];
    float bin[4] = {0, 0, 0, 0};
    if (PASS == 0) {
        win = *reinterpret_cast<const bf16x8*>(wpk + CNN_WPK_IN + ((wave & 1) * 16 + ln) * 32 + 8 * g);
#pragma unroll
        for (int i = 0; i < 4; i++) bin[i] = params[CNN_OFF_IN + C * 9 + (wave & 1) * 16 + 4 * g + i];
    }
#pragma unroll
    for (int ky = 0; ky < 3; ky++) {
        bf16x8 z = {};
        if (PASS == 1 && ln < 3) z = *reinterpret_cast<const bf16x8*>(wpk + CNN_WPK_OUT + (ky * 3 + ln) * C + 8 * g);
        wout[ky] = z;
    }
    const float b_out = params[CNN_OFF_OUT + 9 * C];
    const float s_lum = PASS == 1 && lum ? fmaxf(*sigma_dev, 1e-4f) : 0.0f;
    const int l0 = PASS == 0 ? 0 : 3;  // first middle layer of the pass
    const int mt = wave & 1;
    bf16x8 wa[3][9];
    float bm[3][4];
#pragma unroll
    for (int l = 0; l < 3; l++) {
#pragma unroll
        for (int t = 0; t < 9; t++)
            wa[l][t] = *reinterpret_cast<const bf16x8*>(wpk + (size_t)(l0 + l) * C * 9 * C + ((mt * 16 + (lane & 15)) * 9 + t) * C + 8 * g);
#pragma unroll
        for (int i = 0; i < 4; i++) bm[l][i] = params[CNN_OFF_MID + (size_t)(l0 + l) * CNN_MID_STRIDE + C * 9 * C + mt * 16 + 4 * g + i];
    }
    const int half = wave >> 1;
    const FsLane lo = fs_lane(lane, mt, half);
    // computed column of this lane in class FS_CPW half + q is o_j0 + q
    const int o_j0 = FS_CPW * half + FS_NC * ln;
    bool col_in[FS_CPW];  // the column lies in the image
    bool o_col[FS_CPW];   // this lane sends the column of a stored row to HBM
#pragma unroll
    for (int q = 0; q < FS_CPW; q++) {
        const int j = o_j0 + q, gx = gx0 + j;
        col_in[q] = gx >= 0 && gx < Wd;
        // B: of the wave half's classes the mt = 0 wave owns the first FS_CPW - 1, the
        // mt = 1 wave the last; one lane group (g = 0) holds a column's output
        o_col[q] = j >= FS_HALO && j < FS_HALO + FS_OUT && gx < Wd && (PASS == 0 || ((q < FS_CPW - 1) == (mt == 0) && g == 0));
    }
    // B: the output layer.  A wave stores NQ classes at most: class oq[k] (mt = 0:
    // k; mt = 1: FS_CPW - 1, for k = 0 only) from window classes k0 .. k0 + FS_CPW
    // (k0 = 0 or FS_CPW - 1, at ring byte offsets ob[]; for mt = 1 the last is a
    // repeat that no output uses)
    constexpr int NQ = FS_CPW - 1;
    int oq[NQ];
    bool oq_on[NQ];
#pragma unroll
    for (int k = 0; k < NQ; k++) {
        oq[k] = mt ? FS_CPW - 1 : k;
        oq_on[k] = mt ? k == 0 && o_col[FS_CPW - 1] : o_col[k];
    }
    int ob[FS_CPW + 1];
#pragma unroll
    for (int k = 0; k < FS_CPW + 1; k++) ob[k] = mt ? lo.b[FS_CPW - 1 + k < FS_WIN ? FS_CPW - 1 + k : FS_WIN - 1] : lo.b[k];

    // input rows: A reads x rows (FS_RW floats, element i = global column
    // gx0 - 1 + i), B reads a4 rows (FS_RW pixels of 64 bytes, pixel i =
    // global column gx0 - 1 + i, stored at ring entry fs_entry(i - 1))
    constexpr int IN_UNITS = PASS == 0 ? FS_RW : FS_RW * 4;  // floats / 16-byte chunks per row
    constexpr int IN_K = (IN_UNITS + FS_NT - 1) / FS_NT;
    float xv[PASS == 0 ? IN_K : 1];
    uint4 av[PASS == 0 ? 1 : IN_K];
    int cofs[IN_K];  // B: uint4 index of this thread's chunks in a ring row
#pragma unroll
    for (int k = 0; k < IN_K; k++) {
        const int u = tid + k * FS_NT;
        cofs[k] = fs_chunk(fs_entry((u >> 2) - 1), u & 3);
    }
    auto fetch = [&](int row) {  // global -> registers
#pragma unroll
        for (int k = 0; k < IN_K; k++) {
            const int u = tid + k * FS_NT;
            if (PASS == 0) {
                const int gx = gx0 - 1 + u;
                xv[k] = (u < IN_UNITS && row >= 0 && row < H && gx >= 0 && gx < Wd) ? x[ioff + (size_t)row * Wd + gx] : 0.0f;
            } else {
                const int gx = gx0 - 1 + (u >> 2);
                av[k] = (u < IN_UNITS && row >= 0 && row < H && gx >= 0 && gx < Wd)
                            ? reinterpret_cast<const uint4*>(act_in + (ioff + (size_t)row * Wd + gx) * C)[u & 3]
                            : make_uint4(0, 0, 0, 0);
            }
        }
    };
    auto commit = [&](int slot) {  // registers -> input ring slot
#pragma unroll
        for (int k = 0; k < IN_K; k++) {
            const int u = tid + k * FS_NT;
            if (u < IN_UNITS) {
                if (PASS == 0)
                    xr[slot * FS_RW + u] = xv[k];
                else
                    reinterpret_cast<uint4*>(RING(0, slot))[cofs[k]] = av[k];
            }
        }
    };
    // A: input-layer operands of an a1 row into operand slot si from the x
    // rows above, at and below it (x ring slots s0, s1, s2), cast to bf16 as
    // k_conv_in does.  Entry e holds computed column j = (e % 16) FS_NC + e / 16
    // (the inverse of fs_entry), so consecutive lanes write consecutive
    // entries.  Waves 0 and 1 write taps 0..3 of entries 0..95, waves 2 and 3
    // taps 4..8.
    auto build_im = [&](int si, int s0, int s1, int s2) {
        if (PASS == 0 && lane < FS_W / 2) {
            uint8_t* d = im + si * FS_IM_SLOT;
            const int e = (wave & 1) * (FS_W / 2) + lane;
            const int j = (e & 15) * FS_NC + (e >> 4);
            typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
            if (wave < 2) {
                const float* a = xr + s0 * FS_RW + j;
                const float* b = xr + s1 * FS_RW + j;
                const bf16x4 v = {(__bf16)a[0], (__bf16)a[1], (__bf16)a[2], (__bf16)b[0]};
                *reinterpret_cast<bf16x4*>(d + e * 16) = v;
            } else {
                const float* b = xr + s1 * FS_RW + j;
                const float* c = xr + s2 * FS_RW + j;
                const bf16x4 v = {(__bf16)b[1], (__bf16)b[2], (__bf16)c[0], (__bf16)c[1]};
                *reinterpret_cast<bf16x4*>(d + e * 16 + 8) = v;
                *reinterpret_cast<__bf16*>(d + FS_IM_CHUNK + e * 16) = (__bf16)c[2];  // the chunk's other elements stay 0
            }
        }
    };
    // A: byte offset of this lane's input-layer operand in an operand slot
    // (class FS_CPW half + q at + 256 q): chunk g, the 16 lanes of a group on
    // consecutive entries
    const int im_off = g * (int)FS_IM_CHUNK + (FS_CPW * half * 16 + ln) * 16;

    // The first step is r = y0 - 3, a multiple of 3.  A: step r builds the
    // operands of a1 row r + 1 from x rows r .. r + 2 and commits x row r + 3;
    // B: step r reads a4 rows r - 2 .. r and commits a4 row r + 1.
    const int r_first = y0 - 3;
    const int r_end = y0 + FS_SH + (PASS == 0 ? 6 : 7);
    fetch(r_first - 1);
    __syncthreads();  // ring zeroing done before any commit
    commit(2);        // row r_first - 1, slot 2
    fetch(r_first);
    commit(0);
    fetch(r_first + 1);
    if (PASS == 0) {
        commit(1);
        fetch(r_first + 2);
        __syncthreads();
        build_im(0, 2, 0, 1);  // a1 row r_first from x rows r_first - 1 .. r_first + 1
        __syncthreads();       // x row r_first - 1 read
        commit(2);             // x row r_first + 2
        fetch(r_first + 3);
    }
    __syncthreads();

    constexpr size_t IM_SLOT = FS_IM_SLOT;
    auto step = [&](auto sub_tag, int r) {
        constexpr int SUB = decltype(sub_tag)::value;
        auto slot = [](int c) { return ((SUB + c) % 3 + 3) % 3; };  // slot of row r + c; constant-folded
        // the row each layer computes at this step, and whether a stored row needs it
        // (S = FS_SH; each 3x3 layer needs one more row either side than the next):
        // A: a1 r in [y0-3, y0+S+3), a2 r-2 in [y0-2, y0+S+2), a3 r-4 in [y0-1, y0+S+1), a4 r-6 in [y0, y0+S)
        // B: a5 r-1 in [y0-3, y0+S+3), a6 r-3 in [y0-2, y0+S+2), a7 r-5 in [y0-1, y0+S+1), y r-7 in [y0, y0+S)
        auto need = [&](int row, int lo_, int hi_) { return row >= y0 + lo_ && row < y0 + FS_SH + hi_; };
        constexpr int D0 = PASS == 0 ? 2 : 1;  // layer l computes row r - D0 - 2 l
        const int row_l[3] = {r - D0, r - D0 - 2, r - D0 - 4};
        bool on_l[3];
#pragma unroll
        for (int l = 0; l < 3; l++) {
            constexpr int NL0 = PASS == 0 ? 2 : 3;  // rows needed around [y0, y0 + S) by layer 0
            on_l[l] = need(row_l[l], l - NL0, NL0 - l);
        }
        // B: output row and its x (and L), loaded now so their latency passes under phase 1
        const int orow = r - 7;
        const bool orow_on = PASS == 1 && need(orow, 0, 0) && orow < H;
        const size_t orow_p = ioff + (size_t)orow * Wd + gx0 + o_j0;
        float o_x[NQ], o_l[NQ];
#pragma unroll
        for (int k = 0; k < NQ; k++) {
            const bool on = orow_on && oq_on[k];
            o_x[k] = on ? x[orow_p + oq[k]] : 0.0f;
            o_l[k] = on && lum ? lum[orow_p + oq[k]] : 0.0f;
        }

        // ---------------- phase 1: reads and MFMAs
        // A: a1 row r (one MFMA per class, K = taps; as k_conv_in); operands of a1 row r + 1
        uint2 p_in[FS_CPW];
        if (PASS == 0) {
            bf16x8 bv[FS_CPW];
            f32x4 a1[FS_CPW];
#pragma unroll
            for (int q = 0; q < FS_CPW; q++) bv[q] = lds_at<bf16x8>(im, slot(0) * IM_SLOT + im_off + q * 256);
#pragma unroll
            for (int q = 0; q < FS_CPW; q++) a1[q] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(win, bv[q], f32x4{0, 0, 0, 0}, 0, 0, 0);
            fs_pack_row(a1, bin, r >= 0 && r < H, col_in, p_in);
            build_im(slot(1), slot(0), slot(1), slot(2));
        }
        // middle layers: layer l reads ring l rows row_l[l] - 1 .. row_l[l] + 1
        uint2 p_mid[3][FS_CPW];
#pragma unroll
        for (int l = 0; l < 3; l++) {
            const int c0 = row_l[l] - 1 - r;  // source rows r + c0 .. r + c0 + 2
            f32x4 acc[FS_CPW];
#pragma unroll
            for (int q = 0; q < FS_CPW; q++) acc[q] = f32x4{0, 0, 0, 0};
#pragma unroll
            for (int ky = 0; ky < 3; ky++) fs_mma_row(acc, wa[l], ky, RING(l, slot(c0 + ky)), lo);
            fs_pack_row(acc, bm[l], row_l[l] >= 0 && row_l[l] < H, col_in, p_mid[l]);
        }
        // B: output row r - 7 (as k_conv_out): E_kx of window class i in row kx of a
        // three-MFMA chain over ky (a7 rows r - 8 .. r - 6), then
        // y = x + (((E_0(col - 1) + E_1(col)) + E_2(col + 1)) + b_out), for the wave's
        // output classes from its window classes k0 .. k0 + FS_CPW
        float ov[NQ];
        if (PASS == 1) {
            f32x4 e[FS_CPW + 1];
#pragma unroll
            for (int k = 0; k < FS_CPW + 1; k++) e[k] = f32x4{0, 0, 0, 0};
#pragma unroll
            for (int ky = 0; ky < 3; ky++) {
                const uint16_t* src = RING(3, slot(-8 + ky));
#pragma unroll
                for (int k = 0; k < FS_CPW + 1; k++)
                    e[k] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(wout[ky], lds_at<bf16x8>(src, ob[k]), e[k], 0, 0, 0);
            }
#pragma unroll
            for (int k = 0; k < NQ; k++) {
                // window classes of output class oq[k]: k .. k + 2 (mt = 0), 0 .. 2 (mt = 1)
                const float e0 = mt ? e[0][0] : e[k][0], e1 = mt ? e[1][1] : e[k + 1][1], e2 = mt ? e[2][2] : e[k + 2][2];
                const float y = o_x[k] + (((e0 + e1) + e2) + b_out);
                ov[k] = lum ? o_l[k] + s_lum * (y - o_x[k]) : y;
            }
        }
        __syncthreads();

        // ---------------- phase 2: writes (each into the slot of the row three above, read in phase 1)
        if (PASS == 0) fs_put_row(p_in, RING(0, slot(0)), lo);
#pragma unroll
        for (int l = 0; l < 3; l++) {
            if (PASS == 0 && l == 2) continue;  // A: a4 goes to HBM
            fs_put_row(p_mid[l], RING(l + 1, slot(row_l[l] - r)), lo);
        }
        if (PASS == 0) {
            commit(slot(0));  // x row r + 3
            fetch(r + 4);
            if (on_l[2] && row_l[2] < H) {
                uint16_t* p = act_out + (ioff + (size_t)row_l[2] * Wd + gx0 + o_j0) * C + mt * 16 + 4 * g;
#pragma unroll
                for (int q = 0; q < FS_CPW; q++)
                    if (o_col[q]) *reinterpret_cast<uint2*>(p + q * C) = p_mid[2][q];
            }
        } else {
            commit(slot(1));  // a4 row r + 1
            fetch(r + 2);
            if (orow_on) {
#pragma unroll
                for (int k = 0; k < NQ; k++)
                    if (oq_on[k]) out[orow_p + oq[k]] = ov[k];
            }
        }
        __syncthreads();
    };
    for (int rb = r_first; rb < r_end; rb += 3) {
        step(std::integral_constant<int, 0>{}, rb);
        if (rb + 1 < r_end) step(std::integral_constant<int, 1>{}, rb + 1);
        if (rb + 2 < r_end) step(std::integral_constant<int, 2>{}, rb + 2);
    }
}

unsigned nb(size_t n, int bs = 256) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace

// ---------------------------------------------------------------- host side

size_t cnn_param_count() { return CNN_NPARAM; }

hipError_t cnn_pack(const float* params, uint16_t* wpk, hipStream_t st) {
    hipLaunchKernelGGL(k_pack_mid, dim3(nb(6 * C * 9 * C)), dim3(256), 0, st, params + CNN_OFF_MID, wpk, wpk + CNN_WPK_BWD);
    hipLaunchKernelGGL(k_pack_io, dim3(nb(C * 32 + 9 * C)), dim3(256), 0, st, params, wpk + CNN_WPK_IN, wpk + CNN_WPK_OUT);
    return hipGetLastError();
}

// Forward of n_img images H x W: x (fp32) -> out (fp32); acts: 7 activation
// buffers of n_img*H*W*32 bf16 (kept for the backward pass).
hipError_t cnn_forward(const float* params, const uint16_t* wpk, const float* x, int n_img, int H, int W, uint16_t* const* acts, float* out,
                       hipStream_t st) {
    dim3 grid((unsigned)((W + TW - 1) / TW), (unsigned)((H + TH - 1) / TH), (unsigned)n_img);
    hipLaunchKernelGGL(k_conv_in, grid, dim3(256), 0, st, x, wpk + CNN_WPK_IN, params + CNN_OFF_IN + C * 9, acts[0], H, W);
    for (int l = 0; l < 6; l++) {
        const float* bias = params + CNN_OFF_MID + (size_t)l * (C * 9 * C + C) + C * 9 * C;
        hipLaunchKernelGGL(k_conv_mid<0>, grid, dim3(NT), 0, st, acts[l], wpk + (size_t)l * C * 9 * C, bias, (const uint16_t*)nullptr, acts[l + 1],
                           H, W);
    }
    hipLaunchKernelGGL(k_conv_out, grid, dim3(256), 0, st, acts[6], x, wpk + CNN_WPK_OUT, params + CNN_OFF_OUT + 9 * C, out, H, W);
    return hipGetLastError();
}

size_t cnn_fused_act_bytes(int n_img, int H, int W) { return (size_t)n_img * H * W * C * sizeof(uint16_t); }

hipError_t cnn_forward_fused(const float* params, const uint16_t* wpk, const float* x, int n_img, int H, int W, uint16_t* act, float* out,
                             hipStream_t st, const float* lum, const float* sigma_dev) {
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void*)k_cnn_fused<0>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)FS_LDS_A);
        if (e == hipSuccess) e = hipFuncSetAttribute((const void*)k_cnn_fused<1>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)FS_LDS_B);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    dim3 grid((unsigned)((W + FS_OUT - 1) / FS_OUT), (unsigned)((H + FS_SH - 1) / FS_SH), (unsigned)n_img);
    hipLaunchKernelGGL(k_cnn_fused<0>, grid, dim3(FS_NT), FS_LDS_A, st, x, (const uint16_t*)nullptr, params, wpk, act, (float*)nullptr, H, W,
                       (const float*)nullptr, (const float*)nullptr);
    hipLaunchKernelGGL(k_cnn_fused<1>, grid, dim3(FS_NT), FS_LDS_B, st, x, (const uint16_t*)act, params, wpk, (uint16_t*)nullptr, out, H, W, lum,
                       sigma_dev);
    return hipGetLastError();
}

size_t cnn_wgrad_blocks(int n_img, int H, int W) {
    const long segs = (long)((W + WG_PX - 1) / WG_PX) * H * n_img;
    const long per = 16;
    return (size_t)((segs + per - 1) / per);
}

size_t cnn_scratch_floats(int n_img, int H, int W) {
    const size_t npx = (size_t)n_img * H * W;
    const size_t wg = cnn_wgrad_blocks(n_img, H, W) * (C * 9 * C + C);
    const size_t io_blk = (npx + 4095) / 4096;
    const size_t io = io_blk * (9 * C + 1 + C * 9 + C);
    size_t m = wg > io ? wg : io;
    return npx /* dout */ + nb(npx) /* loss partials */ + m;
}

// Loss (MSE over the crop interior) and gradients of all parameters into grad.
// gacts: 2 scratch activation-gradient buffers; scratch: cnn_scratch_floats.
hipError_t cnn_backward(const float* params, const uint16_t* wpk, const float* x, const float* tgt, int n_img, int H, int W, int margin,
                        uint16_t* const* acts, const float* out, uint16_t* const* gacts, float* scratch, float* grad, float* loss_part_out,
                        hipStream_t st) {
    const size_t npx = (size_t)n_img * H * W;
    const long nvalid = (long)n_img * (H - 2 * margin) * (W - 2 * margin);
    float* dout = scratch;
    float* lossp = scratch + npx;
    float* part = lossp + nb(npx);
    hipLaunchKernelGGL(k_mse, dim3(nb(npx)), dim3(256), 0, st, out, tgt, H, W, n_img, margin, 1.0f / (float)(nvalid > 0 ? nvalid : 1), dout,
                       loss_part_out ? loss_part_out : lossp);
    // a7 -> g7
    hipLaunchKernelGGL(k_out_dgrad, dim3(nb(npx)), dim3(256), 0, st, dout, wpk + CNN_WPK_OUT, acts[6], gacts[0], H, W, n_img);
    dim3 grid((unsigned)((W + TW - 1) / TW), (unsigned)((H + TH - 1) / TH), (unsigned)n_img);
    const int wblk = (int)cnn_wgrad_blocks(n_img, H, W);
    const int npar = C * 9 * C + C;
    int cur = 0;
    for (int l = 5; l >= 0; l--) {
        // layer l maps acts[l] -> acts[l+1]; gacts[cur] holds dL/d(pre-activation of l+1)
        hipLaunchKernelGGL(k_wgrad_mid, dim3(wblk), dim3(NT), 0, st, acts[l], gacts[cur], H, W, n_img, 16, part);
        hipLaunchKernelGGL(k_reduce_parts, dim3(nb(npar)), dim3(256), 0, st, part, wblk, npar, grad + CNN_OFF_MID + (size_t)l * npar, 0,
                           (float*)nullptr);
        hipLaunchKernelGGL(k_conv_mid<1>, grid, dim3(NT), 0, st, gacts[cur], wpk + CNN_WPK_BWD + (size_t)l * C * 9 * C, (const float*)nullptr, acts[l],
                           gacts[cur ^ 1], H, W);
        cur ^= 1;
    }
    // input and output layers
    constexpr int NOUT = 9 * C + 1, NIN = C * 9 + C, NP = NOUT + NIN;
    const int px_per = 4096;
    const int iob = (int)((npx + px_per - 1) / px_per);
    hipLaunchKernelGGL(k_io_wgrad, dim3(iob, (NP + 255) / 256), dim3(256), 0, st, dout, acts[6], gacts[cur], x, H, W, n_img, px_per, part);
    // part layout [blk][NP]: out-layer params first, then in-layer params
    hipLaunchKernelGGL(k_reduce_parts, dim3(nb(NP)), dim3(256), 0, st, part, iob, NP, grad + CNN_OFF_OUT, NOUT, grad + CNN_OFF_IN);
    return hipGetLastError();
}

hipError_t cnn_adam(float* params, const float* grad, float* m, float* v, int step, float lr, hipStream_t st) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float bc1 = 1.0f - powf(b1, (float)step), bc2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3(nb(CNN_NPARAM)), dim3(256), 0, st, params, grad, m, v, (int)CNN_NPARAM, lr, b1, b2, eps, bc1, bc2);
    return hipGetLastError();
}

hipError_t cnn_input(const float* L, int w, int h, const float* sigma_dev, float* x, hipStream_t st) {
    if (w <= 0 || h <= 0) return hipSuccess;
    const dim3 grid((unsigned)((w + CI_TW - 1) / CI_TW), (unsigned)((h + CI_TH - 1) / CI_TH));
    hipLaunchKernelGGL(k_cnn_input, grid, dim3(256), 0, st, L, w, h, sigma_dev, x);
    return hipGetLastError();
}

hipError_t cnn_target(const float* x, const float* L, const float* LF, size_t n, const float* sigma_dev, float* t, hipStream_t st) {
    hipLaunchKernelGGL(k_cnn_target, dim3(nb(n)), dim3(256), 0, st, x, L, LF, n, sigma_dev, t);
    return hipGetLastError();
}

hipError_t cnn_crops(const float* a, const float* b, int w, int h, int c, int n, uint64_t seed, float* ca, float* cb, hipStream_t st) {
    hipLaunchKernelGGL(k_crops, dim3(nb((size_t)c * c * n)), dim3(256), 0, st, a, b, w, h, c, n, seed, ca, cb);
    return hipGetLastError();
}

}  // namespace ds
