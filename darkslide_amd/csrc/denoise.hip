// Denoise composite (enhance.rs:74-92) for gfx950: Oklab split, noise
// estimate (enhance.rs:9-32), box filters (image.rs:268-313) and the two
// guided filters (image.rs:384-405) around the NLM of the L channel.
//
// Box filters run as separable passes.  The horizontal pass stages a row
// segment plus its clamped halo in LDS and slides a (2r+1)-window along it;
// the vertical pass gives each thread one column and slides down a row
// segment, so each load instruction of a wave reads one contiguous row
// piece.  Inputs of the first pass and outputs of the last are computed in
// functors, which fuses the guided filter's products, coefficients and the
// inverse Oklab into the passes.

#include "common.h"
#include "denoise.h"
#include "nlm.h"

#pragma clang fp contract(off)

namespace ds {

namespace {

constexpr int BH_NT = 256;
constexpr int BH_SEG = 8;                    // outputs per thread, horizontal
constexpr int BH_SPAN = BH_NT * BH_SEG;      // outputs per block
constexpr int BV_NT = 256;
constexpr int BV_ROWS = 64;                  // outputs per thread, vertical

template <int NP>
struct Planes {
    float* p[NP];
};
template <int NP>
struct CPlanes {
    const float* p[NP];
};

// ---- input functors: value(s) at in-image pixel i
template <int NP>
struct InPlanes {
    CPlanes<NP> src;
    __device__ __forceinline__ void operator()(size_t i, float* v) const {
#pragma unroll
        for (int k = 0; k < NP; k++) v[k] = src.p[k][i];
    }
};

// guided filter inputs for both chroma planes: I = l2 = L + (lf - L) m
// (enhance.rs:86), then [I, a, b, I a, I b, I I] (image.rs:386-391)
struct InGuided {
    const float* L;
    const float* LF;
    const float* A;
    const float* B;
    float m;
    __device__ __forceinline__ void operator()(size_t i, float* v) const {
        float o = L[i];
        float I = o + (LF[i] - o) * m;
        float a = A[i], b = B[i];
        v[0] = I;
        v[1] = a;
        v[2] = b;
        v[3] = I * a;
        v[4] = I * b;
        v[5] = I * I;
    }
};

// ---- output functors of the vertical pass
template <int NP>
struct OutPlanes {
    Planes<NP> dst;
    __device__ __forceinline__ void operator()(size_t i, const float* v) const {
#pragma unroll
        for (int k = 0; k < NP; k++) dst.p[k][i] = v[k];
    }
};

// image.rs:395-400 for the a and b planes, eps = (2 max(sigma,1e-4))^2
// (enhance.rs:82, :88)
struct OutCoef {
    float* a_a;
    float* b_a;
    float* a_b;
    float* b_b;
    const float* sigma_dev;
    __device__ __forceinline__ void operator()(size_t i, const float* v) const {
        float s2 = fmaxf(*sigma_dev, 1e-4f) * 2.0f;
        float eps = s2 * s2;
        float mi = v[0], mpa = v[1], mpb = v[2], mipa = v[3], mipb = v[4], mii = v[5];
        float var = mii - mi * mi;
        float cov_a = mipa - mi * mpa;
        float aa = cov_a / (var + eps);
        a_a[i] = aa;
        b_a[i] = mpa - aa * mi;
        float cov_b = mipb - mi * mpb;
        float ab = cov_b / (var + eps);
        a_b[i] = ab;
        b_b[i] = mpb - ab * mi;
    }
};

// image.rs:404 a*I + b for both planes, then Oklab -> ProPhoto and max(0)
// (enhance.rs:91)
struct OutFinal {
    const float* L;
    const float* LF;
    float m;
    OklabM ok;
    float* rgb;
    __device__ __forceinline__ void operator()(size_t i, const float* v) const {
        float o = L[i];
        float I = o + (LF[i] - o) * m;
        float lab[3] = {I, v[0] * I + v[1], v[2] * I + v[3]};
        float l[3], cube[3], p[3];
        mul_mv(ok.m2_inv, lab, l);
        cube[0] = l[0] * l[0] * l[0];
        cube[1] = l[1] * l[1] * l[1];
        cube[2] = l[2] * l[2] * l[2];
        mul_mv(ok.from_lms, cube, p);
        rgb[3 * i + 0] = fmaxr(p[0], 0.0f);
        rgb[3 * i + 1] = fmaxr(p[1], 0.0f);
        rgb[3 * i + 2] = fmaxr(p[2], 0.0f);
    }
};

// Horizontal (2r+1)-box, output scaled by norm (image.rs:276-288).
template <int NP, class In>
__global__ __launch_bounds__(BH_NT) void k_box_h(In in, Planes<NP> out, int w, int h, int r, float norm) {
    extern __shared__ float lds[];
    const int span = BH_SPAN + 2 * r;
    const int y = blockIdx.y;
    const int x0 = blockIdx.x * BH_SPAN;
    const size_t row = (size_t)y * (size_t)w;
    for (int i = threadIdx.x; i < span; i += BH_NT) {
        int gx = clampi(x0 - r + i, 0, w - 1);
        float v[NP];
        in(row + (size_t)gx, v);
#pragma unroll
        for (int k = 0; k < NP; k++) lds[k * span + i] = v[k];
    }
    __syncthreads();
    const int b = threadIdx.x * BH_SEG;
    if (x0 + b >= w) return;
#pragma unroll
    for (int k = 0; k < NP; k++) {
        const float* s = lds + k * span + b;
        float acc = 0.0f;
        for (int i = 0; i <= 2 * r; i++) acc += s[i];
#pragma unroll
        for (int j = 0; j < BH_SEG; j++) {
            int gx = x0 + b + j;
            if (gx < w) out.p[k][row + (size_t)gx] = acc * norm;
            if (j + 1 < BH_SEG) acc += s[j + 2 * r + 1] - s[j];
        }
    }
}

// Vertical (2r+1)-box, output scaled by norm (image.rs:290-311).
template <int NP, class Out>
__global__ __launch_bounds__(BV_NT) void k_box_v(CPlanes<NP> in, Out out, int w, int h, int r, float norm) {
    const int x = blockIdx.x * BV_NT + threadIdx.x;
    if (x >= w) return;
    const int y0 = blockIdx.y * BV_ROWS;
    const int y1 = min(y0 + BV_ROWS, h);
    float acc[NP];
#pragma unroll
    for (int k = 0; k < NP; k++) acc[k] = 0.0f;
    for (int j = -r; j <= r; j++) {
        size_t i = (size_t)clampi(y0 + j, 0, h - 1) * (size_t)w + (size_t)x;
#pragma unroll
        for (int k = 0; k < NP; k++) acc[k] += in.p[k][i];
    }
    for (int y = y0; y < y1; y++) {
        float v[NP];
#pragma unroll
        for (int k = 0; k < NP; k++) v[k] = acc[k] * norm;
        out((size_t)y * (size_t)w + (size_t)x, v);
        size_t ia = (size_t)clampi(y + r + 1, 0, h - 1) * (size_t)w + (size_t)x;
        size_t ib = (size_t)clampi(y - r, 0, h - 1) * (size_t)w + (size_t)x;
#pragma unroll
        for (int k = 0; k < NP; k++) acc[k] += in.p[k][ia] - in.p[k][ib];
    }
}

template <int NP, class In>
hipError_t box_h(In in, Planes<NP> out, size_t w, size_t h, size_t r, hipStream_t st) {
    const float norm = 1.0f / (float)(2 * r + 1);
    size_t lds = (size_t)NP * (BH_SPAN + 2 * r) * sizeof(float);
    if (lds > 150 * 1024) return hipErrorInvalidValue;
    static bool attr = false;
    if (!attr) {
        hipError_t e = hipFuncSetAttribute((const void*)k_box_h<NP, In>, hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024);
        if (e != hipSuccess) return e;
        attr = true;
    }
    dim3 grid((unsigned)((w + BH_SPAN - 1) / BH_SPAN), (unsigned)h);
    hipLaunchKernelGGL((k_box_h<NP, In>), grid, dim3(BH_NT), lds, st, in, out, (int)w, (int)h, (int)r, norm);
    return hipGetLastError();
}

template <int NP, class Out>
hipError_t box_v(CPlanes<NP> in, Out out, size_t w, size_t h, size_t r, hipStream_t st) {
    const float norm = 1.0f / (float)(2 * r + 1);
    dim3 grid((unsigned)((w + BV_NT - 1) / BV_NT), (unsigned)((h + BV_ROWS - 1) / BV_ROWS));
    hipLaunchKernelGGL((k_box_v<NP, Out>), grid, dim3(BV_NT), 0, st, in, out, (int)w, (int)h, (int)r, norm);
    return hipGetLastError();
}

// ---- Oklab split (color.rs:408-412)
__global__ void k_oklab_fwd(const float* __restrict__ rgb, size_t n, OklabM ok, float* __restrict__ L, float* __restrict__ A,
                            float* __restrict__ B) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float p[3] = {rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]};
    float lms[3], c[3], lab[3];
    mul_mv(ok.to_lms, p, lms);
    c[0] = cbrtf(lms[0]);
    c[1] = cbrtf(lms[1]);
    c[2] = cbrtf(lms[2]);
    mul_mv(ok.m2, c, lab);
    L[i] = lab[0];
    A[i] = lab[1];
    B[i] = lab[2];
}

// ---- noise estimate: every step-th pixel off the 1-px border, |laplacian|
// as ordered u32 bits; non-samples get +inf and sort last.
constexpr uint32_t kInfBits = 0x7f800000u;

__global__ void k_noise_sample(const float* __restrict__ p, size_t w, size_t h, size_t step, size_t count, uint32_t* __restrict__ vals) {
    size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    size_t i = k * step;
    size_t x = i % w, y = i / w;
    uint32_t bits = kInfBits;
    if (x >= 1 && y >= 1 && x + 1 < w && y + 1 < h) {
        float c = p[i];
        float lap = 4.0f * c - p[i - 1] - p[i + 1] - p[i - w] - p[i + w];
        bits = __float_as_uint(fabsf(lap));
    }
    vals[k] = bits;
}

// Exact upper median of the valid samples (select_nth_unstable at len/2) by
// a 4 x 8-bit MSB-first radix select in one workgroup, then
// sigma = v / 0.6745 / sqrt(20).
constexpr int SEL_NT = 1024;
__global__ __launch_bounds__(SEL_NT) void k_noise_select(const uint32_t* __restrict__ vals, size_t count, float sqrt20,
                                                         float* __restrict__ sigma_out) {
    __shared__ uint32_t hist[256];
    __shared__ uint32_t s_red[SEL_NT / 64];
    __shared__ uint32_t s_prefix, s_target;
    const int tid = threadIdx.x;
    uint32_t nv = 0;
    for (size_t k = tid; k < count; k += SEL_NT) nv += vals[k] != kInfBits ? 1u : 0u;
    for (int off = 32; off > 0; off >>= 1) nv += __shfl_down(nv, off, 64);
    if ((tid & 63) == 0) s_red[tid >> 6] = nv;
    __syncthreads();
    if (tid == 0) {
        uint32_t t = 0;
        for (int i = 0; i < SEL_NT / 64; i++) t += s_red[i];
        s_target = t / 2;
        s_prefix = 0;
        if (t == 0) *sigma_out = 0.0f;  // enhance.rs:25-27
        s_red[0] = t;
    }
    __syncthreads();
    if (s_red[0] == 0) return;
    uint32_t mask = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int i = tid; i < 256; i += SEL_NT) hist[i] = 0;
        __syncthreads();
        const uint32_t prefix = s_prefix;
        for (size_t k = tid; k < count; k += SEL_NT) {
            uint32_t v = vals[k];
            if ((v & mask) == prefix) atomicAdd(&hist[(v >> shift) & 255u], 1u);
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t target = s_target, cum = 0;
            int b = 0;
            for (; b < 256; b++) {
                if (cum + hist[b] > target) break;
                cum += hist[b];
            }
            s_target = target - cum;
            s_prefix = prefix | ((uint32_t)b << shift);
        }
        mask |= 255u << shift;
        __syncthreads();
    }
    if (tid == 0) *sigma_out = __uint_as_float(s_prefix) / 0.6745f / sqrt20;
}

}  // namespace

hipError_t launch_oklab_fwd(const float* rgb, size_t n, const OklabM& ok, float* L, float* A, float* B, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_oklab_fwd, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, rgb, n, ok, L, A, B);
    return hipGetLastError();
}

size_t noise_sample_count(size_t w, size_t h) {
    size_t n = w * h;
    size_t step = n / 200000 > 1 ? n / 200000 : 1;  // enhance.rs:12
    return (n + step - 1) / step;
}

hipError_t launch_estimate_noise(const float* plane, size_t w, size_t h, uint32_t* scratch, float* sigma_out, hipStream_t st) {
    size_t n = w * h;
    if (n == 0) {
        hipLaunchKernelGGL(k_noise_select, dim3(1), dim3(SEL_NT), 0, st, scratch, (size_t)0, 1.0f, sigma_out);
        return hipGetLastError();
    }
    size_t step = n / 200000 > 1 ? n / 200000 : 1;
    size_t count = (n + step - 1) / step;
    hipLaunchKernelGGL(k_noise_sample, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, plane, w, h, step, count, scratch);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_noise_select, dim3(1), dim3(SEL_NT), 0, st, scratch, count, sqrtf(20.0f), sigma_out);
    return hipGetLastError();
}

hipError_t launch_box_blur(const float* in, float* tmp, float* out, size_t w, size_t h, size_t r, hipStream_t st) {
    if (w == 0 || h == 0) return hipSuccess;
    if (r == 0) return hipMemcpyAsync(out, in, w * h * sizeof(float), hipMemcpyDeviceToDevice, st);
    InPlanes<1> hin{{{in}}};
    hipError_t e = box_h<1>(hin, Planes<1>{{tmp}}, w, h, r, st);
    if (e != hipSuccess) return e;
    return box_v<1>(CPlanes<1>{{tmp}}, OutPlanes<1>{{{out}}}, w, h, r, st);
}

// Guided filters of a and b by l2 = L + (lf - L) m, then Oklab^-1 into rgb.
// ws: 10 planes of w*h floats.
hipError_t launch_guided_pair(const float* L, const float* LF, const float* A, const float* B, float m, size_t r, const float* sigma_dev,
                              const OklabM& ok, float* ws, float* rgb, size_t w, size_t h, hipStream_t st) {
    const size_t n = w * h;
    if (n == 0) return hipSuccess;
    float* t[10];
    for (int k = 0; k < 10; k++) t[k] = ws + (size_t)k * n;
    // r = 0 gives norm 1 and a one-tap window: the identity, as box_blur's
    // r == 0 early return (image.rs:270-272)
    InGuided gin{L, LF, A, B, m};
    hipError_t e = box_h<6>(gin, Planes<6>{{t[0], t[1], t[2], t[3], t[4], t[5]}}, w, h, r, st);
    if (e != hipSuccess) return e;
    OutCoef coef{t[6], t[7], t[8], t[9], sigma_dev};
    e = box_v<6>(CPlanes<6>{{t[0], t[1], t[2], t[3], t[4], t[5]}}, coef, w, h, r, st);
    if (e != hipSuccess) return e;
    e = box_h<4>(InPlanes<4>{{{t[6], t[7], t[8], t[9]}}}, Planes<4>{{t[0], t[1], t[2], t[3]}}, w, h, r, st);
    if (e != hipSuccess) return e;
    OutFinal fin{L, LF, m, ok, rgb};
    return box_v<4>(CPlanes<4>{{t[0], t[1], t[2], t[3]}}, fin, w, h, r, st);
}

}  // namespace ds
