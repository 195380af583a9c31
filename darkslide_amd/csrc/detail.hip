// The detail stages of the render after the raw path (pipeline/detail.rs):
// noise_reduction (Oklab luminance guided filters, chroma guided filters on
// a subsampled grid) and sharpen (display-encoded luminance unsharp mask
// with halo damping and edge masking), plus the Plane::resize and
// fast_guided_filter primitives they use (image.rs:217-262, :408-431).
//
// Per-pixel arithmetic keeps the reference's operation order and is
// compiled without FMA contraction; the guided filters and box passes are
// the exact running-sum kernels of the denoise path.  The halo damping's
// expf is ds_expf (glibc's bits); powf in the sRGB curves is the device's,
// within an ulp or two of glibc's.

#include <cmath>

#include "box.h"
#include "common.h"
#include "denoise.h"
#include "detail.h"
#include "libm_ref.h"
#include "stages.h"

#pragma clang fp contract(off)

namespace ds {

namespace {

unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

// Plane::resize, shrinking path: the box [x0, x1) x [y0, y1) of the source,
// bounds from f32 products truncated as the reference does.
__global__ void k_resize_box(const float* __restrict__ in, int w, int h, float* __restrict__ out, int w2, int h2, float sx, float sy) {
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w2) return;
    const int x0 = (int)((float)x * sx);
    const int x1 = min(max((int)((float)(x + 1) * sx), x0 + 1), w);
    const int y0 = (int)((float)y * sy);
    const int y1 = min(max((int)((float)(y + 1) * sy), y0 + 1), h);
    float acc = 0.0f;
    for (int v = y0; v < y1; v++)
        for (int u = x0; u < x1; u++) acc += in[(size_t)v * w + u];
    out[(size_t)y * w2 + x] = acc / (float)((x1 - x0) * (y1 - y0));
}

// Plane::sample_bilinear (image.rs:217-228) at the centre of output pixel (x, y).
__device__ __forceinline__ float bilinear(const float* __restrict__ p, int w, int h, int x, int y, float sx, float sy) {
    const float fx = ((float)x + 0.5f) * sx - 0.5f, fy = ((float)y + 0.5f) * sy - 0.5f;
    const float fx0 = floorf(fx), fy0 = floorf(fy);
    const float ax = fx - fx0, ay = fy - fy0;
    const int ix = (int)fx0, iy = (int)fy0;
    auto at = [&](int u, int v) { return p[(size_t)clampi(v, 0, h - 1) * w + clampi(u, 0, w - 1)]; };
    const float a = at(ix, iy) * (1.0f - ax) + at(ix + 1, iy) * ax;
    const float b = at(ix, iy + 1) * (1.0f - ax) + at(ix + 1, iy + 1) * ax;
    return a + (b - a) * ay;
}

__global__ void k_resize_bilinear(const float* __restrict__ in, int w, int h, float* __restrict__ out, int w2, int h2, float sx, float sy) {
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w2) return;
    out[(size_t)y * w2 + x] = bilinear(in, w, h, x, y, sx, sy);
}

// fast_guided_filter's last step (image.rs:428-430): both coefficient
// planes resized up to (w, h), then a * I + b.
__global__ void k_fgf_final(const float* __restrict__ I, const float* __restrict__ ma, const float* __restrict__ mb, int sw, int sh, int w,
                            int h, float sx, float sy, float* __restrict__ out) {
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t i = (size_t)y * w + x;
    const float a = bilinear(ma, sw, sh, x, y, sx, sy), b = bilinear(mb, sw, sh, x, y, sx, sy);
    out[i] = a * I[i] + b;
}

// The bilinear taps of one output pixel (as in bilinear above), shared by
// several coefficient planes of the same size.
struct BilinearTap {
    size_t i00, i10, i01, i11;
    float ax, ay;
    __device__ __forceinline__ BilinearTap(int w, int h, int x, int y, float sx, float sy) {
        const float fx = ((float)x + 0.5f) * sx - 0.5f, fy = ((float)y + 0.5f) * sy - 0.5f;
        const float fx0 = floorf(fx), fy0 = floorf(fy);
        ax = fx - fx0;
        ay = fy - fy0;
        const int ix = (int)fx0, iy = (int)fy0;
        const size_t r0 = (size_t)clampi(iy, 0, h - 1) * w, r1 = (size_t)clampi(iy + 1, 0, h - 1) * w;
        const int u0 = clampi(ix, 0, w - 1), u1 = clampi(ix + 1, 0, w - 1);
        i00 = r0 + u0;
        i10 = r0 + u1;
        i01 = r1 + u0;
        i11 = r1 + u1;
    }
    __device__ __forceinline__ float operator()(const float* __restrict__ p) const {
        const float a = p[i00] * (1.0f - ax) + p[i10] * ax;
        const float b = p[i01] * (1.0f - ax) + p[i11] * ax;
        return a + (b - a) * ay;
    }
};

// image.rs:395-401 for two filter inputs that share the guide: pa, pb ->
// (a1, b1), (a2, b2) from the means of [i, pa, pb, i pa, i pb, i i].
__global__ void k_fgf_pair_products(const float* __restrict__ is, const float* __restrict__ pa, const float* __restrict__ pb, size_t n,
                                    float* __restrict__ ipa, float* __restrict__ ipb, float* __restrict__ ii) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float v = is[i];
    ipa[i] = v * pa[i];
    ipb[i] = v * pb[i];
    ii[i] = v * v;
}

__global__ void k_fgf_pair_coef(const float* __restrict__ mi_, const float* __restrict__ mpa_, const float* __restrict__ mpb_,
                                const float* __restrict__ mipa_, const float* __restrict__ mipb_, const float* __restrict__ mii_, size_t n,
                                float eps, float* __restrict__ a1, float* __restrict__ b1, float* __restrict__ a2, float* __restrict__ b2) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float mi = mi_[i];
    const float var = mii_[i] - mi * mi;
    {
        const float mp = mpa_[i];
        const float av = (mipa_[i] - mi * mp) / (var + eps);
        a1[i] = av;
        b1[i] = mp - av * mi;
    }
    {
        const float mp = mpb_[i];
        const float av = (mipb_[i] - mi * mp) / (var + eps);
        a2[i] = av;
        b2[i] = mp - av * mi;
    }
}

// image.rs:428-430 (or :404 without subsampling) for both planes, fused
// with detail.rs:80-81: o + (a I + b - o) k, in place.
template <bool FAST>
__global__ void k_fgf_pair_final_mix(const float* __restrict__ I, const float* __restrict__ ma1, const float* __restrict__ mb1,
                                     const float* __restrict__ ma2, const float* __restrict__ mb2, int sw, int sh, int w, int h, float sx,
                                     float sy, float k, float* __restrict__ A, float* __restrict__ B) {
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t i = (size_t)y * w + x;
    float a1, b1, a2, b2;
    if (FAST) {
        const BilinearTap t(sw, sh, x, y, sx, sy);
        a1 = t(ma1);
        b1 = t(mb1);
        a2 = t(ma2);
        b2 = t(mb2);
    } else {
        a1 = ma1[i];
        b1 = mb1[i];
        a2 = ma2[i];
        b2 = mb2[i];
    }
    const float iv = I[i];
    const float fa = a1 * iv + b1, fb = a2 * iv + b2;
    const float oa = A[i], ob = B[i];
    A[i] = oa + (fa - oa) * k;
    B[i] = ob + (fb - ob) * k;
}

// detail.rs:44-55: the filtered luminance blended in by the local strength.
__global__ void k_nr_luma(float* __restrict__ L, const float* __restrict__ f2, const float* __restrict__ field, size_t n, float lum, float contrast,
                          float max_strength) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float s = lum + (field ? field[i] : 0.0f) / 100.0f;
    s = clampr(s, 0.0f, 1.0f);
    const float l = L[i], f = f2[i];
    const float filt = f + (l - f) * contrast * 0.35f;
    L[i] = l + (filt - l) * fminf(s / fmaxr(max_strength, 1e-6f), 1.0f);
}

// detail.rs:85-89: Oklab back to ProPhoto (color.rs:415-418), clamped at 0.
__global__ void k_nr_compose(const float* __restrict__ L, const float* __restrict__ A, const float* __restrict__ B, size_t n, OklabM ok,
                             float* __restrict__ rgb) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float v[3] = {L[i], A[i], B[i]};
    float l[3], cube[3], p[3];
    mul_mv(ok.m2_inv, v, l);
    cube[0] = l[0] * l[0] * l[0];
    cube[1] = l[1] * l[1] * l[1];
    cube[2] = l[2] * l[2] * l[2];
    mul_mv(ok.from_lms, cube, p);
    rgb[3 * i] = fmaxr(p[0], 0.0f);
    rgb[3 * i + 1] = fmaxr(p[1], 0.0f);
    rgb[3 * i + 2] = fmaxr(p[2], 0.0f);
}

// sRGB curves (color.rs:185-192) and their extension above 1 (pipeline/tone.rs:328-335)
__device__ __forceinline__ float tone_enc(float v) {
    if (v <= 1.0f) return v <= 0.0031308f ? 12.92f * fmaxr(v, 0.0f) : 1.055f * powf(v, 1.0f / 2.4f) - 0.055f;
    return 1.0f + (v - 1.0f) / 2.4f;
}
__device__ __forceinline__ float tone_dec(float e) {
    if (e <= 1.0f) return e <= 0.04045f ? e / 12.92f : powf((e + 0.055f) / 1.055f, 2.4f);
    return 1.0f + (e - 1.0f) * 2.4f;
}

// detail.rs:105: enc(luma_prophoto(p).max(0)) (color.rs:196-198)
__global__ void k_sharp_enc(const float* __restrict__ rgb, size_t n, float* __restrict__ e) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float y = 0.2880402f * rgb[3 * i] + 0.7118741f * rgb[3 * i + 1] + 0.0000857f * rgb[3 * i + 2];
    e[i] = tone_enc(fmaxr(y, 0.0f));
}

// detail.rs:108-138: edge mask from the blurred encoding s (when masking),
// halo-damped difference, luminance ratio applied to all three channels.
__global__ void k_sharp_apply(float* __restrict__ rgb, const float* __restrict__ e, const float* __restrict__ blur, const float* __restrict__ s,
                              const float* __restrict__ field, int w, int h, float amount, float detail, float t) {
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t i = (size_t)y * w + x;
    const float amt = (amount + (field ? field[i] : 0.0f) * 1.5f) / 100.0f * 1.6f;
    if (amt == 0.0f) return;
    const float ev = e[i];
    const float d = ev - blur[i];
    const float damp = detail + (1.0f - detail) * ds_expf(-fabsf(d) / 0.04f);
    float dv = d * damp;
    if (s) {
        auto at = [&](int u, int v) { return s[(size_t)clampi(v, 0, h - 1) * w + clampi(u, 0, w - 1)]; };
        const float gx = at(x + 1, y) - at(x - 1, y);
        const float gy = at(x, y + 1) - at(x, y - 1);
        const float g = sqrtf(gx * gx + gy * gy);
        // smoothstep (pipeline/tone.rs:16-19)
        const float a = t * 0.5f, b = t * 1.5f + 1e-4f;
        const float q = clampr((g - a) / (b - a), 0.0f, 1.0f);
        dv *= q * q * (3.0f - 2.0f * q);
    }
    const float e2 = fmaxr(ev + amt * dv, 0.0f);
    const float y0 = fmaxr(tone_dec(ev), 1e-6f);
    const float y1 = tone_dec(e2);
    const float r = y1 / y0;
    rgb[3 * i] = rgb[3 * i] * r;
    rgb[3 * i + 1] = rgb[3 * i + 1] * r;
    rgb[3 * i + 2] = rgb[3 * i + 2] * r;
}

// The fold at detail.rs:34: max over the field starting from 0 (f32::max
// skips NaN, as fmaxf does); partials per block, then one block.
__global__ __launch_bounds__(256) void k_field_max(const float* __restrict__ f, size_t n, float* __restrict__ partial) {
    float m = 0.0f;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) m = fmaxr(m, f[i]);
    for (int o = 32; o > 0; o >>= 1) m = fmaxr(m, __shfl_xor(m, o));
    __shared__ float red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = fmaxr(fmaxr(red[0], red[1]), fmaxr(red[2], red[3]));
}

__global__ __launch_bounds__(256) void k_field_max_final(const float* __restrict__ partial, int parts, float* __restrict__ out) {
    float m = 0.0f;
    for (int i = threadIdx.x; i < parts; i += 256) m = fmaxr(m, partial[i]);
    for (int o = 32; o > 0; o >>= 1) m = fmaxr(m, __shfl_xor(m, o));
    __shared__ float red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = fmaxr(fmaxr(red[0], red[1]), fmaxr(red[2], red[3]));
}

}  // namespace

hipError_t launch_field_max(const float* field, size_t n, float* partial, float* out, hipStream_t st) {
    hipLaunchKernelGGL(k_field_max, dim3(FIELD_MAX_PARTS), dim3(256), 0, st, field, n, partial);
    hipLaunchKernelGGL(k_field_max_final, dim3(1), dim3(256), 0, st, partial, FIELD_MAX_PARTS, out);
    return hipGetLastError();
}

hipError_t launch_plane_resize(const float* in, size_t w, size_t h, size_t w2, size_t h2, float* out, hipStream_t st) {
    if (w2 == 0 || h2 == 0) return hipSuccess;
    if (w2 == w && h2 == h) return hipMemcpyAsync(out, in, w * h * sizeof(float), hipMemcpyDeviceToDevice, st);
    const float sx = (float)w / (float)w2, sy = (float)h / (float)h2;
    const dim3 grid(nblk(w2), (unsigned)h2);
    if (sx > 1.5f || sy > 1.5f)
        hipLaunchKernelGGL(k_resize_box, grid, dim3(256), 0, st, in, (int)w, (int)h, out, (int)w2, (int)h2, sx, sy);
    else
        hipLaunchKernelGGL(k_resize_bilinear, grid, dim3(256), 0, st, in, (int)w, (int)h, out, (int)w2, (int)h2, sx, sy);
    return hipGetLastError();
}

size_t fgf_ws_floats(size_t w, size_t h, size_t sub) {
    if (sub <= 1 || w / sub < 8 || h / sub < 8) return 8 * w * h;
    return 10 * (w / sub) * (h / sub);
}

hipError_t launch_fast_guided_filter(const float* I, const float* P, size_t w, size_t h, size_t r, float eps, size_t sub, float* ws,
                                     float* out, hipStream_t st) {
    if (w == 0 || h == 0) return hipSuccess;
    if (sub <= 1 || w / sub < 8 || h / sub < 8) return launch_guided_filter(I, P, w, h, r, eps, ws, out, st);
    const size_t sw = w / sub, sh = h / sub, ns = sw * sh, rs = r / sub > 1 ? r / sub : 1;
    float* is = ws;
    float* ps = ws + ns;
    float* gws = ws + 2 * ns;  // 8 planes; the blurred coefficients end in planes 4 and 5
    hipError_t e = launch_plane_resize(I, w, h, sw, sh, is, st);
    if (e == hipSuccess) e = launch_plane_resize(P, w, h, sw, sh, ps, st);
    if (e == hipSuccess) e = launch_guided_coeffs(is, ps, sw, sh, rs, eps, gws, st);
    if (e != hipSuccess) return e;
    const float sx = (float)sw / (float)w, sy = (float)sh / (float)h;
    hipLaunchKernelGGL(k_fgf_final, dim3(nblk(w), (unsigned)h), dim3(256), 0, st, I, gws + 4 * ns, gws + 5 * ns, (int)sw, (int)sh, (int)w,
                       (int)h, sx, sy, out);
    return hipGetLastError();
}

size_t fgf_pair_ws_floats(size_t w, size_t h, size_t sub) {
    const bool fast = !(sub <= 1 || w / sub < 8 || h / sub < 8);
    return 12 * (fast ? (w / sub) * (h / sub) : w * h);
}

hipError_t launch_fgf_pair_mix(const float* I, float* A, float* B, size_t w, size_t h, size_t r, float eps, size_t sub, float k, float* ws,
                               hipStream_t st) {
    if (w == 0 || h == 0) return hipSuccess;
    const bool fast = !(sub <= 1 || w / sub < 8 || h / sub < 8);
    const size_t sw = fast ? w / sub : w, sh = fast ? h / sub : h, ns = sw * sh;
    const size_t rs = fast ? (r / sub > 1 ? r / sub : 1) : r;
    float* q[12];
    for (int j = 0; j < 12; j++) q[j] = ws + (size_t)j * ns;
    const float *is = I, *pa = A, *pb = B;
    hipError_t e = hipSuccess;
    if (fast) {
        e = launch_plane_resize(I, w, h, sw, sh, q[0], st);
        if (e == hipSuccess) e = launch_plane_resize(A, w, h, sw, sh, q[1], st);
        if (e == hipSuccess) e = launch_plane_resize(B, w, h, sw, sh, q[2], st);
        if (e != hipSuccess) return e;
        is = q[0];
        pa = q[1];
        pb = q[2];
    }
    hipLaunchKernelGGL(k_fgf_pair_products, dim3(nblk(ns)), dim3(256), 0, st, is, pa, pb, ns, q[3], q[4], q[5]);
    CBoxPlanes hin{};
    BoxPlanes hout{}, vout{};
    const float* src[6] = {is, pa, pb, q[3], q[4], q[5]};
    for (int j = 0; j < 6; j++) {
        hin.p[j] = src[j];
        hout.p[j] = q[6 + j];
        vout.p[j] = q[j];
    }
    CBoxPlanes vin{};
    for (int j = 0; j < 6; j++) vin.p[j] = q[6 + j];
    e = box_h_planes(hin, hout, 6, sw, sh, rs, st);
    if (e == hipSuccess) e = box_v_planes(vin, vout, 6, sw, sh, rs, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_fgf_pair_coef, dim3(nblk(ns)), dim3(256), 0, st, q[0], q[1], q[2], q[3], q[4], q[5], ns, eps, q[6], q[7], q[8],
                       q[9]);
    CBoxPlanes cin{}, cmid{};
    BoxPlanes mid{}, fin{};
    for (int j = 0; j < 4; j++) {
        cin.p[j] = q[6 + j];
        mid.p[j] = q[j];
        cmid.p[j] = q[j];
        fin.p[j] = q[6 + j];
    }
    e = box_h_planes(cin, mid, 4, sw, sh, rs, st);
    if (e == hipSuccess) e = box_v_planes(cmid, fin, 4, sw, sh, rs, st);
    if (e != hipSuccess) return e;
    const dim3 grid(nblk(w), (unsigned)h);
    if (fast)
        hipLaunchKernelGGL(k_fgf_pair_final_mix<true>, grid, dim3(256), 0, st, I, q[6], q[7], q[8], q[9], (int)sw, (int)sh, (int)w, (int)h,
                           (float)sw / (float)w, (float)sh / (float)h, k, A, B);
    else
        hipLaunchKernelGGL(k_fgf_pair_final_mix<false>, grid, dim3(256), 0, st, I, q[6], q[7], q[8], q[9], (int)sw, (int)sh, (int)w, (int)h,
                           1.0f, 1.0f, k, A, B);
    return hipGetLastError();
}

size_t nr_ws_floats(size_t w, size_t h) {
    const size_t n = w * h;
    return 15 * n;  // L A B f1 f2 and the luminance guided filters' 8 planes, or L A B and the chroma pair's 12
}

hipError_t launch_noise_reduction(float* rgb, size_t w, size_t h, const NrParams& nr, const float* field, float field_max, float scale,
                                  const OklabM& ok, float* ws, hipStream_t st) {
    const bool local = field != nullptr;
    const float lum = nr.luminance / 100.0f, col = nr.color / 100.0f;
    if (lum <= 0.0f && col <= 0.0f && !local) return hipSuccess;
    if (scale < 0.34f) return hipSuccess;  // the downscale has averaged the noise away
    const size_t n = w * h;
    if (n == 0) return hipSuccess;
    float eff = scale;
    if (eff < 0.05f) eff = 0.05f;
    if (eff > 1.0f) eff = 1.0f;
    float* L = ws;
    float* A = ws + n;
    float* B = ws + 2 * n;
    float* f1 = ws + 3 * n;
    float* f2 = ws + 4 * n;
    float* gws = ws + 5 * n;
    hipError_t e = launch_oklab_fwd(rgb, n, ok, L, A, B, st);
    if (e != hipSuccess) return e;
    bool changed = false;
    if (lum > 0.0f || local) {
        float ms = lum + (local ? field_max / 100.0f : 0.0f);
        ms = ms < 0.0f ? 0.0f : (ms > 1.0f ? 1.0f : ms);
        if (ms > 0.0f) {
            const float detail = nr.luminance_detail / 100.0f;
            const float sigma = 0.004f + ms * 0.045f * sqrtf(eff);
            const float q = sigma * (1.4f - detail);
            const float eps = q * q;
            const float rr1 = roundf(1.5f * eff), rr2 = roundf(4.0f * eff);
            const size_t r1 = rr1 > 1.0f ? (size_t)rr1 : 1, r2 = rr2 > 1.0f ? (size_t)rr2 : 1;
            e = launch_guided_filter(L, L, w, h, r1, eps, gws, f1, st);
            if (e == hipSuccess) e = launch_guided_filter(f1, f1, w, h, r2, eps * 0.5f, gws, f2, st);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(k_nr_luma, dim3(nblk(n)), dim3(256), 0, st, L, f2, field, n, lum, nr.luminance_contrast / 100.0f, ms);
            changed = true;
        }
    }
    if (col > 0.0f) {
        const float detail = nr.color_detail / 100.0f, smooth = nr.color_smoothness / 100.0f;
        const float rf = fmaxf(roundf((2.0f + col * 10.0f) * (1.0f + smooth * 1.5f) * eff), 1.0f);
        const size_t r = (size_t)rf;
        const float q = 0.002f + (1.0f - detail) * 0.02f;
        const float eps = q * q;
        const size_t sub = r >= 8 ? 4 : (r >= 3 ? 2 : 1);
        e = launch_fgf_pair_mix(L, A, B, w, h, r, eps, sub, fminf(col, 1.0f), ws + 3 * n, st);
        if (e != hipSuccess) return e;
        changed = true;
    }
    if (changed) hipLaunchKernelGGL(k_nr_compose, dim3(nblk(n)), dim3(256), 0, st, L, A, B, n, ok, rgb);
    return hipGetLastError();
}

size_t sharpen_ws_floats(size_t w, size_t h) { return 5 * w * h; }

hipError_t launch_sharpen(float* rgb, size_t w, size_t h, const SharpParams& sh, const float* field, float scale, float* ws,
                          hipStream_t st) {
    const bool local = field != nullptr;
    if (sh.amount <= 0.0f && !local) return hipSuccess;
    const float sigma = sh.radius * fminf(scale, 1.0f);
    if (sigma < 0.25f) return hipSuccess;
    const size_t n = w * h;
    if (n == 0) return hipSuccess;
    float* ev = ws;
    float* blur = ws + n;
    float* s = ws + 2 * n;
    float* gws = ws + 3 * n;
    hipLaunchKernelGGL(k_sharp_enc, dim3(nblk(n)), dim3(256), 0, st, rgb, n, ev);
    hipError_t e = launch_gaussian_blur(ev, gws, blur, w, h, sigma, st);
    if (e != hipSuccess) return e;
    const bool mask = sh.masking > 0.0f;
    float t = 0.0f;
    if (mask) {
        e = launch_gaussian_blur(ev, gws, s, w, h, 1.0f * fmaxf(scale, 0.5f), st);
        if (e != hipSuccess) return e;
        t = sh.masking / 100.0f * 0.06f;
    }
    hipLaunchKernelGGL(k_sharp_apply, dim3(nblk(w), (unsigned)h), dim3(256), 0, st, rgb, ev, blur, mask ? s : nullptr, field, (int)w, (int)h,
                       sh.amount, sh.detail / 100.0f, t);
    return hipGetLastError();
}

// enhance.rs:123: y = luminance().max(0).sqrt()
__global__ void k_superres_y(const float* __restrict__ rgb, size_t n, float* __restrict__ y) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float l = 0.2880402f * rgb[3 * i] + 0.7118741f * rgb[3 * i + 1] + 0.0000857f * rgb[3 * i + 2];
    y[i] = sqrtf(fmaxr(l, 0.0f));
}

// enhance.rs:125-130
__global__ void k_superres_apply(float* __restrict__ rgb, const float* __restrict__ y, const float* __restrict__ b, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float yv = y[i];
    const float d = clampr(yv - b[i], -0.05f, 0.05f);
    const float nv = fmaxr(yv + d * 0.9f, 0.0f);
    const float r = clampr((nv * nv) / fmaxr(yv * yv, 1e-8f), 0.5f, 2.0f);
    rgb[3 * i] = rgb[3 * i] * r;
    rgb[3 * i + 1] = rgb[3 * i + 1] * r;
    rgb[3 * i + 2] = rgb[3 * i + 2] * r;
}

// export.rs:216-222
__global__ void k_outsharp_apply(float* __restrict__ rgb, const float* __restrict__ l, const float* __restrict__ b, size_t n, float k) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float lv = l[i];
    const float d = lv - b[i];
    const float e2 = fmaxr(lv + k * clampr(d, -0.1f, 0.1f), 0.0f);
    const float y0 = fmaxr(tone_dec(lv), 1e-6f);
    const float r = tone_dec(e2) / y0;
    rgb[3 * i] = rgb[3 * i] * r;
    rgb[3 * i + 1] = rgb[3 * i + 1] * r;
    rgb[3 * i + 2] = rgb[3 * i + 2] * r;
}

size_t export_detail_ws_floats(size_t w, size_t h) { return 4 * w * h; }

hipError_t launch_super_res_refine(float* rgb, size_t w, size_t h, float* ws, hipStream_t st) {
    const size_t n = w * h;
    if (n == 0) return hipSuccess;
    float* y = ws;
    float* b = ws + n;
    hipLaunchKernelGGL(k_superres_y, dim3(nblk(n)), dim3(256), 0, st, rgb, n, y);
    hipError_t e = launch_gaussian_blur(y, ws + 2 * n, b, w, h, 1.2f, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_superres_apply, dim3(nblk(n)), dim3(256), 0, st, rgb, y, b, n);
    return hipGetLastError();
}

hipError_t launch_output_sharpen(float* rgb, size_t w, size_t h, int mode, int amount, float* ws, hipStream_t st) {
    float radius, base;
    switch (mode) {
        case OUT_SHARPEN_SCREEN: radius = 0.6f, base = 0.45f; break;
        case OUT_SHARPEN_MATTE: radius = 1.0f, base = 0.8f; break;
        case OUT_SHARPEN_GLOSSY: radius = 0.8f, base = 0.6f; break;
        default: return hipSuccess;
    }
    const float k = base * (amount == OUT_AMOUNT_LOW ? 0.6f : amount == OUT_AMOUNT_HIGH ? 1.5f : 1.0f);
    const size_t n = w * h;
    if (n == 0) return hipSuccess;
    float* l = ws;
    float* b = ws + n;
    hipLaunchKernelGGL(k_sharp_enc, dim3(nblk(n)), dim3(256), 0, st, rgb, n, l);
    hipError_t e = launch_gaussian_blur(l, ws + 2 * n, b, w, h, radius, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_outsharp_apply, dim3(nblk(n)), dim3(256), 0, st, rgb, l, b, n, k);
    return hipGetLastError();
}

}  // namespace ds
