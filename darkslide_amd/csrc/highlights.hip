// Highlight reconstruction (demosaic.rs:435-502) for gfx950.
//
// Stages, all on one stream and all gated on a device flag so that a frame
// with no clipped sample costs two tiny launches and no host round trip:
//   1. flag := any(clip > 0)                      (the reference's early return)
//   2. near = box_blur(clip, 3)                   (exact running-sum passes)
//   3. weight planes wgt / rr / br per pixel      (demosaic.rs:448-459)
//   4. three-box gaussian of the three planes at once, sigma = max(w,h)/120 >= 2
//   5. per-pixel rebuild of the clipped channels  (demosaic.rs:463-501)
// Per-pixel arithmetic keeps the reference's operation order and is compiled
// without FMA contraction, so the result equals the FP32 oracle bit for bit.

#include <algorithm>

#include "box.h"
#include "common.h"
#include "highlights.h"
#include "stages.h"

#pragma clang fp contract(off)

namespace ds {

namespace {

constexpr size_t HL_HDR = 64;  // floats before the planes (flag at word 0)
constexpr float kF32Max = 3.40282347e+38f;

struct Mul3 {
    float m[3];
};

// Clip thresholds of demosaic.rs:439-445, from raw_clip (device) and mul.
struct Levels {
    float lvl[3];
    float lvl8[3];  // lvl * 0.8, the threshold next to a clipped CFA sample
    float lmin;
    __device__ __forceinline__ void load(const float* raw_clip, const Mul3& mul) {
        for (int c = 0; c < 3; c++) {
            const float rc = raw_clip[c];
            lvl[c] = rc == kF32Max ? kF32Max : rc * mul.m[c] * 0.97f;
            lvl8[c] = lvl[c] * 0.8f;
        }
        lmin = fminf(fminf(fminf(lvl[0], lvl[1]), lvl[2]), 4.0f);
    }
    __device__ __forceinline__ bool clipped(float v, int c, bool cfa_clip) const { return v >= lvl[c] || (cfa_clip && v >= lvl8[c]); }
};

__global__ void k_hl_flag_clear(int* flag) {
    if (threadIdx.x == 0) flag[0] = 0;
}

// flag := 1 when any clip value is positive.  Every writer stores the same
// value, so plain stores suffice.
__global__ __launch_bounds__(256) void k_hl_any(const float* __restrict__ clip, size_t n, int* flag) {
    bool hit = false;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) hit |= clip[i] > 0.0f;
    if (hit) flag[0] = 1;
}

// demosaic.rs:448-459: bright, unclipped pixels away from clipped samples
// vote for the local R/G and B/G ratios with weight k; the rest vote 0.
__global__ __launch_bounds__(256) void k_hl_weights(const float* __restrict__ rgb, const float* __restrict__ clip,
                                                    const float* __restrict__ near, size_t n, Mul3 mul, const float* __restrict__ raw_clip,
                                                    const int* __restrict__ flag, float* __restrict__ wgt, float* __restrict__ rr,
                                                    float* __restrict__ br) {
    if (flag[0] == 0) return;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    Levels L;
    L.load(raw_clip, mul);
    const float p0 = rgb[3 * i], p1 = rgb[3 * i + 1], p2 = rgb[3 * i + 2];
    const bool cc = clip[i] > 0.0f;
    float k = 0.0f, r = 0.0f, b = 0.0f;
    if (!(near[i] > 0.0f || L.clipped(p0, 0, cc) || L.clipped(p1, 1, cc) || L.clipped(p2, 2, cc) || p1 <= 1e-4f)) {
        const float y = (p0 + p1 + p2) / 3.0f / L.lmin;
        k = fminf(y * y, 1.0f);
        r = clampr(p0 / p1, 0.1f, 10.0f) * k;
        b = clampr(p2 / p1, 0.1f, 10.0f) * k;
    }
    wgt[i] = k;
    rr[i] = r;
    br[i] = b;
}

// demosaic.rs:463-501: clipped channels rebuilt from the blurred ratios.
__global__ __launch_bounds__(256) void k_hl_apply(float* __restrict__ rgb, const float* __restrict__ clip, const float* __restrict__ wb,
                                                  const float* __restrict__ rrb, const float* __restrict__ brb, size_t n, Mul3 mul,
                                                  const float* __restrict__ raw_clip, const int* __restrict__ flag) {
    if (flag[0] == 0) return;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    Levels L;
    L.load(raw_clip, mul);
    float p[3] = {rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]};
    const bool cc = clip[i] > 0.0f;
    const bool cl[3] = {L.clipped(p[0], 0, cc), L.clipped(p[1], 1, cc), L.clipped(p[2], 2, cc)};
    const int nc = (int)cl[0] + (int)cl[1] + (int)cl[2];
    if (nc == 0) return;
    if (nc == 3) {
        const float m = fmaxf(fmaxf(p[0], p[1]), p[2]);
        rgb[3 * i] = m;
        rgb[3 * i + 1] = m;
        rgb[3 * i + 2] = m;
        return;
    }
    const float wv = wb[i];
    float r = 1.0f, b = 1.0f;
    if (wv > 1e-4f) {
        r = rrb[i] / wv;
        b = brb[i] / wv;
    }
    // little nearby evidence -> towards neutral
    const float trust = fminf(wv * 3.0f, 1.0f);
    r = 1.0f + (r - 1.0f) * trust;
    b = 1.0f + (b - 1.0f) * trust;
    const float ratio[3] = {r, 1.0f, b};
    float g = p[1];
    for (int c = 0; c < 3; c++)
        if (!cl[c]) g = fmaxf(g, p[c] / ratio[c]);
    for (int c = 0; c < 3; c++)
        if (cl[c]) p[c] = fmaxf(p[c], g * ratio[c]);
    const float top = fmaxf(fmaxf(p[0], p[1]), p[2]);
    const float q = (float)nc / 3.0f;
    const float t = q * q;  // powi(2)
    for (int c = 0; c < 3; c++) p[c] += (top - p[c]) * t * 0.5f;
    rgb[3 * i] = p[0];
    rgb[3 * i + 1] = p[1];
    rgb[3 * i + 2] = p[2];
}

// r == 0 box pass of the gaussian: box_blur returns its input unchanged.
__global__ __launch_bounds__(256) void k_hl_copy3(CBoxPlanes in, BoxPlanes out, size_t n) {
    if (out.gate && out.gate[0] == 0) return;
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    for (int k = 0; k < 3; k++) out.p[k][i] = in.p[k][i];
}

unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

// One gated box_blur (image.rs:269-311) of three planes: in -> tmp -> out.
hipError_t box3(float* const in[3], float* const tmp[3], float* const out[3], size_t w, size_t h, size_t r, const int* flag,
                hipStream_t st) {
    CBoxPlanes ci{}, ct{};
    BoxPlanes t{}, o{};
    for (int k = 0; k < 3; k++) {
        ci.p[k] = in[k];
        t.p[k] = tmp[k];
        ct.p[k] = tmp[k];
        o.p[k] = out[k];
    }
    t.gate = flag;
    o.gate = flag;
    if (r == 0) {
        hipLaunchKernelGGL(k_hl_copy3, dim3(nblk(w * h)), dim3(256), 0, st, ci, o, w * h);
        return hipGetLastError();
    }
    hipError_t e = box_h_planes(ci, t, 3, w, h, r, st);
    if (e == hipSuccess) e = box_v_planes(ct, o, 3, w, h, r, st);
    return e;
}

}  // namespace

size_t highlights_ws_floats(size_t w, size_t h) { return HL_HDR + 9 * w * h; }

hipError_t launch_reconstruct_highlights(float* rgb, const float* clip, size_t w, size_t h, const float mul[3], const float* raw_clip,
                                         float* ws, hipStream_t st) {
    const size_t n = w * h;
    if (n == 0) return hipSuccess;
    Mul3 m{{mul[0], mul[1], mul[2]}};
    int* flag = reinterpret_cast<int*>(ws);
    float* P = ws + HL_HDR;
    // planes: S = wgt / rr / br, T = box scratch, A = blurred; near uses T[1] with T[0] as its scratch
    float* S[3] = {P, P + n, P + 2 * n};
    float* T[3] = {P + 3 * n, P + 4 * n, P + 5 * n};
    float* A[3] = {P + 6 * n, P + 7 * n, P + 8 * n};
    hipLaunchKernelGGL(k_hl_flag_clear, dim3(1), dim3(64), 0, st, flag);
    hipLaunchKernelGGL(k_hl_any, dim3((unsigned)std::min<size_t>(nblk(n), 2048)), dim3(256), 0, st, clip, n, flag);
    // near_clip = box_blur(clip, 3) (demosaic.rs:447)
    {
        CBoxPlanes ci{}, ct{};
        BoxPlanes t{}, o{};
        ci.p[0] = clip;
        t.p[0] = T[0];
        ct.p[0] = T[0];
        o.p[0] = T[1];
        t.gate = flag;
        o.gate = flag;
        hipError_t e = box_h_planes(ci, t, 1, w, h, 3, st);
        if (e == hipSuccess) e = box_v_planes(ct, o, 1, w, h, 3, st);
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_hl_weights, dim3(nblk(n)), dim3(256), 0, st, rgb, clip, T[1], n, m, raw_clip, flag, S[0], S[1], S[2]);
    // gaussian_blur at sigma = max(w, h)/120 >= 2 is always the three-box
    // path (image.rs:327-337): S -> A -> S -> A
    const float sigma = fmaxf((float)(w > h ? w : h) / 120.0f, 2.0f);
    size_t radii[3];
    box_radii_for_gauss(sigma, radii);
    hipError_t e = box3(S, T, A, w, h, radii[0], flag, st);
    if (e == hipSuccess) e = box3(A, T, S, w, h, radii[1], flag, st);
    if (e == hipSuccess) e = box3(S, T, A, w, h, radii[2], flag, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_hl_apply, dim3(nblk(n)), dim3(256), 0, st, rgb, clip, A[0], A[1], A[2], n, m, raw_clip, flag);
    return hipGetLastError();
}

}  // namespace ds
