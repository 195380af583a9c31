// Plane-level primitives of the detail path, each a stand-alone stage of
// the C ABI: gaussian_blur (image.rs:326-361), guided_filter
// (image.rs:384-405), Oklab to ProPhoto (color.rs:415-418) and luminance
// (image.rs:174-176).  All keep the reference's operation order (no FMA
// contraction) and are bit-reproducible against it.

#include "box.h"
#include "common.h"
#include "libm_ref.h"
#include "rl.h"
#include "stages.h"

#include <initializer_list>

#pragma clang fp contract(off)

namespace ds {

namespace {

struct Taps13 {
    float k[13];
};

// gaussian_blur_exact passes (image.rs:347-360): taps in order, clamped edges
__global__ void k_gauss_h(const float* __restrict__ in, float* __restrict__ out, int w, int h, int r, Taps13 taps) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const size_t row = (size_t)y * (size_t)w;
    float acc = 0.0f;
    for (int j = 0; j <= 2 * r; j++) acc += taps.k[j] * in[row + (size_t)clampi(x + j - r, 0, w - 1)];
    out[row + (size_t)x] = acc;
}

__global__ void k_gauss_v(const float* __restrict__ in, float* __restrict__ out, int w, int h, int r, Taps13 taps) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    float acc = 0.0f;
    for (int j = 0; j <= 2 * r; j++) acc += taps.k[j] * in[(size_t)clampi(y + j - r, 0, h - 1) * (size_t)w + (size_t)x];
    out[(size_t)y * (size_t)w + (size_t)x] = acc;
}

// image.rs:395-400 with a scalar eps
__global__ void k_gf1_coef(const float* __restrict__ mi_, const float* __restrict__ mp_, const float* __restrict__ mip_,
                           const float* __restrict__ mii_, size_t n, float eps, float* __restrict__ a, float* __restrict__ b) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float mi = mi_[i], mp = mp_[i];
    const float cov = mip_[i] - mi * mp;
    const float var = mii_[i] - mi * mi;
    const float av = cov / (var + eps);
    a[i] = av;
    b[i] = mp - av * mi;
}

// image.rs:388-389: i p (when ip is set) and i i
__global__ void k_gf1_products(const float* __restrict__ I, const float* __restrict__ P, size_t n, float* __restrict__ ip,
                               float* __restrict__ ii) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float v = I[i];
    if (ip) ip[i] = v * P[i];
    ii[i] = v * v;
}

// image.rs:404
__global__ void k_gf1_final(const float* __restrict__ I, const float* __restrict__ ma, const float* __restrict__ mb, size_t n,
                            float* __restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = ma[i] * I[i] + mb[i];
}

__global__ void k_oklab_inv(const float* __restrict__ lab, size_t n, OklabM ok, float* __restrict__ rgb) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v[3] = {lab[3 * i], lab[3 * i + 1], lab[3 * i + 2]};
    float l[3], cube[3], p[3];
    mul_mv(ok.m2_inv, v, l);
    cube[0] = l[0] * l[0] * l[0];
    cube[1] = l[1] * l[1] * l[1];
    cube[2] = l[2] * l[2] * l[2];
    mul_mv(ok.from_lms, cube, p);
    rgb[3 * i] = p[0];
    rgb[3 * i + 1] = p[1];
    rgb[3 * i + 2] = p[2];
}

__global__ void k_oklab_fwd_il(const float* __restrict__ rgb, size_t n, OklabM ok, float* __restrict__ lab) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float p[3] = {rgb[3 * i], rgb[3 * i + 1], rgb[3 * i + 2]};
    float lms[3], c[3], o[3];
    mul_mv(ok.to_lms, p, lms);
    c[0] = ds_cbrtf(lms[0]);
    c[1] = ds_cbrtf(lms[1]);
    c[2] = ds_cbrtf(lms[2]);
    mul_mv(ok.m2, c, o);
    lab[3 * i] = o[0];
    lab[3 * i + 1] = o[1];
    lab[3 * i + 2] = o[2];
}

__global__ void k_luminance(const float* __restrict__ rgb, size_t n, float* __restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = 0.2880402f * rgb[3 * i] + 0.7118741f * rgb[3 * i + 1] + 0.0000857f * rgb[3 * i + 2];
}

unsigned nblk(size_t n) { return (unsigned)((n + 255) / 256); }

__global__ void k_u64_zero(unsigned long long* p) { *p = 0; }

// 2^32 inputs: 2^18 blocks of 256 threads, 64 inputs per thread at a
// stride of 256; one atomic per wave.
constexpr int LH_PER = 64;
__global__ __launch_bounds__(256) void k_libm_hash(int fn, unsigned long long* __restrict__ acc) {
    const uint32_t base = blockIdx.x * (256u * LH_PER) + threadIdx.x;
    unsigned long long sum = 0;
    for (int k = 0; k < LH_PER; k++) {
        const uint32_t i = base + 256u * (uint32_t)k;
        float x;
        memcpy(&x, &i, 4);
        sum += libm_hash_term(i, fn == 0 ? ds_cbrtf(x) : ds_expf(x));
    }
    for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(acc, sum);
}

}  // namespace

// Wells' three-box radii (image.rs:363-373)
void box_radii_for_gauss(float sigma, size_t radii[3]) {
    const int n = 3;
    float w_ideal = sqrtf(12.0f * sigma * sigma / (float)n + 1.0f);
    int wl = (int)floorf(w_ideal);
    if (wl % 2 == 0) wl -= 1;
    int wu = wl + 2;
    float m_ideal = (12.0f * sigma * sigma - (float)(n * wl * wl) - 4.0f * (float)n * (float)wl - 3.0f * (float)n) / (-4.0f * (float)wl - 4.0f);
    int m = (int)roundf(m_ideal);
    for (int i = 0; i < n; i++) {
        int v = ((i < m ? wl : wu) - 1) / 2;
        radii[i] = (size_t)(v > 0 ? v : 0);
    }
}

hipError_t launch_gaussian_blur(const float* in, float* ws, float* out, size_t w, size_t h, float sigma, hipStream_t st) {
    const size_t n = w * h;
    if (n == 0) return hipSuccess;
    if (sigma < 0.3f) return hipMemcpyAsync(out, in, n * sizeof(float), hipMemcpyDeviceToDevice, st);
    if (sigma < 2.0f) {
        Taps13 t{};
        int r = gaussian_taps(sigma, t.k);
        dim3 grid((unsigned)((w + 255) / 256), (unsigned)h);
        hipLaunchKernelGGL(k_gauss_h, grid, dim3(256), 0, st, in, ws, (int)w, (int)h, r, t);
        hipLaunchKernelGGL(k_gauss_v, grid, dim3(256), 0, st, ws, out, (int)w, (int)h, r, t);
        return hipGetLastError();
    }
    // three boxes (image.rs:334-337): ws holds 2 planes
    size_t radii[3];
    box_radii_for_gauss(sigma, radii);
    float* a = ws;
    float* b = ws + n;
    // box_blur needs its scratch plane distinct from input and output
    hipError_t e = launch_box_blur(in, out, a, w, h, radii[0], st);
    if (e == hipSuccess) e = launch_box_blur(a, out, b, w, h, radii[1], st);
    if (e == hipSuccess) e = launch_box_blur(b, a, out, w, h, radii[2], st);
    return e;
}

hipError_t launch_guided_coeffs(const float* I, const float* P, size_t w, size_t h, size_t r, float eps, float* ws, hipStream_t st) {
    const size_t n = w * h;
    if (n == 0) return hipSuccess;
    float* t[8];
    for (int k = 0; k < 8; k++) t[k] = ws + (size_t)k * n;
    auto planes = [&](std::initializer_list<float*> l) {
        BoxPlanes o{};
        int k = 0;
        for (float* q : l) o.p[k++] = q;
        return o;
    };
    auto cplanes = [&](std::initializer_list<const float*> l) {
        CBoxPlanes o{};
        int k = 0;
        for (const float* q : l) o.p[k++] = q;
        return o;
    };
    // The products are formed once (image.rs:388-389) so every blur is a
    // pass over plain planes.  A self-guided filter (p == i) has
    // box(p) == box(i) and box(i p) == box(i i) bit for bit, so those two
    // planes are blurred once and read twice.
    const float *mi, *mp, *mip, *mii;
    hipError_t e;
    if (I == P) {
        hipLaunchKernelGGL(k_gf1_products, dim3(nblk(n)), dim3(256), 0, st, I, I, n, (float*)nullptr, t[2]);
        e = box_h_planes(cplanes({I, t[2]}), planes({t[0], t[1]}), 2, w, h, r, st);
        if (e == hipSuccess) e = box_v_planes(cplanes({t[0], t[1]}), planes({t[2], t[3]}), 2, w, h, r, st);
        mi = mp = t[2];
        mip = mii = t[3];
    } else {
        hipLaunchKernelGGL(k_gf1_products, dim3(nblk(n)), dim3(256), 0, st, I, P, n, t[2], t[3]);
        e = box_h_planes(cplanes({I, P, t[2], t[3]}), planes({t[0], t[1], t[4], t[5]}), 4, w, h, r, st);
        if (e == hipSuccess) e = box_v_planes(cplanes({t[0], t[1], t[4], t[5]}), planes({t[2], t[3], t[6], t[7]}), 4, w, h, r, st);
        mi = t[2];
        mp = t[3];
        mip = t[6];
        mii = t[7];
    }
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_gf1_coef, dim3(nblk(n)), dim3(256), 0, st, mi, mp, mip, mii, n, eps, t[0], t[1]);
    e = hipGetLastError();
    // the coefficients' means (image.rs:402-403) -> planes 4 and 5
    if (e == hipSuccess) e = box_h_planes(cplanes({t[0], t[1]}), planes({t[2], t[3]}), 2, w, h, r, st);
    if (e == hipSuccess) e = box_v_planes(cplanes({t[2], t[3]}), planes({t[4], t[5]}), 2, w, h, r, st);
    return e;
}

hipError_t launch_guided_filter(const float* I, const float* P, size_t w, size_t h, size_t r, float eps, float* ws, float* out,
                                hipStream_t st) {
    const size_t n = w * h;
    if (n == 0) return hipSuccess;
    hipError_t e = launch_guided_coeffs(I, P, w, h, r, eps, ws, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_gf1_final, dim3(nblk(n)), dim3(256), 0, st, I, ws + 4 * n, ws + 5 * n, n, out);
    return hipGetLastError();
}

hipError_t launch_oklab_fwd_interleaved(const float* rgb, size_t n, const OklabM& ok, float* lab, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_oklab_fwd_il, dim3(nblk(n)), dim3(256), 0, st, rgb, n, ok, lab);
    return hipGetLastError();
}

hipError_t launch_oklab_inv(const float* lab, size_t n, const OklabM& ok, float* rgb, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_oklab_inv, dim3(nblk(n)), dim3(256), 0, st, lab, n, ok, rgb);
    return hipGetLastError();
}

hipError_t launch_luminance(const float* rgb, size_t n, float* out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_luminance, dim3(nblk(n)), dim3(256), 0, st, rgb, n, out);
    return hipGetLastError();
}

hipError_t launch_libm_hash(int fn, unsigned long long* acc, hipStream_t st) {
    hipLaunchKernelGGL(k_u64_zero, dim3(1), dim3(1), 0, st, acc);
    hipLaunchKernelGGL(k_libm_hash, dim3((unsigned)((1ull << 32) / (256 * LH_PER))), dim3(256), 0, st, fn, acc);
    return hipGetLastError();
}

}  // namespace ds
