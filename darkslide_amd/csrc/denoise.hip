// Denoise composite (enhance.rs:74-92) for gfx950: Oklab split, noise
// estimate (enhance.rs:9-32), box filters (image.rs:268-313) and the two
// guided filters (image.rs:384-405) around the NLM of the L channel.
//
// The guided filters' box blurs use the exact running-sum passes of box.hip
// (their variance term cancels too much precision for any other summation
// order to stay within the error budget).  The first horizontal pass computes
// its six inputs (I, a, b, I a, I b, I I) on the fly; the coefficients and the
// final a*I + b with the inverse Oklab are pointwise kernels between passes.

#include "box.h"
#include "common.h"
#include "denoise.h"
#include "nlm.h"

#pragma clang fp contract(off)

namespace ds {

namespace {

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

// Exact element of rank `target` (select_nth_unstable at len / 2, the valid
// count coming from the sampling geometry) by a 4 x 8-bit MSB-first radix
// select in one workgroup, then sigma = v / 0.6745 / sqrt(20).  Border
// markers (+inf) sort after every finite sample, so they never reach the
// selected rank.  Each wave counts into its own histogram (the top bytes of
// the samples cluster in a few bins, which a shared histogram serialises),
// and the bucket holding the rank is found with a wave prefix scan.
constexpr int SEL_NT = 1024, SEL_WAVES = SEL_NT / 64;
__global__ void k_sigma_zero(float* __restrict__ sigma_out) {
    if (threadIdx.x == 0) *sigma_out = 0.0f;
}
__global__ __launch_bounds__(SEL_NT) void k_noise_select(const uint32_t* __restrict__ vals, size_t count, uint32_t target0, float sqrt20,
                                                         float* __restrict__ sigma_out) {
    __shared__ uint32_t hist[SEL_WAVES][256];
    __shared__ uint32_t s_wsum[4];
    __shared__ uint32_t s_prefix, s_target;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) {
        s_prefix = 0;
        s_target = target0;
    }
    uint32_t mask = 0;
    for (int shift = 24; shift >= 0; shift -= 8) {
        for (int i = tid; i < SEL_WAVES * 256; i += SEL_NT) (&hist[0][0])[i] = 0;
        __syncthreads();
        // read before the barriers below: one thread rewrites both at the end
        const uint32_t prefix = s_prefix, target = s_target;
        // eight loads in flight per thread before their atomics
        constexpr int SEL_B = 8;
        for (size_t base = 0; base < count; base += (size_t)SEL_B * SEL_NT) {
            uint32_t v[SEL_B];
#pragma unroll
            for (int q = 0; q < SEL_B; q++) {
                const size_t k = base + (size_t)q * SEL_NT + tid;
                v[q] = k < count ? vals[k] : ~0u;
            }
#pragma unroll
            for (int q = 0; q < SEL_B; q++)
                if (base + (size_t)q * SEL_NT + tid < count && (v[q] & mask) == prefix) atomicAdd(&hist[wave][(v[q] >> shift) & 255u], 1u);
        }
        __syncthreads();
        // bins 0..255 on threads 0..255 (waves 0..3): totals, inclusive scan
        // within each wave, then across the four waves
        uint32_t tot = 0, incl = 0;
        if (tid < 256) {
#pragma unroll
            for (int w = 0; w < SEL_WAVES; w++) tot += hist[w][tid];
            incl = tot;
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const uint32_t u = __shfl_up(incl, off, 64);
                if (lane >= off) incl += u;
            }
            if (lane == 63) s_wsum[wave] = incl;
        }
        __syncthreads();
        if (tid < 256) {
            uint32_t before = 0;
            for (int w = 0; w < wave; w++) before += s_wsum[w];
            const uint32_t excl = before + incl - tot;
            // the one bin whose cumulative range holds the rank
            if (excl <= target && target < excl + tot) {
                s_prefix = prefix | ((uint32_t)tid << shift);
                s_target = target - excl;
            }
        }
        mask |= 255u << shift;
        __syncthreads();
    }
    if (tid == 0) *sigma_out = __uint_as_float(s_prefix) / 0.6745f / sqrt20;
}

// The guide I = L + (lf - L) m (enhance.rs:86) and the products the guided
// filters box (image.rs:388-389): I, I a, I b, I I.
__global__ void k_gf_inputs(const float* __restrict__ L, const float* __restrict__ LF, const float* __restrict__ A,
                            const float* __restrict__ B, float m, size_t n, float* __restrict__ I_, float* __restrict__ IA,
                            float* __restrict__ IB, float* __restrict__ II) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float o = L[i];
    const float I = o + (LF[i] - o) * m;
    I_[i] = I;
    IA[i] = I * A[i];
    IB[i] = I * B[i];
    II[i] = I * I;
}

// image.rs:395-400 for the a and b planes, eps = (2 max(sigma,1e-4))^2
// (enhance.rs:82, :88)
__global__ void k_gf_coef(const float* __restrict__ mi_, const float* __restrict__ mpa_, const float* __restrict__ mpb_,
                          const float* __restrict__ mipa_, const float* __restrict__ mipb_, const float* __restrict__ mii_, size_t n,
                          const float* __restrict__ sigma_dev, float* __restrict__ a_a, float* __restrict__ b_a, float* __restrict__ a_b,
                          float* __restrict__ b_b) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float s2 = fmaxr(*sigma_dev, 1e-4f) * 2.0f;
    const float eps = s2 * s2;
    const float mi = mi_[i], mpa = mpa_[i], mpb = mpb_[i];
    const float var = mii_[i] - mi * mi;
    const float cov_a = mipa_[i] - mi * mpa;
    const float aa = cov_a / (var + eps);
    a_a[i] = aa;
    b_a[i] = mpa - aa * mi;
    const float cov_b = mipb_[i] - mi * mpb;
    const float ab = cov_b / (var + eps);
    a_b[i] = ab;
    b_b[i] = mpb - ab * mi;
}

// image.rs:404 a*I + b for both planes, then Oklab -> ProPhoto and max(0)
// (enhance.rs:91)
__global__ void k_gf_final(const float* __restrict__ L, const float* __restrict__ LF, float m, const float* __restrict__ ma_a,
                           const float* __restrict__ mb_a, const float* __restrict__ ma_b, const float* __restrict__ mb_b, size_t n, OklabM ok,
                           float* __restrict__ rgb) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float o = L[i];
    const float I = o + (LF[i] - o) * m;
    float lab[3] = {I, ma_a[i] * I + mb_a[i], ma_b[i] * I + mb_b[i]};
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

// Samples of enhance.rs:12-22 off the 1-pixel border: for each inner row y,
// the multiples of step in [y w + 1, y w + w - 2].
size_t noise_valid_samples(size_t w, size_t h, size_t step) {
    if (w < 3 || h < 3) return 0;
    size_t v = 0;
    for (size_t y = 1; y + 1 < h; y++) {
        const size_t lo = y * w + 1, hi = y * w + w - 2;
        v += hi / step - (lo + step - 1) / step + 1;
    }
    return v;
}

hipError_t launch_estimate_noise(const float* plane, size_t w, size_t h, uint32_t* scratch, float* sigma_out, hipStream_t st) {
    const size_t n = w * h;
    const size_t step = n / 200000 > 1 ? n / 200000 : 1;
    const size_t valid = noise_valid_samples(w, h, step);
    if (valid == 0) {  // enhance.rs:25-27
        hipLaunchKernelGGL(k_sigma_zero, dim3(1), dim3(64), 0, st, sigma_out);
        return hipGetLastError();
    }
    const size_t count = (n + step - 1) / step;
    hipLaunchKernelGGL(k_noise_sample, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, plane, w, h, step, count, scratch);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_noise_select, dim3(1), dim3(SEL_NT), 0, st, scratch, count, (uint32_t)(valid / 2), sqrtf(20.0f), sigma_out);
    return hipGetLastError();
}

// Guided filters of a and b by l2 = L + (lf - L) m, then Oklab^-1 into rgb.
// ws: 12 planes of w*h floats.
hipError_t launch_guided_pair(const float* L, const float* LF, const float* A, const float* B, float m, size_t r, const float* sigma_dev,
                              const OklabM& ok, float* ws, float* rgb, size_t w, size_t h, hipStream_t st) {
    const size_t n = w * h;
    if (n == 0) return hipSuccess;
    float* t[12];
    for (int k = 0; k < 12; k++) t[k] = ws + (size_t)k * n;
    auto planes = [&](int first, int count) {
        BoxPlanes o{};
        for (int k = 0; k < count; k++) o.p[k] = t[first + k];
        return o;
    };
    auto cplanes = [&](int first, int count) {
        CBoxPlanes o{};
        for (int k = 0; k < count; k++) o.p[k] = t[first + k];
        return o;
    };
    const unsigned nb = (unsigned)((n + 255) / 256);
    // the guide I = l2 and its products (image.rs:386-391) -> t0..t3
    hipLaunchKernelGGL(k_gf_inputs, dim3(nb), dim3(256), 0, st, L, LF, A, B, m, n, t[0], t[1], t[2], t[3]);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // means of I, a, b, I a, I b, I I -> t6..t11 (horizontal pass into t4, t5, t8..t11 first)
    CBoxPlanes hin{};
    hin.p[0] = t[0];
    hin.p[1] = A;
    hin.p[2] = B;
    hin.p[3] = t[1];
    hin.p[4] = t[2];
    hin.p[5] = t[3];
    BoxPlanes hout{};
    hout.p[0] = t[4];
    hout.p[1] = t[5];
    hout.p[2] = t[8];
    hout.p[3] = t[9];
    hout.p[4] = t[10];
    hout.p[5] = t[11];
    e = box_h_planes(hin, hout, 6, w, h, r, st);
    if (e != hipSuccess) return e;
    CBoxPlanes vin{};
    for (int k = 0; k < 6; k++) vin.p[k] = hout.p[k];
    BoxPlanes vout{};
    vout.p[0] = t[6];
    vout.p[1] = t[7];
    vout.p[2] = t[0];
    vout.p[3] = t[1];
    vout.p[4] = t[2];
    vout.p[5] = t[3];
    e = box_v_planes(vin, vout, 6, w, h, r, st);
    if (e != hipSuccess) return e;
    // means now: I t6, a t7, b t0, I a t1, I b t2, I I t3
    // coefficients (image.rs:395-400) -> t8..t11
    hipLaunchKernelGGL(k_gf_coef, dim3(nb), dim3(256), 0, st, t[6], t[7], t[0], t[1], t[2], t[3], n, sigma_dev, t[8], t[9], t[10], t[11]);
    e = hipGetLastError();
    // their means (image.rs:402-403) -> t8..t11 (via t4..t7)
    if (e == hipSuccess) e = box_h_planes(cplanes(8, 4), planes(4, 4), 4, w, h, r, st);
    if (e == hipSuccess) e = box_v_planes(cplanes(4, 4), planes(8, 4), 4, w, h, r, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_gf_final, dim3(nb), dim3(256), 0, st, L, LF, m, t[8], t[9], t[10], t[11], n, ok, rgb);
    return hipGetLastError();
}

}  // namespace ds
