// Denoise composite (enhance.rs:74-92) for gfx950: Oklab split, noise
// estimate (enhance.rs:9-32), box filters (image.rs:268-313) and the two
// guided filters (image.rs:384-405) around the NLM of the L channel.
//
// The guided filters' box blurs use the exact running-sum passes of box.hip
// (their variance term cancels too much precision for any other summation
// order to stay within the error budget).  Pointwise kernels form the guide
// and its products (I, I a, I b, I I), the coefficients, and the final
// a*I + b with the inverse Oklab between the passes.

#include "box.h"
#include "common.h"
#include "libm_ref.h"
#include "denoise.h"
#include "guided.h"
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
    c[0] = ds_cbrtf(lms[0]);
    c[1] = ds_cbrtf(lms[1]);
    c[2] = ds_cbrtf(lms[2]);
    mul_mv(ok.m2, c, lab);
    L[i] = lab[0];
    A[i] = lab[1];
    B[i] = lab[2];
}

// ---- noise estimate: every step-th pixel off the 1-px border, |laplacian|
// as ordered u32 bits; non-samples get +inf and sort last.
//
// The element of rank len/2 (select_nth_unstable, enhance.rs:23) is found by
// an exact MSB-first radix select over the bits in three levels of 12, 10
// and 10 bits, each level spread over the whole grid where it touches every
// sample:
//   k_noise_sample  writes the samples and counts level 1 (sign, exponent,
//                   3 mantissa bits) into a global histogram;
//   k_noise_l2      ranks level 1 (every block, redundantly), counts level 2
//                   of the samples in the chosen bin and appends those
//                   samples to a list (one global atomic per block);
//   k_noise_final   ranks levels 1 and 2 and counts level 3 over the list,
//                   which holds only the samples of the chosen level-1 bin.
// Border markers (+inf) sort after every finite sample, and the rank is
// below the number of finite ones, so they never reach it.
constexpr uint32_t kInfBits = 0x7f800000u;
constexpr int NOISE_L1_SHIFT = 20, NOISE_L1_BINS = 1 << (32 - NOISE_L1_SHIFT);
constexpr int NOISE_L2_BINS = 1024;
constexpr int NOISE_CTRL = 64;  // [0]: list length
constexpr int NOISE_HDR = NOISE_L1_BINS + NOISE_L2_BINS + NOISE_CTRL;
static_assert(NOISE_HDR % 4 == 0, "16-byte aligned list and samples");
static_assert((size_t)NOISE_HDR == NOISE_HDR_U32, "scratch header size published in denoise.h");
constexpr int NS_NT = 256, NS_PER = 4;  // sample / level-2 block: 1024 samples

// One histogram increment per active lane of the wave; the lanes sharing
// the first active lane's bin (a hot bin, most of the time) add their count
// with one atomic instead of serialising on its address.
__device__ inline void wave_count(uint32_t* hist, uint32_t bin, bool active) {
    const int lane = threadIdx.x & 63;
    const unsigned long long act = __ballot(active);
    if (!act) return;
    const int leader = __ffsll(act) - 1;
    const uint32_t lb = __shfl(bin, leader, 64);
    const bool same = active && bin == lb;
    const unsigned long long sm = __ballot(same);
    if (same) {
        if (lane == leader) atomicAdd(&hist[lb], (uint32_t)__popcll(sm));
    } else if (active) {
        atomicAdd(&hist[bin], 1u);
    }
}

// Appends the active lanes' values to buf (capacity cap) in lane order.
__device__ inline void wave_append(uint32_t* buf, uint32_t* n, uint32_t cap, uint32_t v, bool active) {
    const int lane = threadIdx.x & 63;
    const unsigned long long act = __ballot(active);
    if (!act) return;
    const int leader = __ffsll(act) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(n, (uint32_t)__popcll(act));
    base = __shfl(base, leader, 64);
    const uint32_t at = base + (uint32_t)__popcll(act & ((1ull << lane) - 1ull));
    if (active && at < cap) buf[at] = v;
}

// Exclusive prefix of v over the workgroup (thread order).  Ends with the
// totals of every wave published; the caller syncs before s_wsum is reused.
__device__ inline uint32_t sel_excl_scan(uint32_t v, uint32_t* s_wsum) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t incl = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t u = __shfl_up(incl, off, 64);
        if (lane >= off) incl += u;
    }
    if (lane == 63) s_wsum[wave] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (int i = 0; i < wave; i++) before += s_wsum[i];
    return before + incl - v;
}

// The bin of a NT * PER-bin histogram holding rank `target`, and the rank
// within it: s_sel = {bin, rank in bin}.  Thread t reads bins PER t ..
// PER t + PER - 1.  Ends with a barrier.
template <int NT, int PER>
__device__ inline void rank_bin(const uint32_t* bins, uint32_t target, uint32_t* s_wsum, uint32_t* s_sel) {
    const int tid = threadIdx.x;
    uint32_t c[PER];
    uint32_t tot = 0;
#pragma unroll
    for (int j = 0; j < PER; j++) {
        c[j] = bins[PER * tid + j];
        tot += c[j];
    }
    const uint32_t excl = sel_excl_scan(tot, s_wsum);
    if (excl <= target && target < excl + tot) {
        uint32_t t = target - excl;
        int j = 0;
        while (t >= c[j]) t -= c[j++];
        s_sel[0] = (uint32_t)(PER * tid + j);
        s_sel[1] = t;
    }
    __syncthreads();
}

__global__ __launch_bounds__(NS_NT) void k_noise_hist_clear(uint32_t* __restrict__ hdr) {
    const int i = blockIdx.x * NS_NT + threadIdx.x;
    if (i < NOISE_HDR) hdr[i] = 0;
}

__global__ __launch_bounds__(NS_NT) void k_noise_sample(const float* __restrict__ p, size_t w, size_t h, size_t step, size_t count,
                                                        uint32_t* __restrict__ vals, uint32_t* __restrict__ hist1) {
    __shared__ uint32_t lh[NOISE_L1_BINS];
    for (int i = threadIdx.x; i < NOISE_L1_BINS; i += NS_NT) lh[i] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * (NS_NT * NS_PER);
#pragma unroll
    for (int q = 0; q < NS_PER; q++) {
        const size_t k = base + (size_t)q * NS_NT + threadIdx.x;
        const bool in = k < count;
        uint32_t bits = kInfBits;
        if (in) {
            const size_t i = k * step;
            const size_t x = i % w, y = i / w;
            if (x >= 1 && y >= 1 && x + 1 < w && y + 1 < h) {
                const float c = p[i];
                const float lap = 4.0f * c - p[i - 1] - p[i + 1] - p[i - w] - p[i + w];
                bits = __float_as_uint(fabsf(lap));
            }
            vals[k] = bits;
        }
        wave_count(lh, bits >> NOISE_L1_SHIFT, in);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < NOISE_L1_BINS; i += NS_NT)
        if (lh[i]) atomicAdd(&hist1[i], lh[i]);
}

__global__ __launch_bounds__(NS_NT) void k_noise_l2(const uint32_t* __restrict__ vals, size_t count, uint32_t target0,
                                                    const uint32_t* __restrict__ hist1, uint32_t* __restrict__ hist2,
                                                    uint32_t* __restrict__ ctrl, uint32_t* __restrict__ list) {
    __shared__ uint32_t h2[NOISE_L2_BINS];
    __shared__ uint32_t keep[NS_NT * NS_PER];
    __shared__ uint32_t s_wsum[NS_NT / 64];
    __shared__ uint32_t s_sel[2], s_n, s_base;
    const int tid = threadIdx.x;
    for (int i = tid; i < NOISE_L2_BINS; i += NS_NT) h2[i] = 0;
    if (tid == 0) s_n = 0;
    // the block's samples are loaded before the level-1 rank is known
    const size_t base = (size_t)blockIdx.x * (NS_NT * NS_PER);
    uint32_t v[NS_PER];
#pragma unroll
    for (int q = 0; q < NS_PER; q++) {
        const size_t k = base + (size_t)q * NS_NT + tid;
        v[q] = k < count ? vals[k] : 0u;
    }
    rank_bin<NS_NT, NOISE_L1_BINS / NS_NT>(hist1, target0, s_wsum, s_sel);
    const uint32_t b1 = s_sel[0];
#pragma unroll
    for (int q = 0; q < NS_PER; q++) {
        const bool m = base + (size_t)q * NS_NT + tid < count && (v[q] >> NOISE_L1_SHIFT) == b1;
        wave_count(h2, (v[q] >> 10) & 1023u, m);
        wave_append(keep, &s_n, NS_NT * NS_PER, v[q], m);
    }
    __syncthreads();
    const uint32_t n = s_n;
    if (tid == 0 && n) s_base = atomicAdd(&ctrl[0], n);
    __syncthreads();
    for (uint32_t i = tid; i < n; i += NS_NT) list[s_base + i] = keep[i];
    for (int i = tid; i < NOISE_L2_BINS; i += NS_NT)
        if (h2[i]) atomicAdd(&hist2[i], h2[i]);
}

constexpr int SEL_NT = 1024;
static_assert(NOISE_L1_BINS == 4 * SEL_NT && NOISE_L2_BINS == SEL_NT, "final ranks: four level-1 bins, one level-2 / level-3 bin per thread");

__global__ void k_sigma_zero(float* __restrict__ sigma_out) {
    if (threadIdx.x == 0) *sigma_out = 0.0f;
}

// sigma = v / 0.6745 / sqrt(20) of the sample of rank target0 (enhance.rs:23-29).
__global__ __launch_bounds__(SEL_NT) void k_noise_final(const uint32_t* __restrict__ hist1, const uint32_t* __restrict__ hist2,
                                                        const uint32_t* __restrict__ ctrl, const uint32_t* __restrict__ list,
                                                        uint32_t target0, float sqrt20, float* __restrict__ sigma_out) {
    __shared__ uint32_t h3[SEL_NT];
    __shared__ uint32_t s_wsum[SEL_NT / 64];
    __shared__ uint32_t s_sel[2];
    const int tid = threadIdx.x;
    h3[tid] = 0;
    rank_bin<SEL_NT, NOISE_L1_BINS / SEL_NT>(hist1, target0, s_wsum, s_sel);
    const uint32_t b1 = s_sel[0], t1 = s_sel[1];
    rank_bin<SEL_NT, 1>(hist2, t1, s_wsum, s_sel);
    const uint32_t p2 = (b1 << 10) | s_sel[0], t2 = s_sel[1];  // bits 31..10 of the answer
    // level 3 over the list: the samples of bin b1, in no particular order
    const uint32_t n = ctrl[0];
    constexpr int B = 8;  // loads in flight per thread
    for (uint32_t base = 0; base < n; base += (uint32_t)B * SEL_NT) {
        uint32_t v[B];
#pragma unroll
        for (int q = 0; q < B; q++) {
            const uint32_t i = base + (uint32_t)q * SEL_NT + tid;
            v[q] = i < n ? list[i] : 0u;
        }
#pragma unroll
        for (int q = 0; q < B; q++) wave_count(h3, v[q] & 1023u, base + (uint32_t)q * SEL_NT + tid < n && (v[q] >> 10) == p2);
    }
    __syncthreads();
    rank_bin<SEL_NT, 1>(h3, t2, s_wsum, s_sel);
    if (tid == 0) *sigma_out = __uint_as_float((p2 << 10) | s_sel[0]) / 0.6745f / sqrt20;
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
    float aa, ba, ab, bb;
    gf_coef_px(mi_[i], mpa_[i], mpb_[i], mipa_[i], mipb_[i], mii_[i], gf_eps(sigma_dev), aa, ba, ab, bb);
    a_a[i] = aa;
    b_a[i] = ba;
    a_b[i] = ab;
    b_b[i] = bb;
}

// image.rs:404 a*I + b for both planes, then Oklab -> ProPhoto and max(0)
// (enhance.rs:91)
__global__ void k_gf_final(const float* __restrict__ I_, const float* __restrict__ ma_a, const float* __restrict__ mb_a,
                           const float* __restrict__ ma_b, const float* __restrict__ mb_b, size_t n, OklabM ok, float* __restrict__ rgb) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    gf_final_px(I_[i], ma_a[i], mb_a[i], ma_b[i], mb_b[i], ok, rgb + 3 * i);
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

size_t noise_scratch_u32(size_t w, size_t h) { return NOISE_HDR + 2 * noise_sample_count(w, h); }

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
    // scratch: level-1 and level-2 histograms, control words, the list of
    // samples in the chosen level-1 bin, the samples
    uint32_t* hist1 = scratch;
    uint32_t* hist2 = hist1 + NOISE_L1_BINS;
    uint32_t* ctrl = hist2 + NOISE_L2_BINS;
    uint32_t* list = scratch + NOISE_HDR;
    uint32_t* vals = list + count;
    const uint32_t target = (uint32_t)(valid / 2);
    const size_t per = (size_t)NS_NT * NS_PER;
    const unsigned blocks = (unsigned)((count + per - 1) / per);
    hipLaunchKernelGGL(k_noise_hist_clear, dim3((NOISE_HDR + NS_NT - 1) / NS_NT), dim3(NS_NT), 0, st, scratch);
    hipLaunchKernelGGL(k_noise_sample, dim3(blocks), dim3(NS_NT), 0, st, plane, w, h, step, count, vals, hist1);
    hipLaunchKernelGGL(k_noise_l2, dim3(blocks), dim3(NS_NT), 0, st, vals, count, target, hist1, hist2, ctrl, list);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_noise_final, dim3(1), dim3(SEL_NT), 0, st, hist1, hist2, ctrl, list, target, sqrtf(20.0f), sigma_out);
    return hipGetLastError();
}

// Guided filters of a and b by l2 = L + (lf - L) m, then Oklab^-1 into rgb.
// ws: 12 planes of w*h floats.  LF is dead once the guide is formed and
// serves as a spare plane.
hipError_t launch_guided_pair(const float* L, float* LF, const float* A, const float* B, float m, size_t r, const float* sigma_dev,
                              const OklabM& ok, float* ws, float* rgb, size_t w, size_t h, hipStream_t st, bool guide_ready) {
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
    // the guide I = l2 and its products (image.rs:386-391) -> t0..t3; I
    // stays in t0 for the final step
    if (!guide_ready) hipLaunchKernelGGL(k_gf_inputs, dim3(nb), dim3(256), 0, st, L, LF, A, B, m, n, t[0], t[1], t[2], t[3]);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    // means of I, a, b, I a, I b, I I (horizontal pass into t4, t5, t8..t11 first)
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
    vout.p[2] = LF;
    vout.p[3] = t[1];
    vout.p[4] = t[2];
    vout.p[5] = t[3];
    e = box_v_planes(vin, vout, 6, w, h, r, st);
    if (e != hipSuccess) return e;
    // means now: I t6, a t7, b LF, I a t1, I b t2, I I t3
    // coefficients (image.rs:395-400) -> t8..t11
    hipLaunchKernelGGL(k_gf_coef, dim3(nb), dim3(256), 0, st, t[6], t[7], LF, t[1], t[2], t[3], n, sigma_dev, t[8], t[9], t[10], t[11]);
    e = hipGetLastError();
    // their means (image.rs:402-403) -> t8..t11 (via t4..t7)
    if (e == hipSuccess) e = box_h_planes(cplanes(8, 4), planes(4, 4), 4, w, h, r, st);
    if (e == hipSuccess) e = box_v_planes(cplanes(4, 4), planes(8, 4), 4, w, h, r, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_gf_final, dim3(nb), dim3(256), 0, st, t[0], t[8], t[9], t[10], t[11], n, ok, rgb);
    return hipGetLastError();
}

}  // namespace ds
