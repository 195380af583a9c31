// Non-local means (enhance.rs:34-72) for gfx950.
//
// The reference computes, for each search offset (dx, dy), the squared
// difference plane, box-blurs it (the patch distance) and accumulates
// exp(-dist/h^2)-weighted samples.  Here one workgroup owns a TW x TH output
// tile and loops over all offsets with the running (acc, wsum) of its pixels
// in registers.  The input tile plus an (S + P)-pixel halo sits in LDS.  Per
// offset:
//   phase 1: every thread takes 8-column row segments of the tile rows
//            [-P, TH+P), forms the squared differences and their horizontal
//            (2P+1)-sums with a sliding window, and writes them to LDS;
//   phase 2: every thread slides a vertical (2P+1)-window down one column
//            segment, turns the window sum into the weight with one exp2, and
//            accumulates.
// The horizontal-sum buffer is double-buffered, so one barrier per offset
// separates the phases.
//
// Border semantics are the reference's two-level clamp: the difference plane
// is defined on in-image pixels with the shifted sample clamped
// (enhance.rs:49-53), and the box blur clamps its window to the image
// (image.rs:279, :298).  Tiles whose halo stays inside the image skip the
// clamps.

#include "common.h"
#include "nlm.h"

// The NLM weights are not bit-reproducible against the CPU anyway (exp2 vs
// expf, window-sum order), so products may fuse into FMAs here.
#pragma clang fp contract(fast)

namespace ds {

namespace {

constexpr int NLM_TW = 64;
constexpr int NLM_TH = 64;
constexpr int NLM_NT = 256;
constexpr int NLM_SEG = 8;        // columns per phase-1 task
constexpr int NLM_ROWS = NLM_TH / (NLM_NT / NLM_TW);  // rows per phase-2 thread

template <int P>
__global__ __launch_bounds__(NLM_NT) void k_nlm_tile(const float* __restrict__ in, float* __restrict__ out, int w, int h, int S,
                                                     NlmWeight wp) {
    const float wscale = wp.scale();
    extern __shared__ float lds[];
    constexpr int TW = NLM_TW, TH = NLM_TH, NT = NLM_NT;
    constexpr int HR = TH + 2 * P;  // rows of horizontal sums
    const int SP = S + P;
    const int PW = TW + 2 * SP;  // input tile row length
    float* pt = lds;                          // input tile, (TH + 2SP) x PW
    float* hs0 = lds + (TH + 2 * SP) * PW;    // horizontal sums, 2 x HR x TW
    const int tid = threadIdx.x;
    const int ox = blockIdx.x * TW, oy = blockIdx.y * TH;

    // input tile with clamped coordinates: slot (u, v) holds p[clamp(ox+u), clamp(oy+v)]
    for (int i = tid; i < (TH + 2 * SP) * PW; i += NT) {
        int u = i % PW - SP, v = i / PW - SP;
        int gx = clampi(ox + u, 0, w - 1), gy = clampi(oy + v, 0, h - 1);
        pt[i] = in[(size_t)gy * w + gx];
    }
    auto P_at = [&](int u, int v) -> float { return pt[(v + SP) * PW + (u + SP)]; };
    const bool interior = ox - SP >= 0 && oy - SP >= 0 && ox + TW + SP <= w && oy + TH + SP <= h;

    const int col = tid % TW;
    const int row0 = (tid / TW) * NLM_ROWS;
    float acc[NLM_ROWS], wsum[NLM_ROWS];
#pragma unroll
    for (int k = 0; k < NLM_ROWS; k++) {
        acc[k] = 0.0f;
        wsum[k] = 0.0f;
    }
    __syncthreads();

    const int side = 2 * S + 1;
    const int noff = side * side;
    for (int o = 0; o < noff; o++) {
        const int dy = o / side - S, dx = o % side - S;
        float* hs = hs0 + (o & 1) * (HR * TW);
        // ---- phase 1: horizontal window sums of the squared differences
        for (int t = tid; t < HR * (TW / NLM_SEG); t += NT) {
            const int v = t / (TW / NLM_SEG) - P;
            const int u0 = (t % (TW / NLM_SEG)) * NLM_SEG;
            float d[NLM_SEG + 2 * P];
            if (interior) {
#pragma unroll
                for (int i = 0; i < NLM_SEG + 2 * P; i++) {
                    float e = P_at(u0 - P + i, v) - P_at(u0 - P + i + dx, v + dy);
                    d[i] = e * e;
                }
            } else {
                const int gv = clampi(oy + v, 0, h - 1);
                const int sv = clampi(gv + dy, 0, h - 1) - oy;
#pragma unroll
                for (int i = 0; i < NLM_SEG + 2 * P; i++) {
                    const int gu = clampi(ox + u0 - P + i, 0, w - 1);
                    const int su = clampi(gu + dx, 0, w - 1) - ox;
                    float e = P_at(gu - ox, gv - oy) - P_at(su, sv);
                    d[i] = e * e;
                }
            }
            float s = 0.0f;
#pragma unroll
            for (int i = 0; i < 2 * P + 1; i++) s += d[i];
            float* dst = hs + (v + P) * TW + u0;
#pragma unroll
            for (int j = 0; j < NLM_SEG; j++) {
                dst[j] = s;
                if (j + 1 < NLM_SEG) s += d[j + 2 * P + 1] - d[j];
            }
        }
        __syncthreads();
        // ---- phase 2: vertical window sums, weights, accumulation
        {
            const float* hc = hs + row0 * TW + col;  // hs row (row0 - P) + P
            float s = 0.0f;
#pragma unroll
            for (int j = 0; j < 2 * P + 1; j++) s += hc[j * TW];
#pragma unroll
            for (int k = 0; k < NLM_ROWS; k++) {
                const float wt = __builtin_amdgcn_exp2f(s * wscale);
                const float q = P_at(col + dx, row0 + k + dy);
                acc[k] += wt * q;
                wsum[k] += wt;
                if (k + 1 < NLM_ROWS) s += hc[(k + 2 * P + 1) * TW] - hc[k * TW];
            }
        }
        // next offset writes the other hs buffer; the barrier after its
        // phase 1 orders it against this phase 2
    }
#pragma unroll
    for (int k = 0; k < NLM_ROWS; k++) {
        const int gx = ox + col, gy = oy + row0 + k;
        if (gx < w && gy < h) out[(size_t)gy * w + gx] = acc[k] / fmaxf(wsum[k], 1e-12f);
    }
}

// Any search / patch radius: direct per-pixel patch distances with the same
// two-level clamp.  Used when the tile kernel's halo or patch limit is
// exceeded.
__global__ void k_nlm_direct(const float* __restrict__ in, float* __restrict__ out, int w, int h, int S, int P, NlmWeight wp,
                             float norm) {
    const float inv_h2 = 1.0f / wp.h2();
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w || y >= h) return;
    auto at = [&](int u, int v) { return in[(size_t)clampi(v, 0, h - 1) * w + clampi(u, 0, w - 1)]; };
    float acc = 0.0f, wsum = 0.0f;
    for (int dy = -S; dy <= S; dy++)
        for (int dx = -S; dx <= S; dx++) {
            float dist = 0.0f;
            for (int j = -P; j <= P; j++) {
                const int cy = clampi(y + j, 0, h - 1);
                float row = 0.0f;
                for (int i = -P; i <= P; i++) {
                    const int cx = clampi(x + i, 0, w - 1);
                    float e = at(cx, cy) - at(cx + dx, cy + dy);
                    row += e * e;
                }
                dist += row * norm;
            }
            dist *= norm;
            float wt = expf(-(dist * inv_h2));
            acc += wt * at(x + dx, y + dy);
            wsum += wt;
        }
    out[(size_t)y * w + x] = acc / fmaxf(wsum, 1e-12f);
}

constexpr int NLM_MAX_SP = 24;  // halo limit of the tile kernel

template <int P>
hipError_t launch_tile(const float* in, float* out, int w, int h, int S, NlmWeight wp, hipStream_t stream) {
    const int SP = S + P;
    size_t lds = (size_t)((NLM_TH + 2 * SP) * (NLM_TW + 2 * SP) + 2 * (NLM_TH + 2 * P) * NLM_TW) * sizeof(float);
    dim3 grid((unsigned)((w + NLM_TW - 1) / NLM_TW), (unsigned)((h + NLM_TH - 1) / NLM_TH));
    static bool attr_set = false;
    if (!attr_set) {
        hipError_t e = hipFuncSetAttribute((const void*)k_nlm_tile<P>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)nlm_tile_lds_bytes(NLM_MAX_SP, P));
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    hipLaunchKernelGGL(k_nlm_tile<P>, grid, dim3(NLM_NT), lds, stream, in, out, w, h, S, wp);
    return hipGetLastError();
}

}  // namespace

size_t nlm_tile_lds_bytes(int S, int P) {
    const int SP = S + P;
    return (size_t)((NLM_TH + 2 * SP) * (NLM_TW + 2 * SP) + 2 * (NLM_TH + 2 * P) * NLM_TW) * sizeof(float);
}

hipError_t launch_nlm(const float* in, float* out, size_t w, size_t h, NlmH hp, long search, size_t patch, hipStream_t stream) {
    if (w == 0 || h == 0) return hipSuccess;
    const int S = (int)search;
    const int P = (int)patch;
    const double norm = 1.0 / (double)(2 * P + 1);
    NlmWeight wp;
    wp.h_param = hp.h_param;
    wp.sigma_dev = hp.sigma_dev;
    wp.sigma_mul = hp.sigma_mul;
    // weight = exp(-(norm^2 * window_sum) / h2) = exp2(window_sum * scale)
    wp.k = (float)(norm * norm * 1.4426950408889634);
    if (S + P <= NLM_MAX_SP) {
        switch (P) {
            case 0: return launch_tile<0>(in, out, (int)w, (int)h, S, wp, stream);
            case 1: return launch_tile<1>(in, out, (int)w, (int)h, S, wp, stream);
            case 2: return launch_tile<2>(in, out, (int)w, (int)h, S, wp, stream);
            case 3: return launch_tile<3>(in, out, (int)w, (int)h, S, wp, stream);
            case 4: return launch_tile<4>(in, out, (int)w, (int)h, S, wp, stream);
            case 5: return launch_tile<5>(in, out, (int)w, (int)h, S, wp, stream);
            default: break;
        }
    }
    dim3 grid((unsigned)((w + 255) / 256), (unsigned)h);
    hipLaunchKernelGGL(k_nlm_direct, grid, dim3(256), 0, stream, in, out, (int)w, (int)h, S, P, wp, (float)norm);
    return hipGetLastError();
}

}  // namespace ds
