// Richardson-Lucy "raw details" (enhance.rs:94-119) for gfx950.
//
// For sigma in [0.3, 2) the PSF blur is gaussian_blur_exact (image.rs:341-361):
// separable, r = ceil(3 sigma) <= 6 taps per side, clamped edges, taps
// accumulated in order.  One launch runs one whole iteration
//     conv = G(est); ratio = y / max(conv, 1e-6); corr = G(ratio);
//     est = clamp(est * corr, 0, 64)
// on a TX x TY tile: est with a 2r halo and y with an r halo are staged in
// LDS, and the four separable passes run LDS to LDS, so an iteration moves
// est in, y in and est out through HBM once.  The arithmetic keeps the
// reference's order and is compiled without FMA contraction, so each
// iteration is bit-identical to the CPU one.
//
// Other sigmas fall back to plane-wide passes: sigma < 0.3 (blur is a copy)
// and sigma >= 2 (three box blurs, image.rs:334-337).

#include "box.h"
#include "common.h"
#include "rl.h"
#include "stages.h"

#pragma clang fp contract(off)

namespace ds {

namespace {

constexpr int RL_TX = 64;

// Start of every tap sum.  The reference starts from +0.0, and +0.0 + t
// equals t for every t other than -0.0; here every term is a positive tap
// times a value >= 0 (est and y stay >= 1e-6, the ratio > 0), so no term is
// -0.0 and starting from -0.0 gives the same bits.  -0.0 + t == t for all t,
// so the compiler drops that first add, one per output and pass.
constexpr float kSumStart = -0.0f;

struct Taps {
    float k[13];
};

// y outside [1e-6, 2^64] (infinite, or too large for div_in_range) seen by
// the first iteration of a call: the later iterations then keep the plain
// division.  y = max(luminance, 1e-6) is never NaN (fmaxf returns the
// non-NaN operand).
constexpr float kDivSafeMax = 0x1p64f;

// a / b, bit for bit the compiler's IEEE division (v_div_scale, v_rcp,
// two Newton steps on the reciprocal and two remainder corrections,
// v_div_fmas, v_div_fixup) with the scale and fixup steps left out, for
// a in [1e-6, 2^64] and b in [1e-6, 64.01].  There v_div_scale returns its
// operand unchanged with VCC clear: neither operand is 0 or denormal, the
// exponents differ by less than 96, 1/b and a/b are normal and a >= 2^-103.
// So v_div_fmas is a plain fused multiply-add, and v_div_fixup returns the
// quotient (no NaN, infinity or zero operand, no overflow or underflow, both
// signs positive).  8 instructions instead of 11.
__device__ __forceinline__ float div_in_range(float a, float b) {
    float r = __builtin_amdgcn_rcpf(b);
    const float e = __builtin_fmaf(-b, r, 1.0f);
    r = __builtin_fmaf(e, r, r);
    float q = a * r;
    float t = __builtin_fmaf(-b, q, a);
    q = __builtin_fmaf(t, r, q);
    t = __builtin_fmaf(-b, q, a);
    return __builtin_fmaf(t, r, q);
}

// luminance (image.rs:174-176, color.rs:196-198), y = max(lum, 1e-6); est = y.
// Clears the call's range flag, which the first iteration sets.
__global__ void k_rl_init(const float* __restrict__ rgb, size_t n, float* __restrict__ y, float* __restrict__ est, int* __restrict__ flag) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *flag = 0;
    if (i >= n) return;
    float l = 0.2880402f * rgb[3 * i] + 0.7118741f * rgb[3 * i + 1] + 0.0000857f * rgb[3 * i + 2];
    float v = fmaxr(l, 1e-6f);
    y[i] = v;
    est[i] = v;
}

// out = p * clamp(est / y, 0.6, 1.6) (enhance.rs:113-115)
__device__ __forceinline__ void apply_px(const float* __restrict__ rgb, float est, float y, size_t i, float* __restrict__ out) {
    float r = clampr(est / y, 0.6f, 1.6f);
    out[3 * i + 0] = rgb[3 * i + 0] * r;
    out[3 * i + 1] = rgb[3 * i + 1] * r;
    out[3 * i + 2] = rgb[3 * i + 2] * r;
}
__global__ void k_rl_apply(const float* __restrict__ rgb, const float* __restrict__ est, const float* __restrict__ y, size_t n,
                           float* __restrict__ out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    apply_px(rgb, est[i], y[i], i, out);
}

// Interior staging: the ROWS x COLS block of a plane at g (row stride w)
// into LDS rows of stride LS.  Rows go to waves, columns to lanes; the row
// base is uniform per wave (scalar arithmetic) and each lane adds its
// column, so a load costs one vector add instead of the clamped 2-D index of
// stage_tile.
template <int ROWS, int COLS, int LS, int NT>
__device__ __forceinline__ void stage_rows(const float* __restrict__ g, int w, float* __restrict__ lds) {
    constexpr int NW = NT / 64, RPW = (ROWS + NW - 1) / NW, CH = (COLS + 63) / 64;
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6), lane = (int)threadIdx.x & 63;
    float r[RPW][CH];
    const float* rp = g + (size_t)wave * (size_t)w;  // row wave + NW i, advanced by NW rows
    const size_t step = (size_t)NW * (size_t)w;
#pragma unroll
    for (int i = 0; i < RPW; i++) {
#pragma unroll
        for (int c = 0; c < CH; c++) {
            const int row = wave + NW * i, col = lane + 64 * c;
            if (row < ROWS && col < COLS) r[i][c] = rp[col];
        }
        rp += step;
    }
#pragma unroll
    for (int i = 0; i < RPW; i++)
#pragma unroll
        for (int c = 0; c < CH; c++) {
            const int row = wave + NW * i, col = lane + 64 * c;
            if (row < ROWS && col < COLS) lds[row * LS + col] = r[i][c];
        }
}

// MODE bit 0 (FIRST): the call's first iteration, with the plain division;
// it checks every y it reads and sets *flag if one is outside [1e-6, 2^64].
// Later iterations read *flag: clear, the interior tiles divide with
// div_in_range (est is then in [0, 64], so conv is in [0, 64.01]).
// MODE bit 1 (LAST): the call's last iteration, which applies its est to rgb
// as k_rl_apply does (into rgb_out) instead of storing est.
template <int R, int TY, int NT, int MODE>
__global__ __launch_bounds__(NT) void k_rl_iter(const float* __restrict__ est, const float* __restrict__ yp, float* __restrict__ est_out,
                                                   int w, int h, Taps taps, TileOrder order, int* __restrict__ flag,
                                                   const float* __restrict__ rgb, float* __restrict__ rgb_out) {
    constexpr bool FIRST = (MODE & 1) != 0, LAST = (MODE & 2) != 0;
    static_assert(TY % 8 == 0, "row blocks of 8");
    constexpr int TX = RL_TX, K = 2 * R + 1, B = 8;
    // est: [-2R, T+2R) both ways; y, ratio: [-R, T+R); tmp: cols [-R, TX+R), rows [-2R, TY+2R)
    constexpr int EW = TX + 4 * R, EH = TY + 4 * R;
    constexpr int YW = TX + 2 * R, YH = TY + 2 * R;
    constexpr int MW = TX + 2 * R, MH = TY + 4 * R;
    constexpr int QW = TX, QH = TY + 2 * R;
    // LDS row strides: odd, so that lanes walking down consecutive rows (the
    // horizontal passes) or along a row (the vertical passes) hit distinct banks
    constexpr int ES = EW | 1, YS = YW | 1, MS = MW | 1, QS = QW | 1;
    constexpr int TMP = (MH * MS > QH * QS) ? MH * MS : QH * QS;
    static_assert(YH * YS <= EH * ES, "ratio fits in the est area");
    // Two LDS planes, so that more blocks fit per CU (the iteration is bound
    // by how many tiles are in flight, not by arithmetic): y is read from
    // global memory where the ratio needs it, and the ratio overwrites est
    // once the first pass has read it (the est values of the final update are
    // kept in registers).
    __shared__ float s_est[EH * ES];
    __shared__ float s_tmp[TMP];
    float* const s_rat = s_est;
    const int tid = threadIdx.x;
    int tx, ty;
    tile_edge_first_xcd(blockIdx.x, order, tx, ty);
    const int ox = tx * TX, oy = ty * TY;
    // The taps are symmetric bit for bit (k[i] and k[2R - i] come from the
    // same expression), so taps i and 2R - i share one register, and a
    // product k[i] x[m] that two outputs of a register-blocked task need is
    // formed once.
    float kk[K];
#pragma unroll
    for (int i = 0; i < K; i++) kk[i] = taps.k[i <= R ? i : K - 1 - i];

    const bool interior = ox >= 2 * R && oy >= 2 * R && ox + TX + 2 * R <= w && oy + TY + 2 * R <= h;
    if (interior) {
        stage_rows<EH, EW, ES, NT>(est + (size_t)(oy - 2 * R) * w + (ox - 2 * R), w, s_est);
    } else {
        stage_tile<EH * EW, NT>(
            [&](int i) {
                int u = i % EW - 2 * R, v = i / EW - 2 * R;
                return est[(size_t)clampi(oy + v, 0, h - 1) * w + clampi(ox + u, 0, w - 1)];
            },
            [&](int i, float v) { s_est[(i / EW) * ES + i % EW] = v; });
    }
    __syncthreads();
    auto E = [&](int u, int v) -> float { return s_est[(v + 2 * R) * ES + (u + 2 * R)]; };
    // est at the outputs of this thread's final-update tasks (interior: task
    // t -> column t % TX, rows 8 (t / TX) ..; edge: pixels tid + k NT)
    constexpr int NV2 = (TX * (TY / B) + NT - 1) / NT, NE = (TX * TY + NT - 1) / NT;
    constexpr int NEV = NV2 * B > NE ? NV2 * B : NE;
    float ev[NEV];
    if (interior) {
#pragma unroll
        for (int q = 0; q < NV2; q++) {
            const int t = tid + q * NT;
#pragma unroll
            for (int j = 0; j < B; j++) ev[q * B + j] = t < TX * (TY / B) ? E(t % TX, (t / TX) * B + j) : 0.0f;
        }
    } else {
#pragma unroll
        for (int k = 0; k < NE; k++) {
            const int i = tid + k * NT;
            ev[k] = i < TX * TY ? E(i % TX, i / TX) : 0.0f;
        }
    }

    // est * corr is >= +0 or NaN: est is >= 0 (y >= 1e-6, then clamped), the
    // taps are positive and every ratio is > 0, so clamp(x, 0, 64) is
    // x > 64 ? 64 : x, with NaN kept as f32::clamp keeps it.
    auto clamp64 = [](float x) { return x > 64.0f ? 64.0f : x; };
    auto interior_passes = [&](auto fast_tag) {
        constexpr bool FAST = decltype(fast_tag)::value;
        bool y_ok = true;
        // Register-blocked passes: each task produces B consecutive outputs
        // from B + 2R staged samples.  Every output still accumulates its
        // taps in order 0..2R with separate multiply and add.  Horizontal
        // passes put consecutive lanes on consecutive rows.
        // horizontal blur of est: rows [-2R, TY+2R), cols [-R, TX+R)
        constexpr int H1R = (MW + B - 1) / B;
        for (int t = tid; t < MH * H1R; t += NT) {
            const int v = t % MH - 2 * R, u0 = (t / MH) * B - R;
            const float* src = s_est + (v + 2 * R) * ES + (u0 - R + 2 * R);
            float x[B + 2 * R];
#pragma unroll
            for (int i = 0; i < B + 2 * R; i++) x[i] = src[min(i, EW - 1 - (u0 - R + 2 * R))];
            float* dst = s_tmp + (v + 2 * R) * MS + (u0 + R);
#pragma unroll
            for (int j = 0; j < B; j++) {
                float acc = kSumStart;
#pragma unroll
                for (int i = 0; i < K; i++) acc += kk[i] * x[j + i];
                if (u0 + j < TX + R) dst[j] = acc;
            }
        }
        __syncthreads();
        // vertical blur -> conv, ratio: rows [-R, TY+R), cols [-R, TX+R)
        constexpr int V1R = (YH + B - 1) / B;
        for (int t = tid; t < YW * V1R; t += NT) {
            const int u = t % YW - R, v0 = (t / YW) * B - R;
            // y of the task's rows, loaded before the taps are summed
            float yv[B];
            const float* yr = yp + (size_t)(oy + v0) * (size_t)w + (ox + u);
#pragma unroll
            for (int j = 0; j < B; j++) {
                yv[j] = v0 + j < TY + R ? *yr : 1.0f;
                yr += w;
            }
            if (FIRST) {
#pragma unroll
                for (int j = 0; j < B; j++) y_ok = y_ok && yv[j] <= kDivSafeMax;
            }
            const float* src = s_tmp + (v0 - R + 2 * R) * MS + (u + R);
            float x[B + 2 * R];
#pragma unroll
            for (int i = 0; i < B + 2 * R; i++) x[i] = src[min(i, MH - 1 - (v0 - R + 2 * R)) * MS];
#pragma unroll
            for (int j = 0; j < B; j++) {
                float acc = kSumStart;
#pragma unroll
                for (int i = 0; i < K; i++) acc += kk[i] * x[j + i];
                const float c = fmaxr(acc, 1e-6f);
                if (v0 + j < TY + R) s_rat[(v0 + j + R) * YS + (u + R)] = FAST ? div_in_range(yv[j], c) : yv[j] / c;
            }
        }
        __syncthreads();
        // horizontal blur of ratio: rows [-R, TY+R), cols [0, TX)
        constexpr int H2R = QW / B;
        for (int t = tid; t < QH * H2R; t += NT) {
            const int v = t % QH - R, u0 = (t / QH) * B;
            const float* src = s_rat + (v + R) * YS + u0;
            float x[B + 2 * R];
#pragma unroll
            for (int i = 0; i < B + 2 * R; i++) x[i] = src[i];
            float* dst = s_tmp + (v + R) * QS + u0;
#pragma unroll
            for (int j = 0; j < B; j++) {
                float acc = kSumStart;
#pragma unroll
                for (int i = 0; i < K; i++) acc += kk[i] * x[j + i];
                dst[j] = acc;
            }
        }
        __syncthreads();
        // vertical blur -> corr, est update: rows [0, TY), cols [0, TX)
        constexpr int V2R = TY / B;
#pragma unroll
        for (int q = 0; q < NV2; q++) {
            const int t = tid + q * NT;
            if (t >= QW * V2R) break;
            const int u = t % QW, v0 = (t / QW) * B;
            const float* src = s_tmp + v0 * QS + u;
            float x[B + 2 * R];
#pragma unroll
            for (int i = 0; i < B + 2 * R; i++) x[i] = src[i * QS];
            const size_t p0 = (size_t)(oy + v0) * (size_t)w + (ox + u);
            float yl[LAST ? B : 1];
            if (LAST) {
#pragma unroll
                for (int j = 0; j < B; j++) yl[j] = yp[p0 + (size_t)j * w];
            }
#pragma unroll
            for (int j = 0; j < B; j++) {
                float acc = kSumStart;
#pragma unroll
                for (int i = 0; i < K; i++) acc += kk[i] * x[j + i];
                const float e = clamp64(ev[q * B + j] * acc);
                if (LAST)
                    apply_px(rgb, e, yl[j], p0 + (size_t)j * w, rgb_out);
                else
                    est_out[p0 + (size_t)j * w] = e;
            }
        }
        if (FIRST && !y_ok) *flag = 1;
    };
    if (interior) {
        if (FIRST || __builtin_amdgcn_readfirstlane(*flag) != 0)
            interior_passes(std::false_type{});
        else
            interior_passes(std::true_type{});
        return;
    }

    // Frame-edge tiles: each stage evaluates at the clamped image coordinate
    // of its slot, so a slot outside the image holds the value of the edge
    // pixel it clamps to.
    auto cu_of = [&](int u) { return clampi(ox + u, 0, w - 1) - ox; };
    auto cv_of = [&](int v) { return clampi(oy + v, 0, h - 1) - oy; };
    for (int i = tid; i < MH * MW; i += NT) {
        int u = i % MW - R, v = i / MW - 2 * R;
        int cu = cu_of(u);
        float acc = kSumStart;
#pragma unroll
        for (int j = 0; j < K; j++) acc += kk[j] * E(cu + j - R, v);
        s_tmp[(v + 2 * R) * MS + (u + R)] = acc;
    }
    __syncthreads();
    for (int i = tid; i < YH * YW; i += NT) {
        int u = i % YW - R, v = i / YW - R;
        int cv = cv_of(v);
        float acc = kSumStart;
#pragma unroll
        for (int j = 0; j < K; j++) acc += kk[j] * s_tmp[(cv + j - R + 2 * R) * MS + (u + R)];
        const float yv = yp[(size_t)(oy + cv) * w + (ox + cu_of(u))];  // y at the clamped pixel
        if (FIRST && !(yv <= kDivSafeMax)) *flag = 1;
        s_rat[(v + R) * YS + (u + R)] = yv / fmaxr(acc, 1e-6f);
    }
    __syncthreads();
    for (int i = tid; i < QH * QW; i += NT) {
        int u = i % QW, v = i / QW - R;
        int cu = cu_of(u);
        float acc = kSumStart;
#pragma unroll
        for (int j = 0; j < K; j++) acc += kk[j] * s_rat[(v + R) * YS + (cu + j - R + R)];
        s_tmp[(v + R) * QS + u] = acc;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NE; k++) {
        const int i = tid + k * NT;
        int u = i % TX, v = i / TX;
        int gx = ox + u, gy = oy + v;
        if (i >= TX * TY || gx >= w || gy >= h) continue;
        float acc = kSumStart;
#pragma unroll
        for (int j = 0; j < K; j++) acc += kk[j] * s_tmp[(v + j - R + R) * QS + u];
        const size_t p = (size_t)gy * w + gx;
        if (LAST)
            apply_px(rgb, clamp64(ev[k] * acc), yp[p], p, rgb_out);
        else
            est_out[p] = clamp64(ev[k] * acc);
    }
}

// ---- plane-wide fallback pieces
__global__ void k_rl_ratio(const float* __restrict__ y, const float* __restrict__ conv, size_t n, float* __restrict__ ratio) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ratio[i] = y[i] / fmaxr(conv[i], 1e-6f);
}
__global__ void k_rl_update(float* __restrict__ est, const float* __restrict__ corr, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) est[i] = clampr(est[i] * corr[i], 0.0f, 64.0f);
}

// mode: bit 0 the call's first iteration, bit 1 its last (rgb -> rgb_out)
template <int R, int TY, int NT>
hipError_t launch_iter_ty(const float* est, const float* y, float* est_out, size_t w, size_t h, const Taps& t, int mode, int* flag,
                          const float* rgb, float* rgb_out, hipStream_t st) {
    // interior tiles: ox >= 2R, oy >= 2R, ox + TX + 2R <= w, oy + TY + 2R <= h
    const TileOrder order = tile_order((int)w, (int)h, RL_TX, TY, 2 * R, 2 * R);
    dim3 grid((unsigned)(order.nx * order.ny));
#define DS_RL_ITER(M) \
    hipLaunchKernelGGL((k_rl_iter<R, TY, NT, M>), grid, dim3(NT), 0, st, est, y, est_out, (int)w, (int)h, t, order, flag, rgb, rgb_out)
    switch (mode & 3) {
        case 0: DS_RL_ITER(0); break;
        case 1: DS_RL_ITER(1); break;
        case 2: DS_RL_ITER(2); break;
        default: DS_RL_ITER(3); break;
    }
#undef DS_RL_ITER
    return hipGetLastError();
}

// Tile-shape variants of the iteration kernel (64 x TY, same results):
// 0 = 64x64 in 512 threads (45 KB, three tiles and 6 waves per SIMD per CU;
// 8 % less halo work than 64x32), 1 = 64x16, 2 = 64x32 (26 KB, six tiles of
// 256 threads per CU).
template <int R>
hipError_t launch_iter(const float* est, const float* y, float* est_out, size_t w, size_t h, const Taps& t, int variant, int mode, int* flag,
                       const float* rgb, float* rgb_out, hipStream_t st) {
    switch (variant) {
        case 1: return launch_iter_ty<R, 16, 256>(est, y, est_out, w, h, t, mode, flag, rgb, rgb_out, st);
        case 2: return launch_iter_ty<R, 32, 256>(est, y, est_out, w, h, t, mode, flag, rgb, rgb_out, st);
        default: return launch_iter_ty<R, 64, 512>(est, y, est_out, w, h, t, mode, flag, rgb, rgb_out, st);
    }
}

}  // namespace

// Taps of gaussian_blur_exact (image.rs:342-345), computed on the host with
// the same f32 operations (glibc expf, sequential sum, in-place division).
int gaussian_taps(float sigma, float* k) {
    int r = (int)ceilf(sigma * 3.0f);
    float s = 0.0f;
    for (int i = -r; i <= r; i++) {
        k[i + r] = expf(-(float)(i * i) / (2.0f * sigma * sigma));
    }
    for (int i = 0; i <= 2 * r; i++) s += k[i];
    for (int i = 0; i <= 2 * r; i++) k[i] /= s;
    return r;
}

// five planes and the range flag (64 floats keep the tail 256-byte aligned)
size_t rl_workspace_floats(size_t w, size_t h) { return 5 * w * h + 64; }

hipError_t launch_raw_details(const float* rgb, size_t w, size_t h, float sigma, size_t iters, float* ws, float* rgb_out, hipStream_t st,
                              int variant) {
    const size_t n = w * h;
    if (n == 0) return hipSuccess;
    float* y = ws;
    float* est = ws + n;
    float* est2 = ws + 2 * n;
    float* t0 = ws + 3 * n;
    float* t1 = ws + 4 * n;
    const unsigned nb = (unsigned)((n + 255) / 256);
    int* flag = reinterpret_cast<int*>(ws + 5 * n);
    hipLaunchKernelGGL(k_rl_init, dim3(nb), dim3(256), 0, st, rgb, n, y, est, flag);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    const bool exact = sigma >= 0.3f && sigma < 2.0f;  // image.rs:328-333
    Taps taps{};
    int r = 0;
    if (exact) r = gaussian_taps(sigma, taps.k);
    for (size_t it = 0; it < iters; it++) {
        if (exact) {
            // the last iteration applies its est to rgb itself (k_rl_apply's step)
            const int mode = (it == 0 ? 1 : 0) | (it + 1 == iters ? 2 : 0);
            switch (r) {
                case 1: e = launch_iter<1>(est, y, est2, w, h, taps, variant, mode, flag, rgb, rgb_out, st); break;
                case 2: e = launch_iter<2>(est, y, est2, w, h, taps, variant, mode, flag, rgb, rgb_out, st); break;
                case 3: e = launch_iter<3>(est, y, est2, w, h, taps, variant, mode, flag, rgb, rgb_out, st); break;
                case 4: e = launch_iter<4>(est, y, est2, w, h, taps, variant, mode, flag, rgb, rgb_out, st); break;
                case 5: e = launch_iter<5>(est, y, est2, w, h, taps, variant, mode, flag, rgb, rgb_out, st); break;
                case 6: e = launch_iter<6>(est, y, est2, w, h, taps, variant, mode, flag, rgb, rgb_out, st); break;
                default: return hipErrorInvalidValue;
            }
            if (e != hipSuccess) return e;
            if (it + 1 == iters) return hipSuccess;
            float* t = est;
            est = est2;
            est2 = t;
            continue;
        }
        // conv = gaussian_blur(est); ratio; corr = gaussian_blur(ratio); update
        for (int pass = 0; pass < 2; pass++) {
            const float* in = pass == 0 ? est : t0;
            float* blurred = pass == 0 ? est2 : t1;
            if (sigma < 0.3f) {
                e = hipMemcpyAsync(blurred, in, n * sizeof(float), hipMemcpyDeviceToDevice, st);
            } else {
                size_t radii[3];
                box_radii_for_gauss(sigma, radii);
                // box(box(box(in, r0), r1), r2); t0 is free as the
                // horizontal buffer once `in` has been read by the first box
                float* scratch = pass == 0 ? t1 : est2;
                e = launch_box_blur(in, scratch, blurred, w, h, radii[0], st);
                if (e == hipSuccess) e = hipMemcpyAsync(scratch, blurred, n * sizeof(float), hipMemcpyDeviceToDevice, st);
                if (e == hipSuccess) e = launch_box_blur(scratch, t0, blurred, w, h, radii[1], st);
                if (e == hipSuccess) e = hipMemcpyAsync(scratch, blurred, n * sizeof(float), hipMemcpyDeviceToDevice, st);
                if (e == hipSuccess) e = launch_box_blur(scratch, t0, blurred, w, h, radii[2], st);
            }
            if (e != hipSuccess) return e;
            if (pass == 0) {
                hipLaunchKernelGGL(k_rl_ratio, dim3(nb), dim3(256), 0, st, y, est2, n, t0);
                e = hipGetLastError();
                if (e != hipSuccess) return e;
            }
        }
        hipLaunchKernelGGL(k_rl_update, dim3(nb), dim3(256), 0, st, est, t1, n);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_rl_apply, dim3(nb), dim3(256), 0, st, rgb, est, y, n, rgb_out);
    return hipGetLastError();
}

}  // namespace ds
