// Non-local means (enhance.rs:34-72) for gfx950.
//
// The reference computes, for each search offset (dx, dy), the squared
// difference plane, box-blurs it (the patch distance) and accumulates
// exp(-dist/h^2)-weighted samples.  Here one workgroup owns a TW x TH output
// tile and loops over all offsets with the running (acc, wsum) of its pixels
// in registers.  The input tile plus an (S + P)-pixel halo sits in LDS.  Per
// offset:
//   phase 1: every thread takes one 16-column row segment of the tile rows
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

#include <type_traits>

#include "common.h"
#include "nlm.h"

// The NLM weights are not bit-reproducible against the CPU anyway (exp2 vs
// expf, window-sum order), so products may fuse into FMAs here.
#pragma clang fp contract(fast)

namespace ds {

namespace {

constexpr int NLM_TW = 64;        // tile columns
constexpr int NLM_HR = 64;        // rows of horizontal sums per tile (TH + 2P)
constexpr int NLM_NT = 512;
constexpr int NLM_SEG = 8;        // columns per phase-1 task
constexpr int NLM_NG = NLM_NT / NLM_TW;  // phase-2 row groups (threads per column)
constexpr int NLM_HS = NLM_TW + 4;  // padded row stride of the horizontal-sum buffer
constexpr int NLM_MAX_SP = 24;    // halo limit of the tile kernel
constexpr int NLM_NARROW_SP = 13; // halo limit of the narrow instance (search 10, patch 3)
constexpr int NLM_CHUNK = 7;      // offsets per shifted-sample window (interior tiles, narrow instance)
// Input tile row stride: fixed per instance (so every LDS offset inside the
// offset loop is an immediate) and odd (lanes walking down one column hit
// distinct banks).  The narrow instance, with one horizontal-sum buffer and a
// second barrier per offset, fits three tiles per CU instead of two.
constexpr int NLM_PW_WIDE = NLM_TW + 2 * NLM_MAX_SP + 1;
constexpr int NLM_PW_NARROW = NLM_TW + 2 * NLM_NARROW_SP + 1;
// Tile rows.  Phase 1 forms NLM_HR rows of horizontal sums per tile (TH + 2P
// of them used), phase 2 slides down NLM_NG * ceil(TH / NLM_NG) rows (TH of
// them stored).  Of TH = NLM_HR - 2P and that rounded down to a multiple of
// NLM_NG, the one that forms fewer rows per stored row.
__host__ __device__ constexpr int nlm_rows_formed(int th) { return NLM_HR + NLM_NG * ((th + NLM_NG - 1) / NLM_NG); }
__host__ __device__ constexpr int nlm_th(int P) {
    return nlm_rows_formed((NLM_HR - 2 * P) / NLM_NG * NLM_NG) * (NLM_HR - 2 * P) < nlm_rows_formed(NLM_HR - 2 * P) * ((NLM_HR - 2 * P) / NLM_NG * NLM_NG)
               ? (NLM_HR - 2 * P) / NLM_NG * NLM_NG
               : NLM_HR - 2 * P;
}
static_assert(nlm_th(3) == 56 && nlm_th(2) == 60 && nlm_th(1) == 62, "tile rows per patch radius");
// phase-2 rows per thread: the tile rows split over NLM_NG row groups
__host__ __device__ constexpr int nlm_rpt(int P) { return (nlm_th(P) + NLM_NG - 1) / NLM_NG; }
// sum rows incl. padding read by the last row group (which runs past the
// tile: NLM_NG * rpt >= TH rows, reading 2P more)
constexpr int NLM_HSR = NLM_HR + NLM_NG;
// input tile (TH + 2 SP rows) plus spare rows read (not stored) by the last
// row group: rows up to NG * rpt + 2 SP - P
__host__ __device__ inline int nlm_pt_rows(int SP, int P) {
    const int a = nlm_th(P) + 2 * SP + 2, b = NLM_NG * nlm_rpt(P) + 2 * SP - P + 1;
    return a > b ? a : b;
}
__host__ __device__ inline int nlm_pt_floats(int SP, int P, int PW) { return (nlm_pt_rows(SP, P) * PW + 3) & ~3; }

// The guide planes at pixel `at` from its NLM result v, each product rounded
// on its own (no contraction), as k_gf_inputs rounds them.
__device__ __forceinline__ void store_guide(const NlmGuide& g, size_t at, float v) {
#pragma clang fp contract(off)
    const float o = g.L[at];
    const float I = o + (v - o) * g.m;
    g.I[at] = I;
    g.IA[at] = I * g.A[at];
    g.IB[at] = I * g.B[at];
    g.II[at] = I * I;
}

// Tile kernel.  Phase 1: thread t owns hs row (t % 64) and the column
// segments t / 64 + NSG k, k < NTASK (TW / (NSG SEG) tasks per thread, NSG =
// NT / 64).  The centre samples of its segments do not depend on the offset
// and stay in registers for the whole offset loop.  Phase 2: thread t owns
// column t % 64 and one of NLM_NG groups of tile rows.
template <int P, int PW, int HSB>
__global__ __launch_bounds__(NLM_NT, HSB == 1 ? 6 : 2) void k_nlm_tile(const float* __restrict__ in, float* __restrict__ out, int w, int h, int S,
                                                              NlmWeight wp, const uint8_t* __restrict__ regions, TileOrder order,
                                                              NlmGuide guide) {
    extern __shared__ float lds[];
    constexpr int TW = NLM_TW, TH = nlm_th(P), NT = NLM_NT, SEG = NLM_SEG, HS = NLM_HS;
    constexpr int DW = SEG + 2 * P;           // differences per phase-1 task
    constexpr int NSG = NT / NLM_HR;          // phase-1 segment groups
    constexpr int NTASK = TW / (SEG * NSG);
    static_assert(NTASK * SEG * NSG == TW && SEG % 4 == 0, "phase-1 tasks must tile the rows");
    constexpr int RPT = nlm_rpt(P);           // phase-2 rows per thread
    static_assert(NLM_NG * RPT + 2 * P <= NLM_HSR, "the last row group's sums are inside the buffer");
    // The tile is staged scaled by cs = sqrt(-scale), so a window sum of
    // squared differences is already the exp2 argument; the accumulated
    // samples carry the same factor, removed once per output.
    const float cs = sqrtf(-wp.scale());
    const int SP = S + P;
    const int PWL = TW + 2 * SP;              // loaded tile row length
    const int PH = TH + 2 * SP;
    float* pt = lds;                          // input tile, PH x PW
    float* hs0 = lds + nlm_pt_floats(SP, P, PW);  // horizontal sums, HSB x HSR x HS
    const int tid = threadIdx.x;
    // frame-edge tiles (the clamped offset loop, ~1.75x the interior one's
    // instructions) first, so they do not stretch the last round
    int tile_x, tile_y;
    tile_edge_first(blockIdx.x, order, tile_x, tile_y);
    const int ox = tile_x * TW, oy = tile_y * TH;
    // region mode: only tiles that touch a flagged region run, and they store
    // only the pixels of flagged regions
    const int nrx = (w + NLM_REGION - 1) / NLM_REGION;
    if (regions) {
        const int rx0 = ox / NLM_REGION, rx1 = (min(ox + TW, w) - 1) / NLM_REGION;
        const int ry0 = oy / NLM_REGION, ry1 = (min(oy + TH, h) - 1) / NLM_REGION;
        bool any = false;
        for (int ry = ry0; ry <= ry1; ry++)
            for (int rx = rx0; rx <= rx1; rx++) any = any || regions[ry * nrx + rx] != 0;
        if (!any) return;
    }

    // input tile with clamped coordinates: slot (u, v) holds p[clamp(ox+u), clamp(oy+v)]
    stage_tile_n<16, NT>(
        PH * PWL,
        [&](int i) {
            int gx = clampi(ox + i % PWL - SP, 0, w - 1), gy = clampi(oy + i / PWL - SP, 0, h - 1);
            return in[(size_t)gy * w + gx] * cs;
        },
        [&](int i, float v) { pt[(i / PWL) * PW + (i % PWL)] = v; });
    auto P_at = [&](int u, int v) -> float { return pt[(v + SP) * PW + (u + SP)]; };
    const bool interior = ox - SP >= 0 && oy - SP >= 0 && ox + TW + SP <= w && oy + TH + SP <= h;
    __syncthreads();

    // phase-1 geometry
    const int v1 = tid % NLM_HR - P;  // tile row of this thread's horizontal sums
    const int u00 = (tid / NLM_HR) * SEG;  // first segment; task k starts at u00 + NSG k SEG
    // window positions (tile coords) of the two segments; at the frame edge the
    // difference plane is evaluated at the clamped position (box_blur's clamp)
    auto cu_of = [&](int u) { return interior ? u : clampi(ox + u, 0, w - 1) - ox; };
    const int cv = interior ? v1 : clampi(oy + v1, 0, h - 1) - oy;
    float cc[NTASK][DW];
#pragma unroll
    for (int task = 0; task < NTASK; task++)
#pragma unroll
        for (int i = 0; i < DW; i++) cc[task][i] = P_at(cu_of(u00 + NSG * SEG * task - P + i), cv);

    // phase-2 geometry
    const int col = tid % TW;
    const int row0 = (tid / TW) * RPT;
    float acc[RPT], wsum[RPT];
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        acc[k] = 0.0f;
        wsum[k] = 0.0f;
    }

    // phase 1 of one task from its differences d (not squared): the sliding
    // window sums, as the reference's box blur slides them, with each square
    // folded into a fused multiply-add: the entering square is added and the
    // leaving one subtracted unrounded, two operations per step instead of a
    // square, a difference and an add (a prefix-sum form saves an operation
    // per sum but moves the denoise composite outside the reference's own
    // spread)
    auto hsums = [&](const float (&d)[DW], float* hs, int u0) {
        float hsum[SEG];
        float s = d[0] * d[0];
#pragma unroll
        for (int i = 1; i < 2 * P + 1; i++) s = __builtin_fmaf(d[i], d[i], s);
#pragma unroll
        for (int j = 0; j < SEG; j++) {
            hsum[j] = s;
            if (j + 1 < SEG) {
                s = __builtin_fmaf(d[j + 2 * P + 1], d[j + 2 * P + 1], s);
                s = __builtin_fmaf(-d[j], d[j], s);
            }
        }
        float4* dst = reinterpret_cast<float4*>(hs + (v1 + P) * HS + u0);
#pragma unroll
        for (int q = 0; q < SEG / 4; q++) dst[q] = make_float4(hsum[4 * q], hsum[4 * q + 1], hsum[4 * q + 2], hsum[4 * q + 3]);
    };
    // phase 2 of the offset at doff (dy * PW + dx): vertical window sums,
    // weights, accumulation
    auto phase2 = [&](const float* hs, int doff) {
        const float* hc = hs + row0 * HS + col;  // hs row r holds tile row r - P
        const float* qp = pt + (row0 + SP) * PW + (col + SP) + doff;
        // -0.0 + t == t for every t, so the first add folds away
        float s = -0.0f;
#pragma unroll
        for (int j = 0; j < 2 * P + 1; j++) s += hc[j * HS];
        // every row segment runs RPT rows (the last one past the tile reads
        // padding rows of hs and is not stored), so the loop has no per-lane
        // branches
#pragma unroll
        for (int k = 0; k < RPT; k++) {
            const float wt = __builtin_amdgcn_exp2f(-s);
            acc[k] += wt * qp[k * PW];
            wsum[k] += wt;
            if (k + 1 < RPT) s += hc[(k + 2 * P + 1) * HS] - hc[k * HS];
        }
    };

    // The offset loop is instantiated once for interior tiles (no clamps)
    // and once for frame-edge tiles, so neither pays for the other's path.
    auto offsets = [&](auto in_tag) {
        constexpr bool IN = decltype(in_tag)::value;
        const int side = 2 * S + 1;
        const int noff = side * side;
        if constexpr (IN && HSB == 1) {
            // Interior tiles, rows of offsets that split into chunks of
            // NLM_CHUNK: the shifted samples of a task for dx .. dx + CHUNK - 1
            // are one window of DW + CHUNK - 1 samples, read once per chunk
            // instead of DW per offset.  Same offsets, same order, same
            // arithmetic as the loop below.
            if (side % NLM_CHUNK == 0) {
                int drow = -S * PW - S;  // dy * PW - S
                for (int dy = -S; dy <= S; dy++, drow += PW) {
                    for (int c0 = 0; c0 < side; c0 += NLM_CHUNK) {
                        const int doff0 = drow + c0;
                        float win[NTASK][DW + NLM_CHUNK - 1];
#pragma unroll
                        for (int task = 0; task < NTASK; task++) {
                            const float* sp = pt + (cv + SP) * PW + (u00 + NSG * SEG * task - P + SP) + doff0;
#pragma unroll
                            for (int i = 0; i < DW + NLM_CHUNK - 1; i++) win[task][i] = sp[i];
                        }
#pragma unroll
                        for (int j = 0; j < NLM_CHUNK; j++) {
#pragma unroll
                            for (int task = 0; task < NTASK; task++) {
                                float d[DW];
#pragma unroll
                                for (int i = 0; i < DW; i++) d[i] = cc[task][i] - win[task][i + j];
                                hsums(d, hs0, u00 + NSG * SEG * task);
                            }
                            __syncthreads();
                            phase2(hs0, doff0 + j);
                            __syncthreads();
                        }
                    }
                }
                return;
            }
        }
        int dx = -S, dy = -S;  // offsets in the reference's order: dy outer, dx inner
        int doff = -S * PW - S;  // dy * PW + dx, advanced with the offset
        for (int o = 0; o < noff; o++) {
            float* hs = HSB == 2 ? hs0 + (o & 1) * (NLM_HSR * HS) : hs0;
            // ---- phase 1: horizontal window sums of the squared differences
#pragma unroll
            for (int task = 0; task < NTASK; task++) {
                const float* c = cc[task];
                const int u0 = u00 + NSG * SEG * task;
                float d[DW];  // differences (not squared)
                if constexpr (IN) {
                    const float* sp = pt + (cv + SP) * PW + (u0 - P + SP) + doff;
#pragma unroll
                    for (int i = 0; i < DW; i++) d[i] = c[i] - sp[i];
                } else {
                    const int sv = clampi(oy + cv + dy, 0, h - 1) - oy;
#pragma unroll
                    for (int i = 0; i < DW; i++) {
                        const int su = clampi(ox + cu_of(u0 - P + i) + dx, 0, w - 1) - ox;
                        d[i] = c[i] - P_at(su, sv);
                    }
                }
                hsums(d, hs, u0);
            }
            __syncthreads();
            // ---- phase 2: vertical window sums, weights, accumulation
            phase2(hs, doff);
            // with two buffers the other one is written next, and the barrier
            // after that phase 1 orders it against this phase 2; with one,
            // phase 1 of the next offset waits here
            if constexpr (HSB == 1) __syncthreads();
            ++doff;
            if (++dx > S) {
                dx = -S;
                ++dy;
                doff += PW - side;
            }
        }
    };
    if (interior)
        offsets(std::true_type{});
    else
        offsets(std::false_type{});
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        const int gx = ox + col, gy = oy + row0 + k;
        if (row0 + k < TH && gx < w && gy < h && (!regions || regions[(gy / NLM_REGION) * nrx + gx / NLM_REGION])) {
            const float v = (acc[k] / cs) / fmaxf(wsum[k], 1e-12f);
            if (guide.I)
                store_guide(guide, (size_t)gy * w + gx, v);
            else
                out[(size_t)gy * w + gx] = v;
        }
    }
}

// NLM at one pixel by direct patch distances, with the two-level clamp of the
// reference.  Also returns the weight sums (sum w, sum w^2).
__device__ __forceinline__ float nlm_pixel(const float* __restrict__ in, int w, int h, int S, int P, float inv_h2, float norm, int x, int y,
                                           float& wsum_out, float& w2sum_out) {
    auto at = [&](int u, int v) { return in[(size_t)clampi(v, 0, h - 1) * w + clampi(u, 0, w - 1)]; };
    float acc = 0.0f, wsum = 0.0f, w2sum = 0.0f;
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
            w2sum += wt * wt;
        }
    wsum_out = wsum;
    w2sum_out = w2sum;
    return acc / fmaxf(wsum, 1e-12f);
}

// Any search / patch radius: one pixel per thread.  Used when the tile
// kernel's halo or patch limit is exceeded.
__global__ void k_nlm_direct(const float* __restrict__ in, float* __restrict__ out, int w, int h, int S, int P, NlmWeight wp,
                             float norm, const uint8_t* __restrict__ regions) {
    const float inv_h2 = 1.0f / wp.h2();
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= w || y >= h) return;
    if (regions && !regions[(y / NLM_REGION) * ((w + NLM_REGION - 1) / NLM_REGION) + x / NLM_REGION]) return;
    float ws, w2;
    out[(size_t)y * w + x] = nlm_pixel(in, w, h, S, P, inv_h2, norm, x, y, ws, w2);
}

// Spot check of a candidate result (the learned denoiser's) against NLM.
// One block per NLM_REGION x NLM_REGION region; its NLM_SPOT x NLM_SPOT
// threads take a lattice of pixels spread evenly over the region.  Each
// computes NLM there and the variance NLM leaves of white noise of the
// estimated sigma, sigma^2 sum w^2 / (sum w)^2.  The region's deviation
// from NLM, in dB of PSNR, is 10 log10(1 + mean (cand - nlm)^2 / mean var):
// the PSNR the candidate gives up if its deviation from NLM is independent
// of NLM's own residual noise, taking that residual at its variance part.
__global__ __launch_bounds__(NLM_SPOT* NLM_SPOT) void k_nlm_spot(const float* __restrict__ in, const float* __restrict__ cand, int w, int h,
                                                                 int S, int P, NlmWeight wp, float norm, float thr_db,
                                                                 float* __restrict__ gap_db, uint8_t* __restrict__ flags) {
    constexpr int NT = NLM_SPOT * NLM_SPOT;
    __shared__ float red[2][NT];
    const int rx = blockIdx.x, ry = blockIdx.y;
    const int x0 = rx * NLM_REGION, y0 = ry * NLM_REGION;
    const int rw = min(NLM_REGION, w - x0), rh = min(NLM_REGION, h - y0);
    const int i = threadIdx.x % NLM_SPOT, j = threadIdx.x / NLM_SPOT;
    const int x = x0 + ((2 * i + 1) * rw) / (2 * NLM_SPOT), y = y0 + ((2 * j + 1) * rh) / (2 * NLM_SPOT);
    const float inv_h2 = 1.0f / wp.h2();
    float ws, w2;
    const float v = nlm_pixel(in, w, h, S, P, inv_h2, norm, x, y, ws, w2);
    const float sg = wp.sigma_dev ? fmaxf(*wp.sigma_dev, 1e-4f) : 1e-4f;
    const float e = cand[(size_t)y * w + x] - v;
    red[0][threadIdx.x] = e * e;
    red[1][threadIdx.x] = sg * sg * w2 / fmaxf(ws * ws, 1e-30f);
    __syncthreads();
    for (int s = NT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            red[0][threadIdx.x] += red[0][threadIdx.x + s];
            red[1][threadIdx.x] += red[1][threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float g = 10.0f * log10f(1.0f + red[0][0] / fmaxf(red[1][0], 1e-30f));
        const int r = ry * gridDim.x + rx;
        gap_db[r] = g;
        flags[r] = (thr_db <= 0.0f || !(g <= thr_db)) ? 1 : 0;  // NaN falls back too
    }
}

// Spot check with the neighbourhood in LDS: one wave per lattice pixel.  A
// block of SPOT_WAVES waves takes SPOT_PER_WAVE pixels per wave; SPOT_BLOCKS
// blocks cover a region's NLM_SPOT x NLM_SPOT lattice.  The wave stages the
// (2(S+P)+1)^2 clamped neighbourhood of its pixel, its lanes split the search
// offsets, and a wave reduction gives NLM and the weight sums.  Each wave
// writes its (sum e^2, sum var) pair; k_nlm_spot_final turns a region's
// pairs into the deviation and the flag.
constexpr int SPOT_WAVES = 4;
constexpr int SPOT_PER_WAVE = 4;
constexpr int SPOT_BLOCKS = NLM_SPOT * NLM_SPOT / (SPOT_WAVES * SPOT_PER_WAVE);
constexpr int SPOT_PARTS = SPOT_BLOCKS * SPOT_WAVES;  // (e^2, var) pairs per region
constexpr int SPOT_MAX_SP = 24;
constexpr int SPOT_ROLL = 7;  // dy offsets per lane in the rolling-window spot check (search 10: 3 x 21 lanes)
constexpr int SPOT_PF = 12;  // prefetched samples per lane: neighbourhoods up to 27 x 27 (S + P <= 13)

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

template <int P>
__global__ __launch_bounds__(SPOT_WAVES * 64) void k_nlm_spot_lds(const float* __restrict__ in, const float* __restrict__ cand, int w, int h,
                                                                  int S, NlmWeight wp, float norm, float* __restrict__ part) {
    extern __shared__ float lds[];
    const int SP = S + P, D = 2 * SP + 1;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    float* nb = lds + wave * D * D;
    const int rx = blockIdx.y, ry = blockIdx.z;
    const int x0 = rx * NLM_REGION, y0 = ry * NLM_REGION;
    const int rw = min(NLM_REGION, w - x0), rh = min(NLM_REGION, h - y0);
    // exp(-(sum of squared differences) norm^2 / h^2) as exp2 of the sum times
    // one factor; the check only sizes a deviation, so it takes fused
    // multiply-adds and the hardware exp2 rather than the reference's
    // operation order
    const float kexp = -(norm * norm) * 1.44269504088896340736f / wp.h2();
    const float sg = wp.sigma_dev ? fmaxf(*wp.sigma_dev, 1e-4f) : 1e-4f;
    const int side = 2 * S + 1, noff = side * side;
    float se = 0.0f, sv = 0.0f;
    // Each wave owns its nb region, so the waves need no block barrier: a
    // wave's LDS accesses complete in program order.  A neighbourhood of up
    // to 64 SPOT_PF samples is fetched into registers one lattice pixel
    // ahead, under the previous pixel's offsets.
    auto pixel = [&](int k, int& x, int& y) {
        const int sidx = (blockIdx.x * SPOT_WAVES + wave) * SPOT_PER_WAVE + k;
        const int i = sidx % NLM_SPOT, j = sidx / NLM_SPOT;
        x = x0 + ((2 * i + 1) * rw) / (2 * NLM_SPOT);
        y = y0 + ((2 * j + 1) * rh) / (2 * NLM_SPOT);
    };
    const bool pf = D * D <= 64 * SPOT_PF;
    float nxt[SPOT_PF];
    auto fetch = [&](int k) {
        int x, y;
        pixel(k, x, y);
#pragma unroll
        for (int q = 0; q < SPOT_PF; q++) {
            const int t = lane + 64 * q;
            nxt[q] = t < D * D ? in[(size_t)clampi(y + t / D - SP, 0, h - 1) * w + clampi(x + t % D - SP, 0, w - 1)] : 0.0f;
        }
    };
    if (pf) fetch(0);
    for (int k = 0; k < SPOT_PER_WAVE; k++) {
        int x, y;
        pixel(k, x, y);
        // nb[(v + SP) * D + (u + SP)] = p[clamp(x + u), clamp(y + v)]
        if (pf) {
#pragma unroll
            for (int q = 0; q < SPOT_PF; q++) {
                const int t = lane + 64 * q;
                if (t < D * D) nb[t] = nxt[q];
            }
        } else {
            for (int base = 0; base < D * D; base += 64 * 8) {
                float r[8];
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const int t = base + lane + 64 * q;
                    r[q] = t < D * D ? in[(size_t)clampi(y + t / D - SP, 0, h - 1) * w + clampi(x + t % D - SP, 0, w - 1)] : 0.0f;
                }
#pragma unroll
                for (int q = 0; q < 8; q++) {
                    const int t = base + lane + 64 * q;
                    if (t < D * D) nb[t] = r[q];
                }
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (pf && k + 1 < SPOT_PER_WAVE) fetch(k + 1);
        const bool interior = x - SP >= 0 && y - SP >= 0 && x + SP < w && y + SP < h;
        // centre patch, as the two-level clamp sees it: window pixel (cx, cy) = clamp(x + i, y + j)
        float cpat[2 * P + 1][2 * P + 1];
        int cxs[2 * P + 1], cys[2 * P + 1];
#pragma unroll
        for (int q = -P; q <= P; q++) {
            cxs[q + P] = clampi(x + q, 0, w - 1) - x;
            cys[q + P] = clampi(y + q, 0, h - 1) - y;
        }
#pragma unroll
        for (int jj = 0; jj < 2 * P + 1; jj++)
#pragma unroll
            for (int ii = 0; ii < 2 * P + 1; ii++) {
                // the same in every lane (one lattice pixel per wave): held in
                // scalar registers, so the vector ones go to the offsets' samples
                const float v = nb[(cys[jj] + SP) * D + cxs[ii] + SP];
                cpat[jj][ii] = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, v)));
            }
        float acc = 0.0f, ws = 0.0f, w2 = 0.0f;
        if (interior && side == 3 * SPOT_ROLL) {
            // Search 10: lane g * side + i takes dx = i - S and the SPOT_ROLL
            // offsets dy = SPOT_ROLL g - S + t.  Going down one dy, the patch
            // under the offset gains one row and loses one, so each step reads
            // only the new row (2P + 1 samples) into a rolling register window
            // instead of the whole patch.
            if (lane < 3 * side) {
                constexpr int PD = 2 * P + 1;
                const int dx = lane % side - S, dy0 = SPOT_ROLL * (lane / side) - S;
                const float* q0 = nb + (SP - P + dy0) * D + (SP - P + dx);
                float win[PD + SPOT_ROLL - 1][PD];
#pragma unroll
                for (int r = 0; r < PD - 1; r++)
#pragma unroll
                    for (int ii = 0; ii < PD; ii++) win[r][ii] = q0[r * D + ii];
#pragma unroll
                for (int t = 0; t < SPOT_ROLL; t++) {
#pragma unroll
                    for (int ii = 0; ii < PD; ii++) win[t + PD - 1][ii] = q0[(t + PD - 1) * D + ii];
                    float dist = 0.0f;
#pragma unroll
                    for (int jj = 0; jj < PD; jj++)
#pragma unroll
                        for (int ii = 0; ii < PD; ii++) {
                            const float e = cpat[jj][ii] - win[t + jj][ii];
                            dist = __builtin_fmaf(e, e, dist);
                        }
                    const float wt = __builtin_amdgcn_exp2f(dist * kexp);
                    acc += wt * nb[(dy0 + t + SP) * D + dx + SP];
                    ws += wt;
                    w2 += wt * wt;
                }
            }
        } else
        for (int o = lane; o < noff; o += 64) {
            const int dy = o / side - S, dx = o % side - S;
            float dist = 0.0f;
            if (interior) {
                const float* q = nb + (SP - P + dy) * D + (SP - P + dx);
#pragma unroll
                for (int jj = 0; jj < 2 * P + 1; jj++)
#pragma unroll
                    for (int ii = 0; ii < 2 * P + 1; ii++) {
                        const float e = cpat[jj][ii] - q[jj * D + ii];
                        dist = __builtin_fmaf(e, e, dist);
                    }
            } else {
#pragma unroll
                for (int jj = 0; jj < 2 * P + 1; jj++) {
                    const int sy = clampi(y + cys[jj] + dy, 0, h - 1) - y + SP;
#pragma unroll
                    for (int ii = 0; ii < 2 * P + 1; ii++) {
                        const int sx = clampi(x + cxs[ii] + dx, 0, w - 1) - x + SP;
                        const float e = cpat[jj][ii] - nb[sy * D + sx];
                        dist = __builtin_fmaf(e, e, dist);
                    }
                }
            }
            const float wt = __builtin_amdgcn_exp2f(dist * kexp);
            acc += wt * nb[(dy + SP) * D + dx + SP];
            ws += wt;
            w2 += wt * wt;
        }
        acc = wave_sum(acc);
        ws = wave_sum(ws);
        w2 = wave_sum(w2);
        const float v = acc / fmaxf(ws, 1e-12f);
        const float e = cand[(size_t)y * w + x] - v;
        se += e * e;
        sv += sg * sg * w2 / fmaxf(ws * ws, 1e-30f);
        // every lane's reads of nb are done before it is restaged
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    if (lane == 0) {
        float* pp = part + ((size_t)(ry * gridDim.y + rx) * SPOT_PARTS + blockIdx.x * SPOT_WAVES + wave) * 2;
        pp[0] = se;
        pp[1] = sv;
    }
}

__global__ __launch_bounds__(64) void k_nlm_spot_final(const float* __restrict__ part, float thr_db, float* __restrict__ gap_db,
                                                       uint8_t* __restrict__ flags) {
    const int r = blockIdx.x;
    float a = 0.0f, b = 0.0f;
    for (int q = threadIdx.x; q < SPOT_PARTS; q += 64) {
        a += part[((size_t)r * SPOT_PARTS + q) * 2];
        b += part[((size_t)r * SPOT_PARTS + q) * 2 + 1];
    }
    a = wave_sum(a);
    b = wave_sum(b);
    if (threadIdx.x == 0) {
        const float g = 10.0f * log10f(1.0f + a / fmaxf(b, 1e-30f));
        gap_db[r] = g;
        flags[r] = (thr_db <= 0.0f || !(g <= thr_db)) ? 1 : 0;  // NaN falls back too
    }
}

template <int P, int PW, int HSB>
hipError_t launch_tile_pw(const float* in, float* out, int w, int h, int S, NlmWeight wp, const uint8_t* regions, const NlmGuide& guide,
                          hipStream_t stream) {
    const size_t lds = (size_t)(nlm_pt_floats(S + P, P, PW) + HSB * NLM_HSR * NLM_HS) * sizeof(float);
    // interior tiles: ox >= S + P, oy >= S + P, ox + TW + S + P <= w, oy + TH + S + P <= h
    const TileOrder order = tile_order(w, h, NLM_TW, nlm_th(P), S + P, S + P);
    dim3 grid((unsigned)(order.nx * order.ny));
    static bool attr_set = false;
    if (!attr_set) {
        const int sp_max = PW == NLM_PW_NARROW ? NLM_NARROW_SP : NLM_MAX_SP;
        const size_t max_lds = (size_t)(nlm_pt_floats(sp_max, P, PW) + HSB * NLM_HSR * NLM_HS) * sizeof(float);
        hipError_t e = hipFuncSetAttribute((const void*)k_nlm_tile<P, PW, HSB>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_lds);
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    hipLaunchKernelGGL((k_nlm_tile<P, PW, HSB>), grid, dim3(NLM_NT), lds, stream, in, out, w, h, S, wp, regions, order, guide);
    return hipGetLastError();
}

template <int P>
hipError_t launch_tile(const float* in, float* out, int w, int h, int S, NlmWeight wp, const uint8_t* regions, const NlmGuide& guide,
                       hipStream_t stream) {
    if (S + P <= NLM_NARROW_SP) return launch_tile_pw<P, NLM_PW_NARROW, 1>(in, out, w, h, S, wp, regions, guide, stream);
    return launch_tile_pw<P, NLM_PW_WIDE, 2>(in, out, w, h, S, wp, regions, guide, stream);
}

}  // namespace

size_t nlm_tile_lds_bytes(int S, int P) {
    if (S + P <= NLM_NARROW_SP) return (size_t)(nlm_pt_floats(S + P, P, NLM_PW_NARROW) + NLM_HSR * NLM_HS) * sizeof(float);
    return (size_t)(nlm_pt_floats(S + P, P, NLM_PW_WIDE) + 2 * NLM_HSR * NLM_HS) * sizeof(float);
}

namespace {
NlmWeight make_weight(NlmH hp, int P, double& norm) {
    norm = 1.0 / (double)(2 * P + 1);
    NlmWeight wp;
    wp.h_param = hp.h_param;
    wp.sigma_dev = hp.sigma_dev;
    wp.sigma_mul = hp.sigma_mul;
    // weight = exp(-(norm^2 * window_sum) / h2) = exp2(window_sum * scale)
    wp.k = (float)(norm * norm * 1.4426950408889634);
    return wp;
}
}  // namespace

namespace {
// (search, patch) the tile kernel takes; the direct kernel takes the rest
bool nlm_tiled(long search, size_t patch) { return search + (long)patch <= NLM_MAX_SP && patch <= 5; }

// NLM through the tile kernel (with its optional guide output stage) or the
// direct kernel (no guide).
hipError_t run_nlm(const float* in, float* out, size_t w, size_t h, NlmH hp, long search, size_t patch, const uint8_t* regions,
                   const NlmGuide& guide, hipStream_t stream) {
    if (w == 0 || h == 0) return hipSuccess;
    const int S = (int)search;
    const int P = (int)patch;
    double norm;
    const NlmWeight wp = make_weight(hp, P, norm);
    if (nlm_tiled(search, patch)) {
        switch (P) {
            case 0: return launch_tile<0>(in, out, (int)w, (int)h, S, wp, regions, guide, stream);
            case 1: return launch_tile<1>(in, out, (int)w, (int)h, S, wp, regions, guide, stream);
            case 2: return launch_tile<2>(in, out, (int)w, (int)h, S, wp, regions, guide, stream);
            case 3: return launch_tile<3>(in, out, (int)w, (int)h, S, wp, regions, guide, stream);
            case 4: return launch_tile<4>(in, out, (int)w, (int)h, S, wp, regions, guide, stream);
            default: return launch_tile<5>(in, out, (int)w, (int)h, S, wp, regions, guide, stream);
        }
    }
    dim3 grid((unsigned)((w + 255) / 256), (unsigned)h);
    hipLaunchKernelGGL(k_nlm_direct, grid, dim3(256), 0, stream, in, out, (int)w, (int)h, S, P, wp, (float)norm, regions);
    return hipGetLastError();
}
}  // namespace

hipError_t launch_nlm_regions(const float* in, float* out, size_t w, size_t h, NlmH hp, long search, size_t patch, const uint8_t* regions,
                              hipStream_t stream) {
    return run_nlm(in, out, w, h, hp, search, patch, regions, NlmGuide{}, stream);
}

hipError_t launch_nlm_guide(const float* L, float* out, size_t w, size_t h, NlmH hp, long search, size_t patch, const NlmGuide& g, bool* fused,
                            hipStream_t stream) {
    *fused = nlm_tiled(search, patch);
    return run_nlm(L, out, w, h, hp, search, patch, nullptr, *fused ? g : NlmGuide{}, stream);
}

hipError_t launch_nlm(const float* in, float* out, size_t w, size_t h, NlmH hp, long search, size_t patch, hipStream_t stream) {
    return launch_nlm_regions(in, out, w, h, hp, search, patch, nullptr, stream);
}

size_t nlm_regions(size_t w, size_t h) { return ((w + NLM_REGION - 1) / NLM_REGION) * ((h + NLM_REGION - 1) / NLM_REGION); }

size_t nlm_spot_scratch_floats(size_t w, size_t h) { return nlm_regions(w, h) * SPOT_PARTS * 2; }

namespace {
template <int P>
hipError_t launch_spot_lds(const float* in, const float* cand, int w, int h, int S, NlmWeight wp, float norm, float* part, hipStream_t stream) {
    const int D = 2 * (S + P) + 1;
    const size_t lds = (size_t)SPOT_WAVES * D * D * sizeof(float);
    static bool attr_set = false;
    if (!attr_set) {
        const int DM = 2 * SPOT_MAX_SP + 1;
        hipError_t e = hipFuncSetAttribute((const void*)k_nlm_spot_lds<P>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)(SPOT_WAVES * DM * DM * sizeof(float)));
        if (e != hipSuccess) return e;
        attr_set = true;
    }
    dim3 grid(SPOT_BLOCKS, (unsigned)((w + NLM_REGION - 1) / NLM_REGION), (unsigned)((h + NLM_REGION - 1) / NLM_REGION));
    hipLaunchKernelGGL(k_nlm_spot_lds<P>, grid, dim3(SPOT_WAVES * 64), lds, stream, in, cand, w, h, S, wp, norm, part);
    return hipGetLastError();
}
}  // namespace

hipError_t launch_nlm_spot_check(const float* in, const float* cand, size_t w, size_t h, NlmH hp, long search, size_t patch, float thr_db,
                                 float* gap_db, uint8_t* flags, float* scratch, hipStream_t stream) {
    if (w == 0 || h == 0) return hipSuccess;
    double norm;
    const int S = (int)search, P = (int)patch;
    const NlmWeight wp = make_weight(hp, P, norm);
    const unsigned nrx = (unsigned)((w + NLM_REGION - 1) / NLM_REGION), nry = (unsigned)((h + NLM_REGION - 1) / NLM_REGION);
    if (S + P <= SPOT_MAX_SP && P <= 5) {
        hipError_t e = hipSuccess;
        switch (P) {
            case 0: e = launch_spot_lds<0>(in, cand, (int)w, (int)h, S, wp, (float)norm, scratch, stream); break;
            case 1: e = launch_spot_lds<1>(in, cand, (int)w, (int)h, S, wp, (float)norm, scratch, stream); break;
            case 2: e = launch_spot_lds<2>(in, cand, (int)w, (int)h, S, wp, (float)norm, scratch, stream); break;
            case 3: e = launch_spot_lds<3>(in, cand, (int)w, (int)h, S, wp, (float)norm, scratch, stream); break;
            case 4: e = launch_spot_lds<4>(in, cand, (int)w, (int)h, S, wp, (float)norm, scratch, stream); break;
            default: e = launch_spot_lds<5>(in, cand, (int)w, (int)h, S, wp, (float)norm, scratch, stream); break;
        }
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_nlm_spot_final, dim3(nrx * nry), dim3(64), 0, stream, scratch, thr_db, gap_db, flags);
        return hipGetLastError();
    }
    dim3 grid(nrx, nry);
    hipLaunchKernelGGL(k_nlm_spot, grid, dim3(NLM_SPOT * NLM_SPOT), 0, stream, in, cand, (int)w, (int)h, S, P, wp, (float)norm, thr_db, gap_db,
                       flags);
    return hipGetLastError();
}

}  // namespace ds
