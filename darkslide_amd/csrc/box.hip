// Box blur (image.rs:268-313) with the reference's exact running sums.
//
// The reference slides a window sum along each row, then down each column:
//     acc = sum of the clamped window at 0, in order -r..r
//     out[x] = acc * norm;  acc += in[x + r + 1] - in[x - r]   (clamped)
// Every out[x] depends on all previous additions, so the result is only
// bit-reproducible by walking each row / column sequentially in the same
// order.  The guided filter needs that: its var = E[I^2] - E[I]^2 cancels
// about two orders of magnitude of the inputs' precision, so a different
// summation order moves its output by ~1e-4 (more than the whole error budget).
//
// Layout: one wave owns 64 rows (horizontal) or 64 columns (vertical) of one
// plane and walks them in 64-element chunks; the serial part is one
// dependent add per element, and the next chunk's loads are in flight while
// the current chunk is walked.  Plain planes go through ring-buffered
// kernels that read every input once (16-byte loads for the horizontal
// pass); rows whose width is not a multiple of 4 (or planes not 16-byte
// aligned) use k_boxx_h, which forms the window differences from two
// coalesced loads per element.  Rows/columns x planes give the parallelism:
// ~(h/64) * planes waves for the horizontal pass, ~(w/64) * planes for the
// vertical one.

#include <atomic>

#include "box.h"
#include "common.h"

#pragma clang fp contract(off)

namespace ds {

namespace {

constexpr int BX = 64;  // chunk length and rows / columns per wave

// Horizontal pass.  grid: (ceil(h/64), planes); block: one wave.  Per
// 64-column chunk:
//   1. lane t forms the window differences d = in[x+r+1] - in[x-r] of column
//      x = 64c + t for the wave's 64 rows (coalesced loads, clamped indices
//      computed once per chunk) and stores them transposed into LDS;
//   2. lane t walks row y0 + t through the chunk: out = acc * norm;
//      acc += d -- the only sequential work, one add per element, reading
//      LDS at immediate offsets;
//   3. the outputs go back through the same LDS tile and out coalesced.
// The next chunk's loads are issued before the walk, and two LDS tiles
// alternate, so each chunk needs two barriers.
template <class In>
__global__ __launch_bounds__(BX) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_boxx_h(In in, BoxPlanes out, int w, int h, int r,
                                                                                          float norm) {
    if (out.gate && *out.gate == 0) return;
    __shared__ float tile[2][BX][BX + 1];
    const int t = threadIdx.x;
    const int y0 = blockIdx.x * BX;
    const int k = blockIdx.y;
    const uint32_t uw = (uint32_t)w;
    const int ntiles = (w + BX - 1) / BX;
    float a[BX], b[BX];
    auto issue = [&](int c) {
        const int x = min(c * BX + t, w - 1);
        const uint32_t xn = (uint32_t)min(x + r + 1, w - 1), xo = (uint32_t)max(x - r, 0);
#pragma unroll
        for (int j = 0; j < BX; j++) {
            const uint32_t row = (uint32_t)min(y0 + j, h - 1) * uw;
            a[j] = in.get(row + xn, k);
            b[j] = in.get(row + xo, k);
        }
    };
    auto commit = [&](int c) {
#pragma unroll
        for (int j = 0; j < BX; j++) tile[c & 1][j][t] = a[j] - b[j];
    };
    // window sum at x = 0 of row y0 + t, in the reference's order
    float acc = 0.0f;
    {
        const uint32_t row = (uint32_t)min(y0 + t, h - 1) * uw;
        for (int i = -r; i <= r; i++) acc += in.get(row + (uint32_t)clampi(i, 0, w - 1), k);
    }
    issue(0);
    commit(0);
    __syncthreads();
    float* dst = out.p[k];
    for (int c = 0; c < ntiles; c++) {
        const bool more = c + 1 < ntiles;
        if (more) issue(c + 1);
        float* rowt = tile[c & 1][t];
#pragma unroll
        for (int j = 0; j < BX; j++) {
            const float d = rowt[j];
            rowt[j] = acc * norm;
            acc += d;
        }
        __syncthreads();
        const int x = c * BX + t;
        if (x < w) {
            const uint32_t obase = (uint32_t)y0 * uw + (uint32_t)x;
#pragma unroll
            for (int j = 0; j < BX; j++)
                if (y0 + j < h) dst[obase + (uint32_t)j * uw] = tile[c & 1][j][t];
        }
        if (more) commit(c + 1);
        __syncthreads();
    }
}

// Ring-buffered passes over plain planes.  The kernels above load every
// input twice (as in[x + r + 1] and as in[x - r]), and one wave both moves
// the data and walks the sums; a wave keeps at most ~63 memory operations in
// flight, which caps it far below the HBM rate.  These kernels read every
// input once, with direct global-to-LDS loads into a ring of recent columns
// (rows), issued several steps ahead by two loader waves; one walker wave
// walks the sums from the ring into an output tile and one storer wave
// writes the previous tile out.  The per-element arithmetic, and so the
// result, is the reference's.
constexpr int MOVERS = 3;
constexpr int MW_THREADS = BX * (MOVERS + 1);

// s_waitcnt immediates (gfx9 encoding: vmcnt in bits 3:0 and 15:14, expcnt
// in 6:4, lgkmcnt in 11:8); a field at its maximum does not wait.
constexpr int waitcnt_vm(int n) { return (n & 15) | ((n >> 4) << 14) | (7 << 4) | (15 << 8); }
constexpr int WAITCNT_LGKM0 = 15 | (3 << 14) | (7 << 4);

// Workgroup barrier without the memory fence of __syncthreads, which would
// make every wave wait for all of its loads into LDS.  Each wave waits with
// s_waitcnt for exactly what the others read after the barrier.
__device__ __forceinline__ void raw_barrier() {
    asm volatile("" ::: "memory");
    __builtin_amdgcn_s_barrier();
    asm volatile("" ::: "memory");
}

// Horizontal pass.  grid: (ceil(h/64), planes); block: 4 waves; walker
// lane t walks row y0 + t.  Input columns sit in a ring of RC columns, in
// 16-column blocks of 4 KB; each block is filled by four 16-byte-per-lane
// direct global-to-LDS loads, one per 16-row group q, whose lane i fetches
// row 16 q + i / 4, column group g = (i % 4) ^ ((i / 8) % 4) of the block
// (the XOR spreads a column's 64 rows over the LDS banks).  So unit
// (row t, group G = c / 4) -- columns 4G .. 4G + 3 of row t -- is at float
//   (G / 4 % (RC / 16)) * 1024 + (t / 16) * 256 + 16 (t % 16) + 4 ((G % 4) ^ ((t / 2) % 4)).
// Steps of HS = 32 columns: wave 0 walks them into an output tile; waves 1
// and 2 each load one block D steps ahead and wait until only the last
// D - 1 steps' loads are in flight; wave 3 stores the previous step's tile
// with 16-byte stores; a barrier ends the step.  Columns [0, A0) are loaded
// first (A0 = 32 D + r + 1 rounded up to 16); step s loads columns
// A0 + 32 s .. + 31.  Groups past the right edge load the last group, so a
// slot past the edge holds the column of the same residue mod 4.  A column
// the walker reads at step s was issued at step s - D or earlier, and
// columns landing during step s never evict one the walk still needs when
// RC > 32 D + 2 r + 47.
// Away from the edges the walker reads whole 16-byte units: the a-columns
// x + r + 1 and b-columns x - r of 4 consecutive x straddle two units at a
// fixed offset given by RM = r % 4 (a template parameter), so 9 + 9 unit
// reads serve 32 columns.  The output tile holds unit (group m, row t) at
// m * 64 + (t ^ m).  Needs w % 4 == 0 and 16-byte aligned planes.
// A block owns rb = 16 nq <= 64 rows (the band is chosen by the host to
// spread the rows x planes evenly over the CUs): only the first nq 16-row
// groups are loaded and stored; walker lanes past the band walk stale ring
// slots and are never stored.
constexpr int HS = 32;

// s_waitcnt until at most 2 nq loads are in flight (nq uniform, 1..4): the
// loaders issue nq loads per step and keep D - 1 = 2 steps in flight.
__device__ __forceinline__ void wait_two_steps(int nq) {
    switch (nq) {
        case 1: __builtin_amdgcn_s_waitcnt(waitcnt_vm(2)); break;
        case 2: __builtin_amdgcn_s_waitcnt(waitcnt_vm(4)); break;
        case 3: __builtin_amdgcn_s_waitcnt(waitcnt_vm(6)); break;
        default: __builtin_amdgcn_s_waitcnt(waitcnt_vm(8)); break;
    }
}

template <int I>
__device__ __forceinline__ float comp(const float4& v) {
    return I == 0 ? v.x : I == 1 ? v.y : I == 2 ? v.z : v.w;
}
// element I (0..7) of the pair (lo, hi)
template <int I>
__device__ __forceinline__ float pick(const float4& lo, const float4& hi) {
    if constexpr (I < 4)
        return comp<I>(lo);
    else
        return comp<I - 4>(hi);
}

template <int RC, int D, int RM>
__global__ __launch_bounds__(MW_THREADS) void k_boxh_ring(CBoxPlanes in, BoxPlanes out, int w, int h, int r, float norm, int nq) {
    static_assert(D == 3, "wait_two_steps keeps two steps of loads in flight");
    if (out.gate && *out.gate == 0) return;
    constexpr int NB = RC / 16;
    __shared__ float ring[RC * BX];
    __shared__ float4 otile[2 * (HS / 4) * BX];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int rb = 16 * nq;
    const int y0 = blockIdx.x * rb;
    const int k = blockIdx.y;
    const float* __restrict__ src = in.p[k];
    float* __restrict__ dst = out.p[k];
    const uint32_t uw = (uint32_t)w;
    // loader lane: row within the 16-row group and column group it fetches
    const int lr = lane >> 2, lg = (lane & 3) ^ ((lane >> 3) & 3);
    // issue one 16-row group q of the block holding columns [L, L + 16)
    auto load_group = [&](int L, int q) {
        const uint32_t row = (uint32_t)min(y0 + 16 * q + lr, h - 1);
        const uint32_t col = (uint32_t)min(L + 4 * lg, w - 4);
        float* lds = ring + ((L >> 4) & (NB - 1)) * 1024 + q * 256;
        __builtin_amdgcn_global_load_lds(src + row * uw + col, lds, 16, 0, 0);
    };
    // walker lane: its row's offset within a block and its group swizzle
    const int wbase = (lane >> 4) * 256 + 16 * (lane & 15), wx = (lane >> 1) & 3;
    auto get = [&](int c) { return ring[((c >> 4) & (NB - 1)) * 1024 + wbase + 4 * (((c >> 2) & 3) ^ wx) + (c & 3)]; };
    const int A0 = (HS * D + r + 1 + 15) & ~15;
    for (int it = wave; it < (A0 >> 4) * 4; it += MOVERS + 1)
        if ((it & 3) < nq) load_group(16 * (it >> 2), it & 3);
    __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    __syncthreads();
    float acc = 0.0f;
    if (wave == 0)
        for (int i = -r; i <= r; i++) acc += get(clampi(i, 0, w - 1));  // window at x = 0, the reference's order
    // unit reads of the fast walk: a-units start at group 8 s + qa, b-units at 8 s + qb
    constexpr int RA = (RM + 1) & 3;        // offset of the a-columns within their first unit
    constexpr int RB = (4 - RM) & 3;        // offset of the b-columns
    constexpr int NA = RA ? 9 : 8, NBU = RB ? 9 : 8;
    const int qa = (r + 1) >> 2;            // r + 1 = 4 qa + RA
    const int qb = -(r >> 2) - (RM ? 1 : 0);  // -r = 4 qb + RB
    int oa[4], ob[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        oa[j] = wbase + 4 * (((qa + j) & 3) ^ wx);
        ob[j] = wbase + 4 * (((qb + j) & 3) ^ wx);
    }
    auto unit = [&](int G, int off) { return *reinterpret_cast<const float4*>(&ring[((G >> 2) & (NB - 1)) * 1024 + off]); };
    // store the tile of step s: lane (row, g2) = (l / 8 + 8 i, l % 8) -> columns 32 s + 4 g2 .. + 3
    auto store_step = [&](int s) {
        const float4* ot = otile + (s & 1) * (HS / 4) * BX;
        const int g2 = lane & 7;
        const int c = HS * s + 4 * g2;
        float4 v[BX / 8];
#pragma unroll
        for (int i = 0; i < BX / 8; i++) {
            const int row = (lane >> 3) + 8 * i;
            v[i] = ot[g2 * BX + (row ^ g2)];
        }
#pragma unroll
        for (int i = 0; i < BX / 8; i++) {
            const int row = (lane >> 3) + 8 * i;
            if (c < w && row < rb && y0 + row < h) *reinterpret_cast<float4*>(dst + (uint32_t)(y0 + row) * uw + (uint32_t)c) = v[i];
        }
    };
    const int steps = (w + HS - 1) / HS;
    for (int s = 0; s < steps; s++) {
        if (wave == 0) {
            const int X0 = HS * s;
            float4* ot = otile + (s & 1) * (HS / 4) * BX;
            if (X0 >= r && X0 + HS + r <= w - 1) {
                // interior: no clamped column
                float4 A[NA], B[NBU];
#pragma unroll
                for (int j = 0; j < NA; j++) A[j] = unit(8 * s + qa + j, oa[j & 3]);
#pragma unroll
                for (int j = 0; j < NBU; j++) B[j] = unit(8 * s + qb + j, ob[j & 3]);
#pragma unroll
                for (int m = 0; m < HS / 4; m++) {
                    float o[4];
                    const float4 a0 = A[m], a1 = A[m + (RA ? 1 : 0)];
                    const float4 b0 = B[m], b1 = B[m + (RB ? 1 : 0)];
                    const float d0 = pick<RA + 0>(a0, a1) - pick<RB + 0>(b0, b1);
                    const float d1 = pick<RA + 1>(a0, a1) - pick<RB + 1>(b0, b1);
                    const float d2 = pick<RA + 2>(a0, a1) - pick<RB + 2>(b0, b1);
                    const float d3 = pick<RA + 3>(a0, a1) - pick<RB + 3>(b0, b1);
                    o[0] = acc * norm;
                    acc += d0;
                    o[1] = acc * norm;
                    acc += d1;
                    o[2] = acc * norm;
                    acc += d2;
                    o[3] = acc * norm;
                    acc += d3;
                    ot[m * BX + (lane ^ m)] = make_float4(o[0], o[1], o[2], o[3]);
                }
            } else {
                float* otf = reinterpret_cast<float*>(ot);
#pragma unroll 4
                for (int j = 0; j < HS; j++) {
                    const int x = X0 + j;
                    const float d = get(min(x + r + 1, w - 1)) - get(max(x - r, 0));
                    otf[((j >> 2) * BX + (lane ^ (j >> 2))) * 4 + (j & 3)] = acc * norm;
                    acc += d;
                }
            }
            __builtin_amdgcn_s_waitcnt(WAITCNT_LGKM0);
        } else if (wave <= 2) {
            const int L = A0 + HS * s + 16 * (wave - 1);
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (q < nq) load_group(L, q);
            wait_two_steps(nq);
        } else if (s > 0) {
            store_step(s - 1);
        }
        raw_barrier();
    }
    if (wave == MOVERS) store_step(steps - 1);
}

// Vertical pass.  grid: (ceil(w/64), planes); block: 4 waves; lane t of
// every wave stands for column x0 + t.  Row y of the block's 64 columns is
// held at ring[(y & (RING - 1)) * 64 + t].  Steps of VS = 32 rows:
//   wave 0 walks the step's rows from the ring into an output tile;
//   waves 1 and 2 issue direct global-to-LDS loads (one 256-byte row per
//     instruction) of the rows D steps ahead, then wait until only the last
//     D - 1 steps' loads are in flight;
//   wave 3 stores the previous step's output tile;
// and a barrier ends the step.  Rows [0, A0 = 32 D + r + 1) are loaded
// before the first step; step s loads rows A0 + 32 s .. + 31 (rows past the
// bottom read the last row).  A row the walker reads at step s was issued at
// step s - D or earlier, and rows landing during step s never evict a row
// the walk still needs when RING > 32 D + 2 r + 32.
constexpr int VS = 32;
// A 16-byte direct global-to-LDS load, outside the kernel template: in the
// kernel body the host pass drops the kernel's launch stub without a
// diagnostic when the size is 16 (a gfx950-only operand).
__device__ __forceinline__ void lds_load16(const float* g, float* lds) { __builtin_amdgcn_global_load_lds(g, lds, 16, 0, 0); }
// V16: rows are read as 16-byte units (w % 4 == 0, planes 16-byte aligned),
// one load covering 4 rows x 64 columns, lane l taking row l / 16 and
// columns 4 (l % 16) .. + 3, so the units land in the ring's row-major slots
// as they are; the prefetch start is rounded up to a whole 4-row group.
// A block owns cb <= 64 columns (cb % 4 == 0 with V16; the band is chosen
// by the host to spread the columns x planes evenly over the CUs); lanes
// past the band shadow its last column and are never stored.
template <int RING, int D, bool V16>
__global__ __launch_bounds__(MW_THREADS) void k_boxv_ring(CBoxPlanes in, BoxPlanes out, int w, int h, int r, float norm, int cb) {
    if (out.gate && *out.gate == 0) return;
    __shared__ float ring[RING * BX];
    __shared__ float otile[2 * VS * BX];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int x0 = blockIdx.x * cb;
    const int x = min(x0 + min(lane, cb - 1), w - 1);  // lanes past the band or the edge shadow its last column
    const bool live = lane < cb && x0 + lane < w;
    const int k = blockIdx.y;
    const float* __restrict__ src = in.p[k];
    float* __restrict__ dst = out.p[k];
    const uint32_t ux = (uint32_t)x;
    auto slot = [&](int y) -> float& { return ring[(y & (RING - 1)) * BX + lane]; };
    auto row_in = [&](int y) { return src + (size_t)y * (size_t)w; };  // uniform row base
    const int A0 = V16 ? (VS * D + r + 1 + 3) & ~3 : VS * D + r + 1;
    // 16-byte loader lane: row within its 4-row group, first column of its unit
    const int qrow = lane >> 4;
    const int qu = min(lane & 15, cb / 4 - 1);  // 16-byte unit of the band
    const uint32_t qcol = (uint32_t)min(x0 + 4 * qu, w - 4);
    auto load_quad = [&](int c) {  // rows c .. c + 3 (c % 4 == 0), clamped per lane
        lds_load16(row_in(min(c + qrow, h - 1)) + qcol, ring + (c & (RING - 1)) * BX);
    };
    if (V16) {
        for (int g = wave; 4 * g < A0; g += MOVERS + 1) load_quad(4 * g);
    } else {
        for (int y = wave; y < min(A0, h); y += MOVERS + 1) __builtin_amdgcn_global_load_lds(row_in(y) + ux, ring + (y & (RING - 1)) * BX, 4, 0, 0);
    }
    __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    __syncthreads();
    float acc = 0.0f;
    if (wave == 0)
        for (int i = -r; i <= r; i++) acc += slot(clampi(i, 0, h - 1));  // window at y = 0, the reference's order
    auto store_step = [&](int s) {
        const float* ot = otile + (s & 1) * VS * BX;
        const int Y0 = VS * s;
        if (V16) {
            // lane l: rows l / 16 + 4 i, columns 4 (l % 16) .. + 3 as one unit
            float4 u[VS / 4];
#pragma unroll
            for (int i = 0; i < VS / 4; i++) u[i] = *reinterpret_cast<const float4*>(ot + (qrow + 4 * i) * BX + 4 * (lane & 15));
            if (4 * (lane & 15) >= cb || x0 + 4 * (lane & 15) >= w) return;
#pragma unroll
            for (int i = 0; i < VS / 4; i++)
                if (Y0 + qrow + 4 * i < h) *reinterpret_cast<float4*>(dst + (size_t)(Y0 + qrow + 4 * i) * (size_t)w + qcol) = u[i];
            return;
        }
        float v[VS];
#pragma unroll
        for (int j = 0; j < VS; j++) v[j] = ot[j * BX + lane];
        if (!live) return;
        if (Y0 + VS <= h) {
#pragma unroll
            for (int j = 0; j < VS; j++) (dst + (size_t)(Y0 + j) * (size_t)w)[ux] = v[j];
        } else {
            for (int j = 0; j < VS; j++)
                if (Y0 + j < h) (dst + (size_t)(Y0 + j) * (size_t)w)[ux] = v[j];
        }
    };
    const int steps = (h + VS - 1) / VS;
    for (int s = 0; s < steps; s++) {
        if (wave == 0) {
            const int Y0 = VS * s;
            float d[VS];
            const int sa = (Y0 + r + 1) & (RING - 1), sb = (Y0 - r) & (RING - 1);
            if (Y0 >= r && Y0 + VS + r <= h - 1 && sa <= RING - VS && sb <= RING - VS) {
                // no clamped row and no wrap: fixed offsets from two bases
                const float* pa = ring + sa * BX + lane;
                const float* pb = ring + sb * BX + lane;
#pragma unroll
                for (int j = 0; j < VS; j++) d[j] = pa[j * BX] - pb[j * BX];
            } else {
#pragma unroll
                for (int j = 0; j < VS; j++) d[j] = slot(min(Y0 + j + r + 1, h - 1)) - slot(max(Y0 + j - r, 0));
            }
            float* ot = otile + (s & 1) * VS * BX;
#pragma unroll
            for (int j = 0; j < VS; j++) {
                ot[j * BX + lane] = acc * norm;
                acc += d[j];
            }
            __builtin_amdgcn_s_waitcnt(WAITCNT_LGKM0);
        } else if (wave <= 2) {
            const int base = A0 + VS * s + (VS / 2) * (wave - 1);
            if (V16) {
#pragma unroll
                for (int i = 0; i < VS / 8; i++) load_quad(base + 4 * i);
                __builtin_amdgcn_s_waitcnt(waitcnt_vm((VS / 8) * (D - 1)));
            } else {
#pragma unroll
                for (int i = 0; i < VS / 2; i++) {
                    const int c = base + i;
                    __builtin_amdgcn_global_load_lds(row_in(min(c, h - 1)) + ux, ring + (c & (RING - 1)) * BX, 4, 0, 0);
                }
                __builtin_amdgcn_s_waitcnt(waitcnt_vm((VS / 2) * (D - 1)));
            }
        } else if (s > 0) {
            store_step(s - 1);
        }
        raw_barrier();
    }
    if (wave == MOVERS) store_step(steps - 1);
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

// Compute units of the current device (cached per device).
int device_cus() {
    static std::atomic<int> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    int n = cache[dev].load(std::memory_order_relaxed);
    if (n == 0) {
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0) n = 256;
        cache[dev].store(n, std::memory_order_relaxed);
    }
    return n;
}

// Band (rows or columns per block, a multiple of `step`, 32..64) of a
// ring pass over `extent` rows or columns of `planes` planes.  A CU moves
// data at about the same rate with one block or two, so a pass lasts as
// long as its busiest CU: ceil(blocks / CUs) bands.  The smallest such
// load wins, ties going to the wider band (fewer blocks).
int ring_band(size_t extent, int planes, int step) {
    const size_t cus = (size_t)device_cus();
    int best = BX;
    size_t best_load = ~(size_t)0;
    for (int b = BX; b >= 32; b -= step) {
        const size_t blocks = (extent + b - 1) / b * (size_t)planes;
        const size_t load = (blocks + cus - 1) / cus * (size_t)b;
        if (load < best_load) {
            best_load = load;
            best = b;
        }
    }
    return best;
}

// Any radius: one thread per row / column reading global memory directly.
template <class In>
__global__ void k_boxx_h_any(In in, BoxPlanes out, int w, int h, int r, float norm, int planes) {
    if (out.gate && *out.gate == 0) return;
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (y >= h || k >= planes) return;
    const size_t row = (size_t)y * (size_t)w;
    float acc = 0.0f;
    for (int i = -r; i <= r; i++) acc += in.get(row + (size_t)clampi(i, 0, w - 1), k);
    for (int x = 0; x < w; x++) {
        out.p[k][row + (size_t)x] = acc * norm;
        acc += in.get(row + (size_t)min(x + r + 1, w - 1), k) - in.get(row + (size_t)max(x - r, 0), k);
    }
}

__global__ void k_boxx_v_any(CBoxPlanes in, BoxPlanes out, int w, int h, int r, float norm) {
    if (out.gate && *out.gate == 0) return;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (x >= w) return;
    const float* src = in.p[k];
    float acc = 0.0f;
    for (int i = -r; i <= r; i++) acc += src[(size_t)clampi(i, 0, h - 1) * (size_t)w + (size_t)x];
    for (int y = 0; y < h; y++) {
        out.p[k][(size_t)y * (size_t)w + (size_t)x] = acc * norm;
        acc += src[(size_t)min(y + r + 1, h - 1) * (size_t)w + (size_t)x] - src[(size_t)max(y - r, 0) * (size_t)w + (size_t)x];
    }
}

template <class In>
hipError_t run_h(In in, BoxPlanes out, int planes, size_t w, size_t h, size_t r, hipStream_t st) {
    const float norm = 1.0f / (float)(2 * r + 1);  // image.rs:274
    if (r < BX) {
        dim3 grid((unsigned)((h + BX - 1) / BX), (unsigned)planes);
        hipLaunchKernelGGL(k_boxx_h<In>, grid, dim3(BX), 0, st, in, out, (int)w, (int)h, (int)r, norm);
    } else {
        dim3 grid((unsigned)((h + 63) / 64), (unsigned)planes);
        hipLaunchKernelGGL(k_boxx_h_any<In>, grid, dim3(64), 0, st, in, out, (int)w, (int)h, (int)r, norm, planes);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t box_h_planes(CBoxPlanes in, BoxPlanes out, int planes, size_t w, size_t h, size_t r, hipStream_t st) {
    bool ring = w % 4 == 0 && r <= 184;
    for (int k = 0; k < planes && ring; k++) ring = aligned16(in.p[k]) && aligned16(out.p[k]);
    if (!ring) return run_h(BoxInPlanes{in}, out, planes, w, h, r, st);
    const float norm = 1.0f / (float)(2 * r + 1);  // image.rs:274
    // ring sizes per the bound RC > 32 D + 2 r + 47 with D = 3
    const int rb = ring_band(h, planes, 16);
    dim3 grid((unsigned)((h + rb - 1) / rb), (unsigned)planes);
    const int ri = (int)r, nq = rb / 16;
#define DS_BOXH(RC, RM) hipLaunchKernelGGL((k_boxh_ring<RC, 3, RM>), grid, dim3(MW_THREADS), 0, st, in, out, (int)w, (int)h, ri, norm, nq)
#define DS_BOXH_RM(RC)              \
    switch (ri & 3) {               \
        case 0: DS_BOXH(RC, 0); break; \
        case 1: DS_BOXH(RC, 1); break; \
        case 2: DS_BOXH(RC, 2); break; \
        default: DS_BOXH(RC, 3); break; \
    }
    if (r <= 56) {
        DS_BOXH_RM(256)
    } else {
        DS_BOXH_RM(512)
    }
#undef DS_BOXH_RM
#undef DS_BOXH
    return hipGetLastError();
}

hipError_t box_v_planes(CBoxPlanes in, BoxPlanes out, int planes, size_t w, size_t h, size_t r, hipStream_t st) {
    const float norm = 1.0f / (float)(2 * r + 1);
    bool v16 = w % 4 == 0;
    for (int k = 0; k < planes && v16; k++) v16 = aligned16(in.p[k]) && aligned16(out.p[k]);
    const int cb = r <= 191 ? ring_band(w, planes, 4) : BX;
    const dim3 grid((unsigned)((w + cb - 1) / cb), (unsigned)planes);
    // ring sizes from the schedule's bound RING > 32 D + 2 r + 32 with D = 3
    // (+ 3 when the prefetch start is rounded to a 4-row group)
#define DS_BOXV(RING, V) hipLaunchKernelGGL((k_boxv_ring<RING, 3, V>), grid, dim3(MW_THREADS), 0, st, in, out, (int)w, (int)h, (int)r, norm, cb)
    if (v16 && r <= 60)
        DS_BOXV(256, true);
    else if (v16 && r <= 188)
        DS_BOXV(512, true);
    else if (r <= 63)
        DS_BOXV(256, false);
    else if (r <= 191)
        DS_BOXV(512, false);
    else
        hipLaunchKernelGGL(k_boxx_v_any, grid, dim3(64), 0, st, in, out, (int)w, (int)h, (int)r, norm);
#undef DS_BOXV
    return hipGetLastError();
}

// box_blur(in, r) -> out (image.rs:269); tmp: w*h floats.  r == 0 is a copy.
hipError_t launch_box_blur(const float* in, float* tmp, float* out, size_t w, size_t h, size_t r, hipStream_t st) {
    if (w == 0 || h == 0) return hipSuccess;
    if (r == 0) return hipMemcpyAsync(out, in, w * h * sizeof(float), hipMemcpyDeviceToDevice, st);
    CBoxPlanes ci{};
    ci.p[0] = in;
    BoxPlanes t{};
    t.p[0] = tmp;
    hipError_t e = box_h_planes(ci, t, 1, w, h, r, st);
    if (e != hipSuccess) return e;
    CBoxPlanes ct{};
    ct.p[0] = tmp;
    BoxPlanes o{};
    o.p[0] = out;
    return box_v_planes(ct, o, 1, w, h, r, st);
}

}  // namespace ds
