// Shared helpers for the gfx950 kernels and the C-ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace ds {

// Bayer colour at (x, y) for a 2x2 pattern stored as pat[(y&1)*2 + (x&1)]
// (0 = R, 1 = G, 2 = B), as CfaPattern::color_at does for a 2x2 pattern
// (decode.rs:26-28).
struct Bayer {
    uint8_t c[4];
    // the four entries packed in a word and shifted out, so that a runtime
    // site index never makes the compiler keep the array in scratch memory
    __host__ __device__ __forceinline__ int at(long x, long y) const {
        const uint32_t u = (uint32_t)c[0] | ((uint32_t)c[1] << 8) | ((uint32_t)c[2] << 16) | ((uint32_t)c[3] << 24);
        return (int)((u >> (8 * (((y & 1) << 1) | (x & 1)))) & 0xffu);
    }
    // parity (x + y) & 1 of the non-green sites; Bayer greens form a checkerboard
    __host__ __device__ __forceinline__ int rb_parity() const { return c[0] == 1 ? 1 : 0; }
};

__host__ __device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// f32::max / f32::clamp semantics (fmaxf returns the non-NaN operand)
__device__ __forceinline__ float fmaxr(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ float clampr(float v, float lo, float hi) {
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}

// Oklab matrices (from / to linear ProPhoto), built on the host exactly as
// color.rs:382-418 builds them, and passed to kernels by value.
struct OklabM {
    float to_lms[3][3];
    float from_lms[3][3];
    float m2[3][3];
    float m2_inv[3][3];
};

__host__ __device__ __forceinline__ void mul_mv(const float m[3][3], const float v[3], float o[3]) {
    o[0] = m[0][0] * v[0] + m[0][1] * v[1] + m[0][2] * v[2];
    o[1] = m[1][0] * v[0] + m[1][1] * v[1] + m[1][2] * v[2];
    o[2] = m[2][0] * v[0] + m[2][1] * v[1] + m[2][2] * v[2];
}

// Stages N elements into LDS with NT threads: every thread issues all of its
// global loads before its first LDS store, so a block pays one memory latency
// for the whole tile instead of one per loop trip.  src(i) returns element i
// (the caller clamps coordinates), dst(i, v) stores it.
template <int N, int NT, class Src, class Dst>
__device__ __forceinline__ void stage_tile(Src src, Dst dst) {
    constexpr int K = (N + NT - 1) / NT;
    float r[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int i = (int)threadIdx.x + k * NT;
        r[k] = (i < N) ? src(i) : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int i = (int)threadIdx.x + k * NT;
        if (i < N) dst(i, r[k]);
    }
}

// Tile orders of 1-D grids over an nx x ny tile grid whose frame-edge tiles
// (halo leaving the image: the kernels' clamped paths, about twice the
// interior path's cost) lie outside the interior rectangle
// [ix0, ix1) x [iy0, iy1).  Dispatched last, the edge row at the bottom of
// the frame stretches a kernel's final round of tiles; these orders put the
// edge tiles first.
struct TileOrder {
    int nx, ny, ix0, ix1, iy0, iy1;  // empty interior: all four zero
};
// Interior tiles of TW x TH tiles over a w x h image with halos hx, hy:
// ox >= hx, oy >= hy, ox + TW + hx <= w, oy + TH + hy <= h.
inline TileOrder tile_order(int w, int h, int TW, int TH, int hx, int hy) {
    TileOrder o;
    o.nx = (w + TW - 1) / TW;
    o.ny = (h + TH - 1) / TH;
    o.ix0 = (hx + TW - 1) / TW;
    o.iy0 = (hy + TH - 1) / TH;
    o.ix1 = w >= TW + hx ? (w - TW - hx) / TW + 1 : 0;
    o.iy1 = h >= TH + hy ? (h - TH - hy) / TH + 1 : 0;
    if (o.ix1 <= o.ix0 || o.iy1 <= o.iy0) o.ix0 = o.ix1 = o.iy0 = o.iy1 = 0;
    return o;
}
// The j-th edge tile in row-major order.
__device__ __forceinline__ void edge_tile(int j, const TileOrder& o, int& tx, int& ty) {
    const int inx = o.ix1 - o.ix0, iny = o.iy1 - o.iy0;
    const int top = o.iy0 * o.nx;
    if (j < top) {
        tx = j % o.nx;
        ty = j / o.nx;
        return;
    }
    j -= top;
    const int m = o.nx - inx;  // edge tiles of a row crossing the interior
    if (j < iny * m) {
        const int c = j % m;
        ty = o.iy0 + j / m;
        tx = c < o.ix0 ? c : o.ix1 + (c - o.ix0);
        return;
    }
    j -= iny * m;
    tx = j % o.nx;
    ty = o.iy1 + j / o.nx;
}
// Block b: the edge tiles in row-major order, then the interior ones.
__device__ __forceinline__ void tile_edge_first(int b, const TileOrder& o, int& tx, int& ty) {
    const int inx = o.ix1 - o.ix0, iny = o.iy1 - o.iy0;
    const int ne = o.nx * o.ny - inx * iny;
    if (b < ne) {
        edge_tile(b, o, tx, ty);
        return;
    }
    b -= ne;
    tx = o.ix0 + b % inx;
    ty = o.iy0 + b / inx;
}
// Block b as the 8 XCDs take it (round-robin by linear id, each XCD with
// its own L2): XCD x first takes every 8th edge tile (x, x + 8, ...), then
// one contiguous row-major run of interior tiles, so neighbouring interior
// tiles share an L2 and the halos they read twice come from it rather than
// HBM.  A bijection of the tile grid.
__device__ __forceinline__ void tile_edge_first_xcd(int b, const TileOrder& o, int& tx, int& ty) {
    const int inx = o.ix1 - o.ix0, iny = o.iy1 - o.iy0;
    const int n = o.nx * o.ny, ne = n - inx * iny;
    const int x = b & 7, k = b >> 3;
    const int q = n >> 3, m = n & 7, qe = ne >> 3, me = ne & 7;
    const int ne_x = qe + (x < me ? 1 : 0);
    if (k < ne_x) {
        edge_tile(8 * k + x, o, tx, ty);
        return;
    }
    // interior items of XCDs 0 .. x-1: their blocks minus their edge tiles
    const int i = x * (q - qe) + (x < m ? x : m) - (x < me ? x : me) + (k - ne_x);
    tx = o.ix0 + i % inx;
    ty = o.iy0 + i / inx;
}

// Runtime-sized variant: batches of KB loads per thread.
template <int KB, int NT, class Src, class Dst>
__device__ __forceinline__ void stage_tile_n(int n, Src src, Dst dst) {
    for (int base = 0; base < n; base += KB * NT) {
        float r[KB];
#pragma unroll
        for (int k = 0; k < KB; k++) {
            const int i = base + (int)threadIdx.x + k * NT;
            r[k] = (i < n) ? src(i) : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < KB; k++) {
            const int i = base + (int)threadIdx.x + k * NT;
            if (i < n) dst(i, r[k]);
        }
    }
}

}  // namespace ds
