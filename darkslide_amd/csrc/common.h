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
    __host__ __device__ __forceinline__ int at(long x, long y) const { return c[((y & 1) << 1) | (x & 1)]; }
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

// Tile of a 2-D grid, renumbered so that each of the 8 XCDs (which take
// workgroups round-robin by linear id, each with its own L2) walks one
// contiguous run of tiles in row-major order.  Neighbouring tiles then share
// an L2, so the halos they read twice come from L2 instead of HBM.  A
// bijection of the grid: XCD x owns tiles [x q + min(x, m), + q + (x < m)),
// q = n / 8, m = n % 8.
__device__ __forceinline__ void xcd_tile(int& tx, int& ty) {
    const int nx = (int)gridDim.x, n = nx * (int)gridDim.y;
    const int b = (int)blockIdx.y * nx + (int)blockIdx.x;
    const int x = b & 7, k = b >> 3, q = n >> 3, m = n & 7;
    const int t = x * q + (x < m ? x : m) + k;
    tx = t % nx;
    ty = t / nx;
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
