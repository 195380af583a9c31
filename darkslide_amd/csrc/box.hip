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
// plane and walks them in 64-element chunks.  The window differences
// in[x+r+1] - in[x-r] are formed with coalesced loads before the walk (each
// is exactly the reference's `at(x + r + 1) - at(x - r)`), so the serial part
// is one dependent add per element; the next chunk's loads are in flight
// while the current chunk is walked.  Rows/columns x planes give the
// parallelism: ~(h/64) * planes waves for the horizontal pass and
// ~(w/64) * planes waves for the vertical one.

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

// Vertical pass.  grid: (ceil(w/64), planes); block: one wave.  Lane t walks
// column x0 + t; its window differences come from two coalesced row loads,
// so the pass needs no LDS: differences and outputs stay in registers and
// the stores of a wave cover one contiguous row piece.
__global__ __launch_bounds__(BX) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_boxx_v(CBoxPlanes in, BoxPlanes out, int w, int h, int r,
                                                                                          float norm) {
    if (out.gate && *out.gate == 0) return;
    const int t = threadIdx.x;
    const int x = min(blockIdx.x * BX + t, w - 1);  // lanes past the edge shadow the last column
    const bool live = blockIdx.x * BX + t < w;
    const int k = blockIdx.y;
    const float* __restrict__ src = in.p[k];
    float* __restrict__ dst = out.p[k];
    const int ntiles = (h + BX - 1) / BX;
    const uint32_t uw = (uint32_t)w;
    float a[BX], b[BX];
    auto issue = [&](int c) {
#pragma unroll
        for (int j = 0; j < BX; j++) {
            const int y = c * BX + j;
            a[j] = src[(uint32_t)min(y + r + 1, h - 1) * uw + (uint32_t)x];
            b[j] = src[(uint32_t)max(y - r, 0) * uw + (uint32_t)x];
        }
    };
    // window sum at y = 0 in the reference's order
    float acc = 0.0f;
    for (int i = -r; i <= r; i++) acc += src[(uint32_t)clampi(i, 0, h - 1) * uw + (uint32_t)x];
    issue(0);
    for (int c = 0; c < ntiles; c++) {
        float d[BX];
#pragma unroll
        for (int j = 0; j < BX; j++) d[j] = a[j] - b[j];
        if (c + 1 < ntiles) issue(c + 1);
        const int yb = c * BX;
        const uint32_t obase = (uint32_t)yb * uw + (uint32_t)x;
#pragma unroll
        for (int j = 0; j < BX; j++) {
            const float o = acc * norm;
            acc += d[j];
            if (live && yb + j < h) dst[obase + (uint32_t)j * uw] = o;
        }
    }
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
    return run_h(BoxInPlanes{in}, out, planes, w, h, r, st);
}

hipError_t box_h_guided(BoxInGuided in, BoxPlanes out, size_t w, size_t h, size_t r, hipStream_t st) {
    return run_h(in, out, 6, w, h, r, st);
}

hipError_t box_h_guide1(BoxInGuide1 in, BoxPlanes out, size_t w, size_t h, size_t r, hipStream_t st) {
    return run_h(in, out, 4, w, h, r, st);
}

hipError_t box_v_planes(CBoxPlanes in, BoxPlanes out, int planes, size_t w, size_t h, size_t r, hipStream_t st) {
    const float norm = 1.0f / (float)(2 * r + 1);
    if (r < BX) {
        dim3 grid((unsigned)((w + BX - 1) / BX), (unsigned)planes);
        hipLaunchKernelGGL(k_boxx_v, grid, dim3(BX), 0, st, in, out, (int)w, (int)h, (int)r, norm);
    } else {
        dim3 grid((unsigned)((w + 63) / 64), (unsigned)planes);
        hipLaunchKernelGGL(k_boxx_v_any, grid, dim3(64), 0, st, in, out, (int)w, (int)h, (int)r, norm);
    }
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
