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
// plane and walks them in 64-element chunks.  Chunks are staged through an
// LDS ring with coalesced loads (three chunk slots: the window spans the
// previous, current and next chunk for r < 64); the next-but-one chunk is
// fetched into registers while the current one is walked.  Within a chunk
// the window differences are gathered first, so the serial part is one
// dependent add per element.  Rows/columns x planes give the parallelism:
// ~(h/64) * planes waves for the horizontal pass and ~(w/64) * planes waves
// for the vertical one.

#include "box.h"
#include "common.h"

#pragma clang fp contract(off)

namespace ds {

namespace {

constexpr int BX = 64;       // chunk length and rows / columns per wave
constexpr int RING = 3 * BX;  // three resident chunks: previous, current, next

// Horizontal pass.  grid: (ceil(h/64), planes); block: one wave.  Lane t owns
// row y0 + t.  Per 64-column chunk c (chunks c-1, c, c+1 resident in the
// ring, x held at ring position x mod 192):
//   1. gather d[j] = in[x+r+1] - in[x-r] for the chunk into registers;
//   2. run the 64 dependent additions acc += d[j], keeping acc * norm;
//   3. stage the outputs in the dead slot of chunk c-1 and store them
//      coalesced, then refill that slot with chunk c+2, fetched into
//      registers before step 1.
template <class In>
__global__ __launch_bounds__(BX) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_boxx_h(In in, BoxPlanes out, int w, int h, int r, float norm) {
    __shared__ float ring[BX][RING + 1];  // odd row stride: lanes of one column hit distinct banks
    const int t = threadIdx.x;
    const int y0 = blockIdx.x * BX;
    const int k = blockIdx.y;
    const size_t myrow = (size_t)min(y0 + t, h - 1) * (size_t)w;
    const int ntiles = (w + BX - 1) / BX;
    float pre[BX];
    auto fetch = [&](int c) {
        const int xx = c * BX + t;
        const uint32_t base = (uint32_t)y0 * (uint32_t)w + (uint32_t)xx;
#pragma unroll
        for (int j = 0; j < BX; j++) {
            const int yy = y0 + j;
            pre[j] = (yy < h && xx < w) ? in.get(base + (uint32_t)j * (uint32_t)w, k) : 0.0f;
        }
    };
    auto commit = [&](int c) {
        const int base = (c % 3) * BX + t;
#pragma unroll
        for (int j = 0; j < BX; j++) ring[j][base] = pre[j];
    };
    fetch(0);
    commit(0);
    if (ntiles > 1) {
        fetch(1);
        commit(1);
    }
    // window sum at x = 0 in the reference's order
    float acc = 0.0f;
    for (int i = -r; i <= r; i++) acc += in.get(myrow + (size_t)clampi(i, 0, w - 1), k);
    __syncthreads();
    const float* row = ring[t];
    for (int c = 0; c < ntiles; c++) {
        if (c + 2 < ntiles) fetch(c + 2);
        const int xb = c * BX;
        float d[BX];
#pragma unroll
        for (int j = 0; j < BX; j++) {
            const int x = xb + j;
            const int xn = min(x + r + 1, w - 1), xo = max(x - r, 0);
            d[j] = row[xn % RING] - row[xo % RING];
        }
#pragma unroll
        for (int j = 0; j < BX; j++) {
            const float dj = d[j];
            d[j] = acc * norm;
            acc += dj;
        }
        // chunk c-1 is no longer read: stage the outputs in its slot
        const int stage = ((c + 2) % 3) * BX;
#pragma unroll
        for (int j = 0; j < BX; j++) ring[t][stage + j] = d[j];
        __syncthreads();
        const int xx = xb + t;
        const uint32_t obase = (uint32_t)y0 * (uint32_t)w + (uint32_t)xx;
        float* dst = out.p[k];
        if (xx < w)
#pragma unroll 16
            for (int j = 0; j < BX; j++) {
                const int yy = y0 + j;
                if (yy < h) dst[obase + (uint32_t)j * (uint32_t)w] = ring[j][stage + t];
            }
        __syncthreads();
        if (c + 2 < ntiles) commit(c + 2);
        __syncthreads();
    }
}

// Vertical pass.  grid: (ceil(w/64), planes); block: one wave.  Lane t walks
// column x0 + t; it only touches its own LDS column (ring[row][t]), so no
// barriers are needed.  Same chunk scheme as the horizontal pass, with the
// outputs stored straight to global memory (a wave's stores cover one
// contiguous row piece).
__global__ __launch_bounds__(BX) __attribute__((amdgpu_waves_per_eu(1, 2))) void k_boxx_v(CBoxPlanes in, BoxPlanes out, int w, int h, int r, float norm) {
    __shared__ float ring[RING][BX];
    const int t = threadIdx.x;
    const int x = min(blockIdx.x * BX + t, w - 1);  // lanes past the edge shadow the last column
    const bool live = blockIdx.x * BX + t < w;
    const int k = blockIdx.y;
    const float* __restrict__ src = in.p[k];
    float* __restrict__ dst = out.p[k];
    const int ntiles = (h + BX - 1) / BX;
    const uint32_t uw = (uint32_t)w;
    float pre[BX];
    // chunk 0 and 1 into the ring; 32-bit element offsets off a uniform base
    // (planes stay below 2^31 elements): one VGPR per in-flight address
    for (int c = 0; c < 2 && c < ntiles; c++) {
        const uint32_t base = (uint32_t)(c * BX) * uw + (uint32_t)x;
#pragma unroll
        for (int j = 0; j < BX; j++) pre[j] = c * BX + j < h ? src[base + (uint32_t)j * uw] : 0.0f;
#pragma unroll
        for (int j = 0; j < BX; j++) ring[c * BX + j][t] = pre[j];
    }
    float acc = 0.0f;
    for (int i = -r; i <= r; i++) acc += src[(uint32_t)clampi(i, 0, h - 1) * uw + (uint32_t)x];
    for (int c = 0; c < ntiles; c++) {
        const bool more = c + 2 < ntiles;
        if (more) {
            const uint32_t base = (uint32_t)((c + 2) * BX) * uw + (uint32_t)x;
#pragma unroll
            for (int j = 0; j < BX; j++) pre[j] = (c + 2) * BX + j < h ? src[base + (uint32_t)j * uw] : 0.0f;
        }
        const int yb = c * BX;
        float d[BX];
#pragma unroll
        for (int j = 0; j < BX; j++) {
            const int y = yb + j;
            const int yn = min(y + r + 1, h - 1), yo = max(y - r, 0);
            d[j] = ring[yn % RING][t] - ring[yo % RING][t];
        }
#pragma unroll
        for (int j = 0; j < BX; j++) {
            const float dj = d[j];
            d[j] = acc * norm;
            acc += dj;
        }
        if (live) {
            const uint32_t obase = (uint32_t)yb * uw + (uint32_t)x;
#pragma unroll
            for (int j = 0; j < BX; j++)
                if (yb + j < h) dst[obase + (uint32_t)j * uw] = d[j];
        }
        if (more) {
            const int slot = ((c + 2) % 3) * BX;
#pragma unroll
            for (int j = 0; j < BX; j++) ring[slot + j][t] = pre[j];
        }
    }
}

// Any radius: one thread per row / column reading global memory directly.
template <class In>
__global__ void k_boxx_h_any(In in, BoxPlanes out, int w, int h, int r, float norm, int planes) {
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
