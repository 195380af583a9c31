// NLM in the reference's own summation order (enhance.rs:34-72), for
// callers that need the CPU path's bits rather than the fastest kernel.
//
// The reference splits the (2S+1)^2 offsets (dy-major) into chunks of
// ceil(n / T) for a rayon pool of T threads; each chunk accumulates its own
// (acc, wsum) planes offset by offset, and the chunk planes are then added
// in chunk order into zero-initialised totals.  Per offset it forms the
// squared differences, box-blurs them with running sums walked along each
// whole row, then down each whole column (image.rs:268-313), and adds
// exp(-(dist / h2)) (glibc expf, restated in libm_ref.h) times the shifted
// sample.  Every chunk here runs in parallel with the others, as rayon's do,
// and within a chunk the offsets run in order:
//   k_nlmref_rows  one thread per (chunk, row): the squared differences of
//                  the chunk's current offset, formed on the fly, and the
//                  row running sum -> tmp (one plane per chunk)
//   k_nlmref_cols  one thread per (chunk, column): the column running sum
//                  of tmp, the weight, and the chunk's acc / wsum update
//   k_nlmref_final the chunk planes added in chunk order, acc / max(wsum, 1e-12)
// The running sums are sequential by definition, so this path does the
// reference's work at the reference's granularity: it is a parity mode, two
// launches per offset step, not the throughput path of nlm.hip.

#include "common.h"
#include "libm_ref.h"
#include "nlm.h"
#include "nlm_ref.h"

#pragma clang fp contract(off)

namespace ds {

namespace {

struct RefGeom {
    int w, h, S, R;  // R: patch radius
    int chunk, nchunks, n_off;
    float norm;      // 1 / (2R + 1)
};

__device__ __forceinline__ float at_clamped(const float* __restrict__ p, int w, int h, int x, int y) {
    return p[(size_t)clampi(y, 0, h - 1) * (size_t)w + (size_t)clampi(x, 0, w - 1)];
}

// offset o of the dy-major list: (dx, dy)
__device__ __forceinline__ void offset_of(int o, int S, int& dx, int& dy) {
    const int side = 2 * S + 1;
    dy = o / side - S;
    dx = o % side - S;
}

__global__ __launch_bounds__(64) void k_nlmref_rows(const float* __restrict__ p, RefGeom g, int step, float* __restrict__ tmp) {
    const int y = blockIdx.x * 64 + threadIdx.x, c = blockIdx.y;
    const int o = c * g.chunk + step;
    if (y >= g.h || step >= g.chunk || o >= g.n_off) return;
    int dx, dy;
    offset_of(o, g.S, dx, dy);
    const float* row = p + (size_t)y * (size_t)g.w;
    // squared difference at column clamp(i) (box_blur's at(), image.rs:281)
    auto diff = [&](int i) {
        const int x = clampi(i, 0, g.w - 1);
        const float d = row[x] - at_clamped(p, g.w, g.h, x + dx, y + dy);
        return d * d;
    };
    float* out = tmp + (size_t)c * (size_t)g.w * (size_t)g.h + (size_t)y * (size_t)g.w;
    float acc = 0.0f;
    for (int i = -g.R; i <= g.R; i++) acc += diff(i);
    for (int x = 0; x < g.w; x++) {
        out[x] = acc * g.norm;
        acc += diff(x + g.R + 1) - diff(x - g.R);
    }
}

__global__ __launch_bounds__(64) void k_nlmref_cols(const float* __restrict__ p, RefGeom g, int step, NlmWeight wt_par,
                                                    const float* __restrict__ tmp, float* __restrict__ acc_planes,
                                                    float* __restrict__ wsum_planes) {
    const int x = blockIdx.x * 64 + threadIdx.x, c = blockIdx.y;
    const int o = c * g.chunk + step;
    if (x >= g.w || step >= g.chunk || o >= g.n_off) return;
    int dx, dy;
    offset_of(o, g.S, dx, dy);
    const size_t n = (size_t)g.w * (size_t)g.h;
    const float h2 = wt_par.h2();
    float* a = acc_planes + (size_t)c * n;
    float* s = wsum_planes + (size_t)c * n;
    auto weight_at = [&](int y, float dist) {
        const size_t i = (size_t)y * (size_t)g.w + (size_t)x;
        const float wt = ds_expf(-(dist / h2));
        a[i] += wt * at_clamped(p, g.w, g.h, x + dx, y + dy);
        s[i] += wt;
    };
    if (g.R == 0) {  // box_blur(p, 0) is the plane itself
        const float* row_src = p;
        for (int y = 0; y < g.h; y++) {
            const float d = row_src[(size_t)y * (size_t)g.w + (size_t)x] - at_clamped(p, g.w, g.h, x + dx, y + dy);
            weight_at(y, d * d);
        }
        return;
    }
    const float* t = tmp + (size_t)c * n;
    auto at = [&](int yy) { return t[(size_t)clampi(yy, 0, g.h - 1) * (size_t)g.w + (size_t)x]; };
    float acc = 0.0f;
    for (int i = -g.R; i <= g.R; i++) acc += at(i);
    for (int y = 0; y < g.h; y++) {
        weight_at(y, acc * g.norm);
        acc += at(y + g.R + 1) - at(y - g.R);
    }
}

__global__ void k_nlmref_zero(float* __restrict__ p, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

__global__ void k_nlmref_final(const float* __restrict__ acc_planes, const float* __restrict__ wsum_planes, size_t n, int nchunks,
                               float* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float acc = 0.0f, ws = 0.0f;
    for (int c = 0; c < nchunks; c++) {
        acc += acc_planes[(size_t)c * n + i];
        ws += wsum_planes[(size_t)c * n + i];
    }
    out[i] = acc / fmaxf(ws, 1e-12f);
}

RefGeom ref_geom(size_t w, size_t h, long search, size_t patch, int threads) {
    RefGeom g{};
    g.w = (int)w;
    g.h = (int)h;
    g.S = (int)search;
    g.R = (int)patch;
    g.n_off = (2 * g.S + 1) * (2 * g.S + 1);
    const int t = threads < 1 ? 1 : threads;
    g.chunk = (g.n_off + t - 1) / t;
    if (g.chunk < 1) g.chunk = 1;
    g.nchunks = (g.n_off + g.chunk - 1) / g.chunk;
    g.norm = 1.0f / (float)(2 * g.R + 1);
    return g;
}

}  // namespace

size_t nlm_ref_workspace_floats(size_t w, size_t h, long search, size_t patch, int threads) {
    if (search < 0) return 0;
    const RefGeom g = ref_geom(w, h, search, patch, threads);
    return 3 * (size_t)g.nchunks * w * h;
}

hipError_t launch_nlm_reference_order(const float* in, float* out, size_t w, size_t h, NlmH hp, long search, size_t patch, int threads,
                                      float* ws, hipStream_t st) {
    const size_t n = w * h;
    if (n == 0 || search < 0) return hipSuccess;
    const RefGeom g = ref_geom(w, h, search, patch, threads);
    float* acc = ws;
    float* wsum = ws + (size_t)g.nchunks * n;
    float* tmp = ws + 2 * (size_t)g.nchunks * n;
    const size_t nz = 2 * (size_t)g.nchunks * n;
    hipLaunchKernelGGL(k_nlmref_zero, dim3((unsigned)((nz + 255) / 256)), dim3(256), 0, st, ws, nz);
    NlmWeight wp{hp.h_param, hp.sigma_dev, hp.sigma_mul, 0.0f};
    const dim3 grid_rows((unsigned)((h + 63) / 64), (unsigned)g.nchunks);
    const dim3 grid_cols((unsigned)((w + 63) / 64), (unsigned)g.nchunks);
    for (int step = 0; step < g.chunk; step++) {
        if (g.R > 0) hipLaunchKernelGGL(k_nlmref_rows, grid_rows, dim3(64), 0, st, in, g, step, tmp);
        hipLaunchKernelGGL(k_nlmref_cols, grid_cols, dim3(64), 0, st, in, g, step, wp, tmp, acc, wsum);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL(k_nlmref_final, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, acc, wsum, n, g.nchunks, out);
    return hipGetLastError();
}

}  // namespace ds
