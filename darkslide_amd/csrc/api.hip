// C ABI of the raw detail path (include/darkslide_gpu.h): contexts,
// device workspace, host<->device staging, stage timing and the entry
// points that stand in for the reference's demosaic / enhance functions.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <atomic>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "darkslide_gpu.h"
#include "demosaic.h"
#include "denoise.h"
#include "detail.h"
#include "highlights.h"
#include "nlm.h"
#include "nlm_ref.h"
#include "box.h"
#include "rl.h"
#include "stages.h"
#include "synth.h"
#include "cnn.h"

#pragma clang fp contract(off)

namespace ds {

// ---- Oklab matrices exactly as color.rs builds them (f32 entries; mul_mm
// and invert accumulate in f64).
namespace {
typedef float M3[3][3];
void mul_mm(const M3 a, const M3 b, M3 o) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) s += (double)a[i][k] * (double)b[k][j];
            o[i][j] = (float)s;
        }
}
void invert(const M3 mi, M3 o) {
    double m[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) m[i][j] = mi[i][j];
    double det = m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                 m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    if (std::fabs(det) < 1e-12) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) o[i][j] = i == j ? 1.0f : 0.0f;
        return;
    }
    double inv = 1.0 / det;
    double r[3][3] = {
        {(m[1][1] * m[2][2] - m[2][1] * m[1][2]) * inv, (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * inv, (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * inv},
        {(m[1][2] * m[2][0] - m[1][0] * m[2][2]) * inv, (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * inv, (m[1][0] * m[0][2] - m[0][0] * m[1][2]) * inv},
        {(m[1][0] * m[2][1] - m[2][0] * m[1][1]) * inv, (m[2][0] * m[0][1] - m[0][0] * m[2][1]) * inv, (m[0][0] * m[1][1] - m[1][0] * m[0][1]) * inv},
    };
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) o[i][j] = (float)r[i][j];
}
void xy_to_xyz(const float xy[2], float o[3]) {
    float x = xy[0], y = std::fmax(xy[1], 1e-6f);
    o[0] = x / y;
    o[1] = 1.0f;
    o[2] = (1.0f - x - y) / y;
}
}  // namespace

OklabM build_oklab() {
    const float d50[2] = {0.34567f, 0.35850f}, d65[2] = {0.31271f, 0.32902f};            // color.rs:62-63
    const M3 brad = {{0.8951f, 0.2664f, -0.1614f}, {-0.7502f, 1.7135f, 0.0367f}, {0.0389f, -0.0685f, 1.0296f}};  // color.rs:79
    const float pr[2] = {0.734699f, 0.265301f}, pg[2] = {0.159597f, 0.840403f}, pb[2] = {0.036598f, 0.000105f};  // color.rs:129
    const M3 m1 = {{0.818933f, 0.36186674f, -0.12885971f}, {0.032984544f, 0.9293119f, 0.03614564f}, {0.0482003f, 0.26436627f, 0.6338517f}};
    const M3 m2 = {{0.21045426f, 0.7936178f, -0.004072047f}, {1.9779985f, -2.4285922f, 0.4505937f}, {0.025904037f, 0.78277177f, -0.80867577f}};
    // bradford(D50, D65) (color.rs:82-87)
    float xs[3], xd[3], s[3], d[3];
    xy_to_xyz(d50, xs);
    xy_to_xyz(d65, xd);
    mul_mv(brad, xs, s);
    mul_mv(brad, xd, d);
    M3 scale = {{d[0] / s[0], 0, 0}, {0, d[1] / s[1], 0}, {0, 0, d[2] / s[2]}};
    M3 binv, sb, bradm;
    invert(brad, binv);
    mul_mm(scale, brad, sb);
    mul_mm(binv, sb, bradm);
    // ProPhoto -> XYZ D50 (color.rs:90-98, white = D50)
    float vr[3], vg[3], vb[3], wx[3], sc[3];
    xy_to_xyz(pr, vr);
    xy_to_xyz(pg, vg);
    xy_to_xyz(pb, vb);
    M3 m = {{vr[0], vg[0], vb[0]}, {vr[1], vg[1], vb[1]}, {vr[2], vg[2], vb[2]}};
    M3 minv;
    invert(m, minv);
    xy_to_xyz(d50, wx);
    mul_mv(minv, wx, sc);
    M3 dg = {{sc[0], 0, 0}, {0, sc[1], 0}, {0, 0, sc[2]}};
    M3 xyz, bx;
    mul_mm(m, dg, xyz);
    mul_mm(bradm, xyz, bx);
    OklabM ok;
    mul_mm(m1, bx, ok.to_lms);  // color.rs:382-385
    invert(ok.to_lms, ok.from_lms);
    std::memcpy(ok.m2, m2, sizeof(M3));
    invert(m2, ok.m2_inv);
    return ok;
}

namespace {

__global__ void k_synth(const SynthScene* __restrict__ sp, uint16_t* __restrict__ out) {
    const SynthScene& s = *sp;
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)s.w * (size_t)s.h;
    if (i >= n) return;
    out[i] = synth_dn(s, (int)(i % (size_t)s.w), (int)(i / (size_t)s.w));
}

// plain copy of n floats (also the known-size reference for HBM counters)
__global__ void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n4) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n4; i += stride) dst[i] = src[i];
}

// plain fill; kept as a kernel so the enqueued work stays graph-capturable
__global__ void k_fill(float* __restrict__ p, size_t n, float v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// number of flagged regions (one block)
__global__ __launch_bounds__(256) void k_count_flags(const uint8_t* __restrict__ flags, int n, int* __restrict__ out) {
    __shared__ int red[256];
    int c = 0;
    for (int i = threadIdx.x; i < n; i += 256) c += flags[i] ? 1 : 0;
    red[threadIdx.x] = c;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// Auto white-balance statistics (pipeline/mod.rs:421-437) in FP64: over
// every third pixel, raw = p * n_as (f32), skip max > 0.9 or min < 0.002,
// w = sqrt(max), acc[c] += ln(raw[c]) w, n += w.  Per-block partial sums
// (acc0, acc1, acc2, n); k_wb_final adds them in block order.
constexpr int WB_NT = 256;
__global__ __launch_bounds__(WB_NT) void k_wb_partial(const float* __restrict__ rgb, size_t npx, float n0, float n1, float n2,
                                                      double* __restrict__ part) {
    __shared__ double red[4][WB_NT];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0, wn = 0.0;
    const size_t cnt = (npx + 2) / 3;
    for (size_t k = (size_t)blockIdx.x * WB_NT + threadIdx.x; k < cnt; k += (size_t)gridDim.x * WB_NT) {
        const float* p = rgb + 3 * (3 * k);
        const float r0 = p[0] * n0, r1 = p[1] * n1, r2 = p[2] * n2;
        const float m = fmaxf(fmaxf(r0, r1), r2);
        const float lo = fminf(fminf(r0, r1), r2);
        if (m > 0.9f || lo < 0.002f) continue;
        const double wt = sqrt((double)m);
        a0 += log((double)r0) * wt;
        a1 += log((double)r1) * wt;
        a2 += log((double)r2) * wt;
        wn += wt;
    }
    red[0][threadIdx.x] = a0;
    red[1][threadIdx.x] = a1;
    red[2][threadIdx.x] = a2;
    red[3][threadIdx.x] = wn;
    __syncthreads();
    for (int s = WB_NT / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int q = 0; q < 4; q++) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 4) part[(size_t)blockIdx.x * 4 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_wb_final(const double* __restrict__ part, int nblk, double* __restrict__ out) {
    if (threadIdx.x >= 4) return;
    double s = 0.0;
    for (int b = 0; b < nblk; b++) s += part[(size_t)b * 4 + threadIdx.x];
    out[threadIdx.x] = s;
}

}  // namespace
}  // namespace ds

using namespace ds;

// ---------------------------------------------------------------------------
// context

namespace {

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
};

enum Slot {
    S_IN = 0, S_IN2, S_OUT, S_CLIP, S_WS, S_L, S_A, S_B, S_LF, S_SIGMA, S_NOISE, S_SCENE, S_X, S_T, S_REGIONS, S_ACT,
    S_PIPE_RGB, S_PIPE_DEN, S_COUNTS, S_STATS, S_HL, S_PIPE_CLIP, S_PIPE_RAW, S_PIPE_OUT, S_NLMREF, S_COUNT
};

// S_HL: clip levels (3 floats) at 0, cfa_clip_levels partials from HL_PART_OFF.
constexpr size_t HL_PART_OFF = 64;
constexpr size_t HL_FLOATS = HL_PART_OFF + 3 * CLIP_PARTS;

// Tile-shape choice for one frame shape (variant indices of the tunable kernels).
struct Tuned {
    int rcd = 0, mhc = 0, rl = 0;
};

// A captured batch: the key is the byte image of its frames, options and
// tile choices; the executable graph is replayed while the key matches.
struct GraphEntry {
    std::vector<uint8_t> key;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int branches = 1;
    size_t n_frames = 0;
    unsigned long gen = 0;  // workspace generation the graph was captured against
    uint64_t net_uid = 0;   // the learned denoiser's network (ds_cnn::uid), 0 for none
};

struct Pending {
    std::string name;
    hipEvent_t a, b;
};

thread_local std::string g_create_error;

}  // namespace

struct ds_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::recursive_mutex mu;
    std::string err;
    Buf slots[S_COUNT];
    std::vector<void*> user_allocs;
    std::vector<void*> host_allocs;  // pinned host buffers from ds_host_alloc
    OklabM ok;
    bool prof = false;
    std::vector<Pending> pending;
    std::map<std::string, std::pair<double, long>> totals;
    // batched pipeline: parallel branches (each with its own stream and
    // workspace), tile choices per frame shape, captured graphs
    ds_ctx* parent = nullptr;
    bool capturing = false;
    std::vector<ds_ctx*> branches;
    std::map<std::pair<size_t, size_t>, Tuned> tuning;
    std::vector<GraphEntry> graphs;
    unsigned long slot_gen = 0;  // bumped when a workspace buffer is replaced
    // host-frame batches: upload and download streams beside `stream`
    hipStream_t copy_in = nullptr, copy_out = nullptr;
};

namespace {

int fail(ds_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c)
        c->err = buf;
    else
        g_create_error = buf;
    return code;
}

int hip_fail(ds_ctx* c, hipError_t e, const char* what) {
    return fail(c, e == hipErrorOutOfMemory ? DS_ERR_ALLOC : DS_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
}

#define HIPCHK(ctx, expr, what)                          \
    do {                                                 \
        hipError_t e_ = (expr);                          \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, what); \
    } while (0)

void* slot(ds_ctx* c, Slot s, size_t bytes, hipError_t* err) {
    *err = hipSuccess;
    Buf& b = c->slots[s];
    if (b.bytes >= bytes && b.p) return b.p;
    if (c->capturing) {  // workspaces are reserved before a graph capture; nothing allocates inside one
        *err = hipErrorStreamCaptureUnsupported;
        return nullptr;
    }
    if (b.p) {
        // captured batches run on the root stream and may hold this buffer:
        // wait for them and retire every graph captured before the change
        ds_ctx* root = c->parent ? c->parent : c;
        if ((*err = hipStreamSynchronize(root->stream)) != hipSuccess) return nullptr;
        root->slot_gen++;
        if ((*err = hipStreamSynchronize(c->stream)) != hipSuccess) return nullptr;
        (void)hipFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
    }
    size_t want = bytes < 256 ? 256 : bytes;
    *err = hipMalloc(&b.p, want);
    if (*err != hipSuccess) {
        b.p = nullptr;
        return nullptr;
    }
    b.bytes = want;
    return b.p;
}

#define SLOT(ctx, var, type, s, bytes)                                   \
    type var;                                                            \
    {                                                                    \
        hipError_t e_;                                                   \
        var = (type)slot(ctx, s, bytes, &e_);                            \
        if (!var) return hip_fail(ctx, e_, "device workspace allocation"); \
    }

// Launch wrapper: with profiling on, brackets the launch with events.
template <class F>
int run(ds_ctx* c, const char* name, F&& f) {
    hipEvent_t a = nullptr, b = nullptr;
    // a launch whose events cannot be made or recorded is not timed
    bool timed = c->prof && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess &&
                 hipEventRecord(a, c->stream) == hipSuccess;
    hipError_t e = f();
    if (timed) timed = hipEventRecord(b, c->stream) == hipSuccess;
    if (timed) {
        c->pending.push_back({name, a, b});
    } else {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    if (e != hipSuccess) return fail(c, DS_ERR_LAUNCH, "%s: %s", name, hipGetErrorString(e));
    return DS_OK;
}

void collect(ds_ctx* c) {
    if (c->pending.empty()) return;
    const bool synced = hipStreamSynchronize(c->stream) == hipSuccess;
    for (auto& p : c->pending) {
        float ms = 0.0f;
        if (synced && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            auto& t = c->totals[p.name];
            t.first += ms;
            t.second += 1;
        }
        (void)hipEventDestroy(p.a);
        (void)hipEventDestroy(p.b);
    }
    c->pending.clear();
}

bool valid_bayer(const uint8_t* p) {
    if (!p) return false;
    for (int i = 0; i < 4; i++)
        if (p[i] > 2) return false;
    bool diag03 = p[0] == 1 && p[3] == 1 && ((p[1] == 0 && p[2] == 2) || (p[1] == 2 && p[2] == 0));
    bool diag12 = p[1] == 1 && p[2] == 1 && ((p[0] == 0 && p[3] == 2) || (p[0] == 2 && p[3] == 0));
    return diag03 || diag12;
}

// Images below 2^32 pixels: the box passes index a plane with 32-bit
// element offsets (row * width).
int check_dims(ds_ctx* c, size_t w, size_t h) {
    if (!c) return DS_ERR_INVALID;
    if (w > (size_t)INT32_MAX / 4 || h > (size_t)INT32_MAX / 4 || (w && h && w * h >= ((size_t)1 << 32)))
        return fail(c, DS_ERR_INVALID, "image %zux%zu too large (2^32 pixels or more)", w, h);
    return DS_OK;
}

int set_device(ds_ctx* c) {
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    return DS_OK;
}

// ---- enqueue-only implementations (device pointers)

Tuned tuned_for(const ds_ctx* c, size_t w, size_t h) {
    const ds_ctx* root = c->parent ? c->parent : c;
    auto it = root->tuning.find({w, h});
    return it == root->tuning.end() ? Tuned{} : it->second;
}

// The CFA input of the demosaic-side kernels: a normalised f32 CFA, or raw
// DN normalised on the fly with black (or, when black_site is set, the
// black level of each CFA site at (y & 1) * 2 + (x & 1)) and white per colour.
DemosaicArgs cfa_args(const float* cfa, const uint16_t* raw, size_t w, size_t h, size_t stride, const uint8_t* pat, float black,
                      const float* black_site, const float* white) {
    DemosaicArgs a{};
    a.cfa_f32 = cfa;
    a.raw_u16 = raw;
    a.stride = stride;
    a.w = w;
    a.h = h;
    std::memcpy(a.pattern, pat, 4);
    a.black = black;
    a.per_site = black_site ? 1 : 0;
    for (int k = 0; k < 4; k++) a.black_site[k] = black_site ? black_site[k] : black;
    for (int k = 0; k < 3; k++) {
        a.white[k] = white ? white[k] : 1.0f;
        a.mul[k] = 1.0f;
    }
    return a;
}

int demosaic_dev(ds_ctx* c, const float* cfa, const uint16_t* raw, size_t w, size_t h, size_t stride, const uint8_t* pat, float black,
                 const float* black_site, const float* white, const float* mul, int algo, float* rgb, float* clip) {
    if (!valid_bayer(pat)) return fail(c, DS_ERR_INVALID, "pattern must be a Bayer 2x2 with a green checkerboard");
    if (algo != DS_ALGO_RCD && algo != DS_ALGO_MHC) return fail(c, DS_ERR_INVALID, "unknown demosaic algorithm %d", algo);
    if (!mul || !rgb || (!cfa && !raw)) return fail(c, DS_ERR_INVALID, "null pointer");
    if (w == 0 || h == 0) return DS_OK;
    if (stride < w) return fail(c, DS_ERR_INVALID, "stride %zu < width %zu", stride, w);
    DemosaicArgs a = cfa_args(cfa, raw, w, h, stride, pat, black, black_site, white);
    for (int k = 0; k < 3; k++) a.mul[k] = mul[k];
    a.algo = algo;
    a.rgb = rgb;
    a.clip = clip;
    const Tuned t = tuned_for(c, w, h);
    const int variant = algo == DS_ALGO_RCD ? t.rcd : t.mhc;
    return run(c, algo == DS_ALGO_RCD && w >= 16 && h >= 16 ? "demosaic_rcd" : "demosaic_mhc",
               [&] { return launch_demosaic(a, c->stream, variant); });
}

// cfa_clip_levels (demosaic.rs:419-430) into levels_dev (3 device floats).
int clip_levels_dev(ds_ctx* c, const float* cfa, const uint16_t* raw, size_t w, size_t h, size_t stride, const uint8_t* pat, float black,
                    const float* black_site, const float* white, float* levels_dev) {
    if (!valid_bayer(pat)) return fail(c, DS_ERR_INVALID, "pattern must be a Bayer 2x2 with a green checkerboard");
    if (!levels_dev || (!cfa && !raw)) return fail(c, DS_ERR_INVALID, "null pointer");
    if (stride < w) return fail(c, DS_ERR_INVALID, "stride %zu < width %zu", stride, w);
    SLOT(c, hl, float*, S_HL, HL_FLOATS * 4);
    const DemosaicArgs a = cfa_args(cfa, raw, w, h, stride, pat, black, black_site, white);
    return run(c, "cfa_clip_levels", [&] { return launch_cfa_clip_levels(a, hl + HL_PART_OFF, levels_dev, c->stream); });
}

// superpixel (demosaic.rs:44-87): rgb (w/2 x h/2 x 3) and clip (w/2 x h/2, or null).
int superpixel_dev(ds_ctx* c, const float* cfa, const uint16_t* raw, size_t w, size_t h, size_t stride, const uint8_t* pat, float black,
                   const float* black_site, const float* white, const float* mul, float* rgb, float* clip) {
    if (!valid_bayer(pat)) return fail(c, DS_ERR_INVALID, "pattern must be a Bayer 2x2 with a green checkerboard");
    if (!mul || !rgb || (!cfa && !raw)) return fail(c, DS_ERR_INVALID, "null pointer");
    if (stride < w) return fail(c, DS_ERR_INVALID, "stride %zu < width %zu", stride, w);
    if (w < 2 || h < 2) return DS_OK;
    DemosaicArgs a = cfa_args(cfa, raw, w, h, stride, pat, black, black_site, white);
    for (int k = 0; k < 3; k++) a.mul[k] = mul[k];
    a.rgb = rgb;
    a.clip = clip;
    return run(c, "superpixel", [&] { return launch_superpixel(a, c->stream); });
}

// Normalised CFA (decode.rs:300-303) of a raw view into out (w x h).
int normalize_dev(ds_ctx* c, const uint16_t* raw, size_t w, size_t h, size_t stride, const uint8_t* pat, float black, const float* black_site,
                  const float* white, float* out) {
    if (!raw || !out) return fail(c, DS_ERR_INVALID, "null pointer");
    const DemosaicArgs a = cfa_args(nullptr, raw, w, h, stride, pat, black, black_site, white);
    return run(c, "cfa_normalize", [&] { return launch_cfa_normalize(a, out, c->stream); });
}

// A ds_raw_cfa reduced to the view the CFA kernels read: pointer at the
// crop origin, the buffer width as stride, the pattern and per-site black
// levels of the crop (decode.rs:258-306).
struct RawView {
    const uint16_t* raw;
    size_t w, h, stride;
    uint8_t pat[4];
    float black_site[4];
    float white[3];
};

int raw_view(ds_ctx* c, const ds_raw_cfa* r, RawView& v) {
    if (!r || !r->data || !r->black) return fail(c, DS_ERR_INVALID, "null pointer");
    if (!valid_bayer(r->pattern)) return fail(c, DS_ERR_INVALID, "pattern must be a Bayer 2x2 with a green checkerboard");
    if (r->crop_x > r->width || r->crop_w > r->width - r->crop_x || r->crop_y > r->height || r->crop_h > r->height - r->crop_y)
        return fail(c, DS_ERR_INVALID, "crop rectangle outside the %zux%zu buffer", r->width, r->height);
    if (int rc = check_dims(c, r->crop_w, r->crop_h)) return rc;
    const size_t bw = r->black_w ? r->black_w : 1, bh = r->black_h ? r->black_h : 1, bcpp = r->black_cpp ? r->black_cpp : 1;
    if (r->nblack == 0) return fail(c, DS_ERR_INVALID, "empty black level table");
    if (r->nblack != 1) {
        if (bw > 2 || bh > 2) return fail(c, DS_ERR_INVALID, "black level table period %zux%zu over 2 is not supported", bw, bh);
        if (r->nblack < bw * bh * bcpp) return fail(c, DS_ERR_INVALID, "black level table shorter than %zux%zux%zu", bw, bh, bcpp);
    }
    for (size_t py = 0; py < 2; py++)
        for (size_t px = 0; px < 2; px++) {
            const size_t sx = r->crop_x + px, sy = r->crop_y + py;
            v.black_site[py * 2 + px] = r->nblack == 1 ? r->black[0] : r->black[((sy % bh) * bw + (sx % bw)) * bcpp];
            v.pat[py * 2 + px] = r->pattern[((sy & 1) << 1) | (sx & 1)];
        }
    v.raw = r->data + r->crop_y * r->width + r->crop_x;
    v.w = r->crop_w;
    v.h = r->crop_h;
    v.stride = r->width;
    for (int k = 0; k < 3; k++) v.white[k] = r->white[k];
    return DS_OK;
}

int demosaic_raw_dev(ds_ctx* c, const RawView& v, const float* mul, int algo, int quality, float* rgb_out, float* clip_out) {
    if (quality == DS_QUALITY_DRAFT)
        return superpixel_dev(c, nullptr, v.raw, v.w, v.h, v.stride, v.pat, 0.0f, v.black_site, v.white, mul, rgb_out, clip_out);
    if (quality != DS_QUALITY_FULL) return fail(c, DS_ERR_INVALID, "unknown quality %d", quality);
    return demosaic_dev(c, nullptr, v.raw, v.w, v.h, v.stride, v.pat, 0.0f, v.black_site, v.white, mul, algo, rgb_out, clip_out);
}

// Uploads the rows of the crop of a host raw buffer; v then views the device copy.
int upload_raw(ds_ctx* c, const ds_raw_cfa* raw, RawView& v) {
    if (int rc = raw_view(c, raw, v)) return rc;
    const size_t n = v.h ? (v.h - 1) * v.stride + v.w : 0;
    SLOT(c, d_raw, uint16_t*, S_IN, (n ? n : 1) * 2);
    if (n) HIPCHK(c, hipMemcpyAsync(d_raw, v.raw, n * 2, hipMemcpyHostToDevice, c->stream), "H2D");
    v.raw = d_raw;
    return DS_OK;
}

NrParams nr_params(const ds_nr_params& p) {
    return NrParams{p.luminance, p.luminance_detail, p.luminance_contrast, p.color, p.color_detail, p.color_smoothness};
}

int noise_reduction_dev(ds_ctx* c, float* rgb, size_t w, size_t h, const ds_nr_params* nr, const float* field, float field_max,
                        float scale) {
    SLOT(c, ws, float*, S_WS, (nr_ws_floats(w, h) ? nr_ws_floats(w, h) : 1) * 4);
    const NrParams P = nr_params(*nr);
    return run(c, "noise_reduction", [&] { return launch_noise_reduction(rgb, w, h, P, field, field_max, scale, c->ok, ws, c->stream); });
}

int sharpen_dev(ds_ctx* c, float* rgb, size_t w, size_t h, const ds_sharpen_params* sh, const float* field, float scale) {
    SLOT(c, ws, float*, S_WS, (sharpen_ws_floats(w, h) ? sharpen_ws_floats(w, h) : 1) * 4);
    const SharpParams P{sh->amount, sh->radius, sh->detail, sh->masking};
    return run(c, "sharpen", [&] { return launch_sharpen(rgb, w, h, P, field, scale, ws, c->stream); });
}

int fgf_dev(ds_ctx* c, const float* I, const float* p, size_t w, size_t h, size_t r, float eps, size_t sub, float* out) {
    if (w * h == 0) return DS_OK;
    if (r > 1024) return fail(c, DS_ERR_INVALID, "radius %zu too large", r);
    SLOT(c, ws, float*, S_WS, fgf_ws_floats(w, h, sub) * 4);
    return run(c, "fast_guided_filter", [&] { return launch_fast_guided_filter(I, p, w, h, r, eps, sub, ws, out, c->stream); });
}

// export.rs:244-247 post-steps in place on a device image.
int export_detail_dev(ds_ctx* c, float* rgb, size_t w, size_t h, bool superres, int mode, int amount) {
    if (!superres && (mode < DS_SHARPEN_FOR_NONE || mode > DS_SHARPEN_FOR_GLOSSY || amount < DS_SHARPEN_AMOUNT_LOW ||
                      amount > DS_SHARPEN_AMOUNT_HIGH))
        return fail(c, DS_ERR_INVALID, "output sharpen mode %d / amount %d out of range", mode, amount);
    if (w * h == 0 || (!superres && mode == DS_SHARPEN_FOR_NONE)) return DS_OK;
    SLOT(c, ws, float*, S_WS, export_detail_ws_floats(w, h) * 4);
    if (superres) return run(c, "super_res_refine", [&] { return launch_super_res_refine(rgb, w, h, ws, c->stream); });
    return run(c, "output_sharpen", [&] { return launch_output_sharpen(rgb, w, h, mode, amount, ws, c->stream); });
}

// reconstruct_highlights (demosaic.rs:435-502) in place; raw_clip on the device.
int highlights_dev(ds_ctx* c, float* rgb, const float* clip, size_t w, size_t h, const float* mul, const float* raw_clip) {
    const size_t n = w * h;
    if (n == 0) return DS_OK;
    if (!rgb || !clip || !mul || !raw_clip) return fail(c, DS_ERR_INVALID, "null pointer");
    SLOT(c, ws, float*, S_WS, highlights_ws_floats(w, h) * 4);
    return run(c, "reconstruct_highlights", [&] { return launch_reconstruct_highlights(rgb, clip, w, h, mul, raw_clip, ws, c->stream); });
}

int noise_dev(ds_ctx* c, const float* plane, size_t w, size_t h, float* sigma_dev) {
    SLOT(c, scratch, uint32_t*, S_NOISE, noise_scratch_u32(w, h) * sizeof(uint32_t));
    return run(c, "estimate_noise", [&] { return launch_estimate_noise(plane, w, h, scratch, sigma_dev, c->stream); });
}

int zero_fill(ds_ctx* c, float* out, size_t n) {
    return run(c, "fill_zero", [&] {
        hipLaunchKernelGGL(k_fill, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, out, n, 0.0f);
        return hipGetLastError();
    });
}

int nlm_dev(ds_ctx* c, const float* in, size_t w, size_t h, NlmH hp, long search, size_t patch, float* out) {
    if (w == 0 || h == 0) return DS_OK;
    // (-S..=S) is empty for S < 0: no weights at all, so acc / max(0, 1e-12) = 0
    if (search < 0) return zero_fill(c, out, w * h);
    if (patch > 64) return fail(c, DS_ERR_INVALID, "patch radius %zu too large", patch);
    if (search > 1024) return fail(c, DS_ERR_INVALID, "search radius %ld too large", search);
    return run(c, "nlm", [&] { return launch_nlm(in, out, w, h, hp, search, patch, c->stream); });
}

// NLM in the reference's summation order for a pool of `threads` threads.
int nlm_ref_dev(ds_ctx* c, const float* in, size_t w, size_t h, NlmH hp, long search, size_t patch, int threads, float* out) {
    if (w == 0 || h == 0) return DS_OK;
    if (search < 0) return zero_fill(c, out, w * h);
    if (threads < 1 || threads > 4096) return fail(c, DS_ERR_INVALID, "thread count %d out of range", threads);
    if (patch > 64) return fail(c, DS_ERR_INVALID, "patch radius %zu too large", patch);
    if (search > 1024) return fail(c, DS_ERR_INVALID, "search radius %ld too large", search);
    SLOT(c, ws, float*, S_NLMREF, nlm_ref_workspace_floats(w, h, search, patch, threads) * 4);
    return run(c, "nlm_reference_order",
               [&] { return launch_nlm_reference_order(in, out, w, h, hp, search, patch, threads, ws, c->stream); });
}

// denoise (enhance.rs:74-92); ref_threads > 0 runs its NLM in the
// reference's summation order for a pool of that many threads.
int denoise_dev(ds_ctx* c, const float* rgb, size_t w, size_t h, float amount, long search, size_t patch, float* out,
                int ref_threads = 0) {
    const size_t n = w * h;
    if (n == 0) return DS_OK;
    SLOT(c, L, float*, S_L, n * 4);
    SLOT(c, A, float*, S_A, n * 4);
    SLOT(c, B, float*, S_B, n * 4);
    SLOT(c, LF, float*, S_LF, n * 4);
    SLOT(c, sig, float*, S_SIGMA, 64);
    SLOT(c, ws, float*, S_WS, 12 * n * 4);
    int rc = run(c, "oklab_fwd", [&] { return launch_oklab_fwd(rgb, n, c->ok, L, A, B, c->stream); });
    if (rc) return rc;
    if ((rc = noise_dev(c, L, w, h, sig))) return rc;
    // enhance.rs:82-88
    const float k = amount / 100.0f;
    NlmH hp{0.0f, sig, 0.6f + 1.6f * k};
    const float m = 0.5f + 0.5f * k;
    // The tile NLM forms the guided pair's guide planes (the first four of
    // ws) as it stores each result; the reference-order NLM, negative search
    // radii and the direct kernel leave them to the guided pair.
    bool guide_ready = false;
    if (ref_threads > 0 || search < 0) {
        rc = ref_threads > 0 ? nlm_ref_dev(c, L, w, h, hp, search, patch, ref_threads, LF) : nlm_dev(c, L, w, h, hp, search, patch, LF);
    } else {
        if (patch > 64) return fail(c, DS_ERR_INVALID, "patch radius %zu too large", patch);
        if (search > 1024) return fail(c, DS_ERR_INVALID, "search radius %ld too large", search);
        const NlmGuide g{L, A, B, m, ws, ws + n, ws + 2 * n, ws + 3 * n};
        rc = run(c, "nlm", [&] { return launch_nlm_guide(L, LF, w, h, hp, search, patch, g, &guide_ready, c->stream); });
    }
    if (rc) return rc;
    float rf = std::fmax(((float)(w > h ? w : h) / 600.0f) * (1.0f + 2.0f * k), 2.0f);
    size_t r = (size_t)rf;
    return run(c, "guided_filter", [&] { return launch_guided_pair(L, LF, A, B, m, r, sig, c->ok, ws, out, w, h, c->stream, guide_ready); });
}

int raw_details_dev(ds_ctx* c, const float* rgb, size_t w, size_t h, float sigma, size_t iters, float* out) {
    const size_t n = w * h;
    if (n == 0) return DS_OK;
    if (!(sigma >= 0.0f) || sigma > 1000.0f) return fail(c, DS_ERR_INVALID, "sigma %g out of range", (double)sigma);
    SLOT(c, ws, float*, S_WS, rl_workspace_floats(w, h) * 4);
    const int variant = tuned_for(c, w, h).rl;
    return run(c, "raw_details", [&] { return launch_raw_details(rgb, w, h, sigma, iters, ws, out, c->stream, variant); });
}

int synth_dev(ds_ctx* c, uint64_t seed, size_t w, size_t h, int bits, float iso, const uint8_t* pattern, uint16_t* raw_out) {
    if (!raw_out || !valid_bayer(pattern) || (bits != 12 && bits != 14) || !(iso >= 1.0f || iso == 0.0f) || w > (1u << 20) || h > (1u << 20))
        return fail(c, DS_ERR_INVALID, "invalid synthetic frame arguments");
    SynthScene s;
    synth_scene_init(&s, seed, (int)w, (int)h, bits, iso, pattern);
    size_t n = w * h;
    if (n == 0) return DS_OK;
    SLOT(c, d_scene, SynthScene*, S_SCENE, sizeof(SynthScene));
    HIPCHK(c, hipMemcpyAsync(d_scene, &s, sizeof(SynthScene), hipMemcpyHostToDevice, c->stream), "H2D");
    int rc = run(c, "synth_bayer", [&] {
        hipLaunchKernelGGL(k_synth, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, d_scene, raw_out);
        return hipGetLastError();
    });
    // the scene lives on this stack frame until the copy has been consumed
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return rc;
}

}  // namespace

// ---------------------------------------------------------------------------
// learned denoiser

struct ds_cnn {
    ds_ctx* ctx = nullptr;
    uint64_t uid = 0;          // never reused, unlike the address: batch graphs are keyed on it
    float* state = nullptr;    // params | grad | adam m | adam v, CNN_NPARAM floats each
    uint16_t* wpk = nullptr;   // bf16 weights of the MFMA convolutions (CNN_WPK values)
    long step = 0;
    uint64_t seed = 0;
    bool packed = false;
    size_t tr_px = 0;          // training buffers, sized for tr_px pixels
    uint16_t* tr_act = nullptr;  // 7 activations + 2 activation gradients, 32 bf16 per pixel each
    float* tr_f = nullptr;       // x | tgt | out | backward scratch
    size_t tr_fcap = 0;          // floats in tr_f
    float* params() { return state; }
    float* grad() { return state + CNN_NPARAM; }
    float* adam_m() { return state + 2 * CNN_NPARAM; }
    float* adam_v() { return state + 3 * CNN_NPARAM; }
};

namespace {

constexpr size_t CNN_ACT_BYTES = 32 * sizeof(uint16_t);  // per pixel

int cnn_pack_now(ds_cnn* net) {
    ds_ctx* c = net->ctx;
    if (net->packed) return DS_OK;
    int rc = run(c, "cnn_pack", [&] { return cnn_pack(net->params(), net->wpk, c->stream); });
    if (rc == DS_OK) net->packed = true;
    return rc;
}

size_t tr_floats(size_t npx, int n, int H, int W) { return 3 * npx + cnn_scratch_floats(n, H, W); }

int cnn_ensure_train(ds_cnn* net, int n, int H, int W) {
    ds_ctx* c = net->ctx;
    const size_t npx = (size_t)n * H * W;
    // the backward scratch depends on the batch shape, not only on its pixel count
    const size_t fl = tr_floats(npx, n, H, W);
    if (net->tr_px >= npx && net->tr_fcap >= fl) return DS_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    if (net->tr_act) (void)hipFree(net->tr_act);
    if (net->tr_f) (void)hipFree(net->tr_f);
    net->tr_act = nullptr;
    net->tr_f = nullptr;
    const size_t px = std::max(npx, net->tr_px), cap = std::max(fl, net->tr_fcap);
    net->tr_px = 0;
    net->tr_fcap = 0;
    HIPCHK(c, hipMalloc(&net->tr_act, 9 * px * CNN_ACT_BYTES), "hipMalloc (CNN activations)");
    HIPCHK(c, hipMalloc(&net->tr_f, cap * sizeof(float)), "hipMalloc (CNN training buffers)");
    net->tr_px = px;
    net->tr_fcap = cap;
    return DS_OK;
}

int cnn_check_shape(ds_ctx* c, size_t n, size_t h, size_t w) {
    if (n == 0 || h == 0 || w == 0) return fail(c, DS_ERR_INVALID, "empty CNN batch");
    if (n > 65535 || h > (1u << 16) || w > (1u << 16) || n * h * w > ((size_t)1 << 31))
        return fail(c, DS_ERR_INVALID, "CNN batch %zux%zux%zu too large", n, h, w);
    return DS_OK;
}

// x (n*h*w, device) -> out; training layout keeps all 7 activations.
int cnn_forward_train(ds_cnn* net, const float* x, int n, int H, int W, float* out) {
    ds_ctx* c = net->ctx;
    const size_t npx = (size_t)n * H * W;
    if (int rc = cnn_pack_now(net)) return rc;
    uint16_t* acts[7];
    for (int k = 0; k < 7; k++) acts[k] = net->tr_act + (size_t)k * npx * 32;
    return run(c, "cnn_forward", [&] { return cnn_forward(net->params(), net->wpk, x, n, H, W, acts, out, c->stream); });
}

// Loss partials of the last backward (host sum in double).
int cnn_loss_host(ds_cnn* net, size_t npx, int n, int H, int W, int margin, double* loss) {
    ds_ctx* c = net->ctx;
    const size_t nparts = (npx + 255) / 256;
    std::vector<float> parts(nparts);
    float* lossp = net->tr_f + 3 * npx + npx;  // scratch: dout | loss partials | ...
    HIPCHK(c, hipMemcpyAsync(parts.data(), lossp, nparts * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    double s = 0.0;
    for (float v : parts) s += v;
    const long nvalid = (long)n * (H - 2 * margin) * (W - 2 * margin);
    *loss = s / (double)(nvalid > 0 ? nvalid : 1);
    return DS_OK;
}

// Forward + backward of a device batch already in tr_f (x at 0, tgt at npx):
// gradient into net->grad().
int cnn_grad_dev(ds_cnn* net, int n, int H, int W, int margin) {
    ds_ctx* c = net->ctx;
    const size_t npx = (size_t)n * H * W;
    float* x = net->tr_f;
    float* tgt = x + npx;
    float* out = tgt + npx;
    float* scratch = out + npx;
    if (int rc = cnn_forward_train(net, x, n, H, W, out)) return rc;
    uint16_t* acts[7];
    for (int k = 0; k < 7; k++) acts[k] = net->tr_act + (size_t)k * npx * 32;
    uint16_t* gacts[2] = {net->tr_act + 7 * npx * 32, net->tr_act + 8 * npx * 32};
    return run(c, "cnn_backward", [&] {
        return cnn_backward(net->params(), net->wpk, x, tgt, n, H, W, margin, acts, out, gacts, scratch, net->grad(), nullptr,
                            c->stream);
    });
}

int cnn_adam_step(ds_cnn* net, float lr) {
    ds_ctx* c = net->ctx;
    net->step += 1;
    net->packed = false;
    int rc = run(c, "cnn_adam", [&] { return cnn_adam(net->params(), net->grad(), net->adam_m(), net->adam_v(), (int)net->step, lr, c->stream); });
    if (rc) return rc;
    return cnn_pack_now(net);
}

int cnn_check_batch(ds_cnn* net, size_t n, size_t h, size_t w, size_t margin) {
    ds_ctx* c = net->ctx;
    if (int rc = cnn_check_shape(c, n, h, w)) return rc;
    if (2 * margin >= h || 2 * margin >= w) return fail(c, DS_ERR_INVALID, "margin %zu leaves no pixels of %zux%zu", margin, w, h);
    return cnn_ensure_train(net, (int)n, (int)h, (int)w);
}

// Learned luminance: LF = L + s (net(x) - x), x = cnn_input(L, s), s from sig.
// Runs on c's stream with c's workspace (activations included), so
// contexts (and the branches of a batch) can share one network.
int cnn_plane_dev(ds_ctx* c, ds_cnn* net, const float* L, size_t w, size_t h, const float* sig, float* lf) {
    const size_t n = w * h;
    if (int rc = cnn_check_shape(c, 1, h, w)) return rc;
    if (!net->packed) {
        int rc = run(c, "cnn_pack", [&] { return cnn_pack(net->params(), net->wpk, c->stream); });
        if (rc) return rc;
        net->packed = true;
    }
    SLOT(c, x, float*, S_X, n * 4);
    int rc = run(c, "cnn_input", [&] { return cnn_input(L, (int)w, (int)h, sig, x, c->stream); });
    if (rc) return rc;
    SLOT(c, act, uint16_t*, S_ACT, cnn_fused_act_bytes(1, (int)h, (int)w));
    // the output pass forms LF = L + s (y - x) itself
    return run(c, "cnn_forward", [&] { return cnn_forward_fused(net->params(), net->wpk, x, 1, (int)h, (int)w, act, lf, c->stream, L, sig); });
}

// denoise with the learned luminance stand-in for NLM, the per-region NLM
// spot check and fallback, then the reference's composite (enhance.rs:74-92).
int denoise_learned_dev(ds_ctx* c, ds_cnn* net, const float* rgb, size_t w, size_t h, float amount, long search, size_t patch, float thr_db,
                        float* out, ds_learned_stats* stats, int* count_dev = nullptr) {
    const size_t n = w * h;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (n == 0) return DS_OK;
    if (search < 0 || search > 1024 || patch > 64) return fail(c, DS_ERR_INVALID, "search %ld / patch %zu out of range", search, patch);
    SLOT(c, L, float*, S_L, n * 4);
    SLOT(c, A, float*, S_A, n * 4);
    SLOT(c, B, float*, S_B, n * 4);
    SLOT(c, LF, float*, S_LF, n * 4);
    SLOT(c, sig, float*, S_SIGMA, 64);
    SLOT(c, ws, float*, S_WS, 12 * n * 4);
    const size_t nreg = nlm_regions(w, h);
    SLOT(c, regbuf, float*, S_REGIONS, (2 * nreg + 64 + nlm_spot_scratch_floats(w, h)) * 4);
    float* gap = regbuf;
    uint8_t* flags = reinterpret_cast<uint8_t*>(regbuf + nreg);
    float* spot = regbuf + 2 * nreg + 64;
    int rc = run(c, "oklab_fwd", [&] { return launch_oklab_fwd(rgb, n, c->ok, L, A, B, c->stream); });
    if (rc) return rc;
    if ((rc = noise_dev(c, L, w, h, sig))) return rc;
    if ((rc = cnn_plane_dev(c, net, L, w, h, sig, LF))) return rc;
    const float k = amount / 100.0f;
    NlmH hp{0.0f, sig, 0.6f + 1.6f * k};
    rc = run(c, "nlm_spot_check", [&] { return launch_nlm_spot_check(L, LF, w, h, hp, search, patch, thr_db, gap, flags, spot, c->stream); });
    if (rc) return rc;
    rc = run(c, "nlm_fallback", [&] { return launch_nlm_regions(L, LF, w, h, hp, search, patch, flags, c->stream); });
    if (rc) return rc;
    if (count_dev) {
        rc = run(c, "count_fallback", [&] {
            hipLaunchKernelGGL(k_count_flags, dim3(1), dim3(256), 0, c->stream, flags, (int)nreg, count_dev);
            return hipGetLastError();
        });
        if (rc) return rc;
    }
    const float m = 0.5f + 0.5f * k;
    float rf = std::fmax(((float)(w > h ? w : h) / 600.0f) * (1.0f + 2.0f * k), 2.0f);
    size_t r = (size_t)rf;
    rc = run(c, "guided_filter", [&] { return launch_guided_pair(L, LF, A, B, m, r, sig, c->ok, ws, out, w, h, c->stream); });
    if (rc || !stats) return rc;
    std::vector<float> g(nreg);
    std::vector<uint8_t> f(nreg);
    HIPCHK(c, hipMemcpyAsync(g.data(), gap, nreg * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipMemcpyAsync(f.data(), flags, nreg, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    stats->regions = nreg;
    double sum = 0.0;
    float mx = 0.0f;
    for (size_t i = 0; i < nreg; i++) {
        stats->fallback += f[i] ? 1 : 0;
        sum += g[i];
        mx = std::max(mx, g[i]);
    }
    stats->mean_gap_db = (float)(sum / (double)nreg);
    stats->max_gap_db = mx;
    return DS_OK;
}

// Training pairs from one synthetic frame: demosaic (RCD), Oklab L, noise
// sigma, NLM at h = 0.6 sigma; x = cnn_input(L), t = x + (LF - L) / s;
// n_crops crops of crop x crop at positions drawn from seed.
int make_pairs_dev(ds_ctx* c, uint64_t seed, size_t fw, size_t fh, float iso, long search, size_t patch, size_t crop, size_t n_crops,
                   float* x_out, float* t_out) {
    if (crop == 0 || crop > fw || crop > fh || n_crops == 0) return fail(c, DS_ERR_INVALID, "crop %zu of %zux%zu", crop, fw, fh);
    if (!(iso >= 1.0f)) return fail(c, DS_ERR_INVALID, "iso %g < 1", (double)iso);
    const size_t n = fw * fh;
    static const uint8_t rggb[4] = {0, 1, 1, 2};
    const float white[3] = {16383.0f, 16383.0f, 16383.0f};
    const float ones[3] = {1.0f, 1.0f, 1.0f};
    SLOT(c, raw, uint16_t*, S_IN, n * 2);
    SLOT(c, rgb, float*, S_OUT, n * 12);
    SLOT(c, L, float*, S_L, n * 4);
    SLOT(c, A, float*, S_A, n * 4);
    SLOT(c, B, float*, S_B, n * 4);
    SLOT(c, LF, float*, S_LF, n * 4);
    SLOT(c, sig, float*, S_SIGMA, 64);
    SLOT(c, x, float*, S_X, n * 4);
    SLOT(c, t, float*, S_T, n * 4);
    int rc = synth_dev(c, seed, fw, fh, 14, iso, rggb, raw);
    if (rc) return rc;
    if ((rc = demosaic_dev(c, nullptr, raw, fw, fh, fw, rggb, 512.0f, nullptr, white, ones, DS_ALGO_RCD, rgb, nullptr))) return rc;
    if ((rc = run(c, "oklab_fwd", [&] { return launch_oklab_fwd(rgb, n, c->ok, L, A, B, c->stream); }))) return rc;
    if ((rc = noise_dev(c, L, fw, fh, sig))) return rc;
    if ((rc = nlm_dev(c, L, fw, fh, NlmH{0.0f, sig, 0.6f}, search, patch, LF))) return rc;
    if ((rc = run(c, "cnn_input", [&] { return cnn_input(L, (int)fw, (int)fh, sig, x, c->stream); }))) return rc;
    if ((rc = run(c, "cnn_target", [&] { return cnn_target(x, L, LF, n, sig, t, c->stream); }))) return rc;
    return run(c, "cnn_crops", [&] { return cnn_crops(x, t, (int)fw, (int)fh, (int)crop, (int)n_crops, seed, x_out, t_out, c->stream); });
}

// He-normal initial weights from a seeded splitmix64 / Box-Muller stream;
// zero biases.
void cnn_init_params(uint64_t seed, std::vector<float>& p) {
    p.assign(CNN_NPARAM, 0.0f);
    SplitMix64 r{seed ^ 0x5DEECE66DULL};
    auto normal = [&]() {
        const double u1 = r.uniform(), u2 = r.uniform();
        return std::sqrt(-2.0 * std::log(u1 + 1e-300)) * std::cos(6.283185307179586 * u2);
    };
    for (size_t i = 0; i < 32 * 9; i++) p[CNN_OFF_IN + i] = (float)(normal() * std::sqrt(2.0 / 9.0));
    for (int l = 0; l < 6; l++)
        for (size_t i = 0; i < 32 * 9 * 32; i++) p[CNN_OFF_MID + l * CNN_MID_STRIDE + i] = (float)(normal() * std::sqrt(2.0 / 288.0));
    // a small output layer: the untrained network starts close to the identity
    for (size_t i = 0; i < 9 * 32; i++) p[CNN_OFF_OUT + i] = (float)(normal() * 0.1 * std::sqrt(2.0 / 288.0));
}

// ---------------------------------------------------------------------------
// tile autotuner

// Median of `reps` timed runs of f (after one untimed run), in ms.
template <class F>
float time_ms(ds_ctx* c, int reps, F&& f) {
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess) return -1.0f;
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return -1.0f;
    }
    std::vector<float> t;
    bool ok = f() == hipSuccess;
    for (int i = 0; ok && i < reps; i++) {
        ok = hipEventRecord(a, c->stream) == hipSuccess && f() == hipSuccess && hipEventRecord(b, c->stream) == hipSuccess &&
             hipEventSynchronize(b) == hipSuccess;
        float ms = 0.0f;
        if (ok) ok = hipEventElapsedTime(&ms, a, b) == hipSuccess;
        t.push_back(ms);
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (!ok) return -1.0f;
    std::sort(t.begin(), t.end());
    return t[t.size() / 2];
}

// Times every tile variant of RCD, MHC and the RL iteration on a synthetic
// frame of this shape and keeps the fastest for the shape.
int autotune_shape(ds_ctx* c, size_t w, size_t h, ds_tuning* out) {
    const size_t n = w * h;
    if (n == 0 || w < 16 || h < 16) return fail(c, DS_ERR_INVALID, "autotune needs a frame of at least 16 x 16");
    if (c->capturing) return fail(c, DS_ERR_INVALID, "autotune inside a capture");
    static const uint8_t rggb[4] = {0, 1, 1, 2};
    const float white[3] = {16383.0f, 16383.0f, 16383.0f};
    const float ones[3] = {1.0f, 1.0f, 1.0f};
    uint16_t* raw = nullptr;
    float* rgb = nullptr;
    float* ws = nullptr;
    float* out_rgb = nullptr;
    hipError_t e = hipMalloc(&raw, n * 2);
    if (e == hipSuccess) e = hipMalloc(&rgb, n * 12);
    if (e == hipSuccess) e = hipMalloc(&out_rgb, n * 12);
    if (e == hipSuccess) e = hipMalloc(&ws, rl_workspace_floats(w, h) * 4);
    auto release = [&] {
        (void)hipStreamSynchronize(c->stream);
        if (raw) (void)hipFree(raw);
        if (rgb) (void)hipFree(rgb);
        if (out_rgb) (void)hipFree(out_rgb);
        if (ws) (void)hipFree(ws);
    };
    if (e != hipSuccess) {
        release();
        return hip_fail(c, e, "hipMalloc (autotune)");
    }
    int rc = synth_dev(c, 0x70E0ull, w, h, 14, 4.0f, rggb, raw);
    if (rc) {
        release();
        return rc;
    }
    ds_tuning t{};
    DemosaicArgs a{};
    a.raw_u16 = raw;
    a.stride = w;
    a.w = w;
    a.h = h;
    std::memcpy(a.pattern, rggb, 4);
    a.black = 512.0f;
    for (int k = 0; k < 3; k++) {
        a.white[k] = white[k];
        a.mul[k] = ones[k];
    }
    a.rgb = rgb;
    a.clip = nullptr;
    float best[3] = {1e30f, 1e30f, 1e30f};
    for (int algo = 0; algo < 2; algo++) {
        a.algo = algo == 0 ? DS_ALGO_RCD : DS_ALGO_MHC;
        for (int v = 0; v < demosaic_variants(a.algo) && v < 4; v++) {
            const float ms = time_ms(c, 3, [&] { return launch_demosaic(a, c->stream, v); });
            if (ms < 0.0f) {
                release();
                return fail(c, DS_ERR_LAUNCH, "autotune: demosaic variant %d failed", v);
            }
            (algo == 0 ? t.rcd_ms : t.mhc_ms)[v] = ms;
            if (ms < best[algo]) {
                best[algo] = ms;
                (algo == 0 ? t.rcd_variant : t.mhc_variant) = v;
            }
        }
    }
    for (int v = 0; v < RL_VARIANTS && v < 4; v++) {
        const float ms = time_ms(c, 3, [&] { return launch_raw_details(rgb, w, h, 1.0f, 4, ws, out_rgb, c->stream, v); });
        if (ms < 0.0f) {
            release();
            return fail(c, DS_ERR_LAUNCH, "autotune: RL variant %d failed", v);
        }
        t.rl_ms[v] = ms;
        if (ms < best[2]) {
            best[2] = ms;
            t.rl_variant = v;
        }
    }
    release();
    c->tuning[{w, h}] = Tuned{t.rcd_variant, t.mhc_variant, t.rl_variant};
    if (out) *out = t;
    return DS_OK;
}

// ---------------------------------------------------------------------------
// batched pipeline

struct FramePlan {
    size_t idx;
    bool highlights, denoise, learned, rl, draft, host;
    size_t w, h;  // size of the demosaiced image
    float amount;
    float rl_sigma;
    size_t rl_iters;
};

FramePlan plan_frame(const ds_frame& f, size_t idx, const ds_pipeline_opts& o) {
    FramePlan p{};
    p.idx = idx;
    p.draft = o.quality == DS_QUALITY_DRAFT;
    p.w = p.draft ? f.w / 2 : f.w;
    p.h = p.draft ? f.h / 2 : f.h;
    p.highlights = o.highlights != 0;
    p.host = o.host_frames != 0;
    p.denoise = o.denoise != DS_DENOISE_NONE;
    p.learned = o.denoise == DS_DENOISE_LEARNED;
    p.amount = o.amount;
    p.rl = o.raw_details != 0;
    p.rl_sigma = o.rl_sigma;
    p.rl_iters = o.rl_iters;
    if (o.reference_gating) {
        // Engine::base_for (pipeline/mod.rs:229-252): denoise(img, amount.max(1.0));
        // raw_details(img, 0.75, 8) only at full quality with a long edge > 3200
        p.amount = std::fmax(o.amount, 1.0f);
        p.rl = p.rl && !p.draft && (p.w > p.h ? p.w : p.h) > 3200;
        p.rl_sigma = 0.75f;
        p.rl_iters = 8;
    }
    return p;
}

// Workspace bytes per slot that one frame's stages ask for (mirrors their SLOT calls).
void frame_needs(const ds_frame& f, const FramePlan& p, size_t need[S_COUNT]) {
    const size_t n = p.w * p.h;
    auto up = [&](Slot s, size_t b) { need[s] = std::max(need[s], b); };
    if (p.host) {
        up(S_PIPE_RAW, f.stride * f.h * 2);
        up(S_PIPE_OUT, n * 12);
    }
    if (p.denoise || p.rl) up(S_PIPE_RGB, n * 12);
    if (p.highlights) {
        up(S_PIPE_CLIP, n * 4);
        up(S_HL, HL_FLOATS * 4);
        up(S_WS, highlights_ws_floats(p.w, p.h) * 4);
    }
    if (p.denoise && p.rl) up(S_PIPE_DEN, n * 12);
    if (p.denoise) {
        up(S_L, n * 4);
        up(S_A, n * 4);
        up(S_B, n * 4);
        up(S_LF, n * 4);
        up(S_SIGMA, 64);
        up(S_WS, 12 * n * 4);
        up(S_NOISE, noise_scratch_u32(p.w, p.h) * sizeof(uint32_t));
        if (p.learned) {
            const size_t nreg = nlm_regions(p.w, p.h);
            up(S_REGIONS, (2 * nreg + 64 + nlm_spot_scratch_floats(p.w, p.h)) * 4);
            up(S_X, n * 4);
            up(S_ACT, cnn_fused_act_bytes(1, (int)p.h, (int)p.w));
        }
    }
    if (p.rl) up(S_WS, rl_workspace_floats(p.w, p.h) * 4);
}

int ensure_branches(ds_ctx* c, int k) {
    while ((int)c->branches.size() < k) {
        ds_ctx* b = new ds_ctx();
        b->device = c->device;
        b->ok = c->ok;
        b->parent = c;
        hipError_t e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            delete b;
            return hip_fail(c, e, "hipStreamCreate (batch branch)");
        }
        c->branches.push_back(b);
    }
    return DS_OK;
}

void destroy_graph(GraphEntry& g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    if (g.graph) (void)hipGraphDestroy(g.graph);
    g.exec = nullptr;
    g.graph = nullptr;
}

// Graphs captured against workspace buffers that have since been replaced.
void retire_stale_graphs(ds_ctx* c) {
    std::vector<GraphEntry> keep;
    for (auto& g : c->graphs) {
        if (g.gen == c->slot_gen)
            keep.push_back(g);
        else
            destroy_graph(g);
    }
    c->graphs.swap(keep);
}

int enqueue_stages(ds_ctx* b, const ds_frame& f, const FramePlan& p, const ds_pipeline_opts& o, int* count_dev);

// One frame's stages on branch context b.  Host frames are copied into the
// branch's staging buffers first and their result copied back last, on the
// branch's stream, so one branch's copies overlap the others' kernels.
int enqueue_frame(ds_ctx* b, const ds_frame& hf, const FramePlan& p, const ds_pipeline_opts& o, int* count_dev) {
    ds_frame f = hf;
    if (p.host) {
        SLOT(b, draw, uint16_t*, S_PIPE_RAW, hf.stride * hf.h * 2);
        SLOT(b, dout, float*, S_PIPE_OUT, p.w * p.h * 12);
        HIPCHK(b, hipMemcpyAsync(draw, hf.raw, hf.stride * hf.h * 2, hipMemcpyHostToDevice, b->stream), "H2D (host frame)");
        f.raw = draw;
        f.rgb_out = dout;
    }
    if (int rc = enqueue_stages(b, f, p, o, count_dev)) return rc;
    if (p.host) HIPCHK(b, hipMemcpyAsync(hf.rgb_out, f.rgb_out, p.w * p.h * 12, hipMemcpyDeviceToHost, b->stream), "D2H (host frame)");
    return DS_OK;
}

int enqueue_stages(ds_ctx* b, const ds_frame& f, const FramePlan& p, const ds_pipeline_opts& o, int* count_dev) {
    const bool direct = !p.denoise && !p.rl;
    const size_t w = p.w, h = p.h;
    const float* bs = f.per_site ? f.black_site : nullptr;
    float* rgb = f.rgb_out;
    if (!direct) {
        SLOT(b, t, float*, S_PIPE_RGB, w * h * 12);
        rgb = t;
    }
    float* clip = nullptr;
    if (p.highlights) {
        SLOT(b, t, float*, S_PIPE_CLIP, w * h * 4);
        clip = t;
    }
    // Engine::base (pipeline/mod.rs:184-188): demosaic(cfa, mul, quality)
    int rc = p.draft ? superpixel_dev(b, nullptr, f.raw, f.w, f.h, f.stride, f.pattern, f.black, bs, f.white, f.mul, rgb, clip)
                     : demosaic_dev(b, nullptr, f.raw, f.w, f.h, f.stride, f.pattern, f.black, bs, f.white, f.mul, o.algo, rgb, clip);
    if (rc) return rc;
    if (p.highlights) {
        // rebuild clipped channels at the clip levels of the whole CFA
        SLOT(b, hl, float*, S_HL, HL_FLOATS * 4);
        if ((rc = clip_levels_dev(b, nullptr, f.raw, f.w, f.h, f.stride, f.pattern, f.black, bs, f.white, hl))) return rc;
        if ((rc = highlights_dev(b, rgb, clip, w, h, f.mul, hl))) return rc;
    }
    const float* cur = rgb;
    if (p.denoise) {
        float* dst = f.rgb_out;
        if (p.rl) {
            SLOT(b, t, float*, S_PIPE_DEN, w * h * 12);
            dst = t;
        }
        rc = p.learned ? denoise_learned_dev(b, o.net, rgb, w, h, p.amount, o.search, o.patch, o.threshold_db, dst, nullptr, count_dev)
                       : denoise_dev(b, rgb, w, h, p.amount, o.search, o.patch, dst);
        if (rc) return rc;
        cur = dst;
    }
    if (p.rl) rc = raw_details_dev(b, cur, w, h, p.rl_sigma, p.rl_iters, f.rgb_out);
    return rc;
}

// Page-locked host memory (hipHostMalloc / ds_host_alloc): copies from and to
// it are asynchronous, so they can sit in the captured graph.
bool pinned_host(const void* p) {
    hipPointerAttribute_t a{};
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

int validate_batch(ds_ctx* c, const ds_frame* frames, size_t n, const ds_pipeline_opts& o) {
    if (!frames && n) return fail(c, DS_ERR_INVALID, "frames is null");
    if (o.algo != DS_ALGO_RCD && o.algo != DS_ALGO_MHC) return fail(c, DS_ERR_INVALID, "unknown demosaic algorithm %d", o.algo);
    if (o.denoise < DS_DENOISE_NONE || o.denoise > DS_DENOISE_LEARNED) return fail(c, DS_ERR_INVALID, "unknown denoiser %d", o.denoise);
    if (o.denoise == DS_DENOISE_LEARNED && (!o.net || o.net->ctx != c))
        return fail(c, DS_ERR_INVALID, "learned denoise needs a network of this context");
    if (o.denoise != DS_DENOISE_NONE && (o.search < 0 || o.search > 1024 || o.patch > 64))
        return fail(c, DS_ERR_INVALID, "search %ld / patch %zu out of range", o.search, o.patch);
    if (o.raw_details && (!(o.rl_sigma >= 0.0f) || o.rl_sigma > 1000.0f)) return fail(c, DS_ERR_INVALID, "RL sigma out of range");
    if (o.streams < 0 || o.streams > DS_MAX_BRANCHES) return fail(c, DS_ERR_INVALID, "streams %d out of range", o.streams);
    if (o.quality != DS_QUALITY_FULL && o.quality != DS_QUALITY_DRAFT) return fail(c, DS_ERR_INVALID, "unknown quality %d", o.quality);
    for (size_t i = 0; i < n; i++) {
        const ds_frame& f = frames[i];
        if (int rc = check_dims(c, f.w, f.h)) return rc;
        if (!f.raw || !f.rgb_out || f.w == 0 || f.h == 0) return fail(c, DS_ERR_INVALID, "frame %zu: null buffer or empty", i);
        if (f.stride < f.w) return fail(c, DS_ERR_INVALID, "frame %zu: stride < width", i);
        if (!valid_bayer(f.pattern)) return fail(c, DS_ERR_INVALID, "frame %zu: not a Bayer pattern", i);
        if (o.quality == DS_QUALITY_DRAFT && (f.w < 2 || f.h < 2)) return fail(c, DS_ERR_INVALID, "frame %zu: too small for superpixels", i);
        if (o.host_frames && (!pinned_host(f.raw) || !pinned_host(f.rgb_out)))
            return fail(c, DS_ERR_INVALID, "frame %zu: host frames must be in pinned host memory (ds_host_alloc)", i);
        if (o.denoise == DS_DENOISE_LEARNED) {  // the network's own limits, checked before any capture starts
            const FramePlan p = plan_frame(f, i, o);
            if (int rc = cnn_check_shape(c, 1, p.h, p.w)) {
                c->err = "frame " + std::to_string(i) + ": " + c->err;
                return rc;
            }
        }
    }
    if (o.host_frames != 0 && o.host_frames != 1) return fail(c, DS_ERR_INVALID, "host_frames must be 0 or 1");
    return DS_OK;
}

// Byte image of everything the captured graph depends on.
std::vector<uint8_t> batch_key(ds_ctx* c, const ds_frame* frames, size_t n, const ds_pipeline_opts& o, int k) {
    std::vector<uint8_t> key;
    auto put = [&](const void* p, size_t b) {
        const uint8_t* q = static_cast<const uint8_t*>(p);
        key.insert(key.end(), q, q + b);
    };
    put(&k, sizeof k);
    put(&n, sizeof n);
    put(&o.algo, sizeof o.algo);
    put(&o.denoise, sizeof o.denoise);
    put(&o.amount, sizeof o.amount);
    put(&o.search, sizeof o.search);
    put(&o.patch, sizeof o.patch);
    put(&o.raw_details, sizeof o.raw_details);
    put(&o.rl_sigma, sizeof o.rl_sigma);
    put(&o.rl_iters, sizeof o.rl_iters);
    put(&o.reference_gating, sizeof o.reference_gating);
    // the network by its unique id (a new network may reuse a freed one's address)
    const uint64_t net_uid = o.denoise == DS_DENOISE_LEARNED && o.net ? o.net->uid : 0;
    put(&net_uid, sizeof net_uid);
    put(&o.threshold_db, sizeof o.threshold_db);
    put(&o.highlights, sizeof o.highlights);
    put(&o.quality, sizeof o.quality);
    put(&o.host_frames, sizeof o.host_frames);
    for (size_t i = 0; i < n; i++) {
        const ds_frame& f = frames[i];
        put(&f.raw, sizeof f.raw);
        put(&f.w, sizeof f.w);
        put(&f.h, sizeof f.h);
        put(&f.stride, sizeof f.stride);
        put(f.pattern, 4);
        put(&f.black, sizeof f.black);
        put(f.white, sizeof f.white);
        put(f.mul, sizeof f.mul);
        put(&f.per_site, sizeof f.per_site);
        if (f.per_site) put(f.black_site, sizeof f.black_site);
        put(&f.rgb_out, sizeof f.rgb_out);
        const Tuned t = tuned_for(c, f.w, f.h);
        put(&t, sizeof t);
    }
    return key;
}

// Captures the whole batch into one graph: frames are spread over k branches
// (longest-processing-time order by pixel count), each branch with its own
// stream and workspace, forked from and joined back into the context stream.
int capture_batch(ds_ctx* c, const ds_frame* frames, size_t n, const ds_pipeline_opts& o, int k, int* counts, GraphEntry& g) {
    if (int rc = ensure_branches(c, k)) return rc;
    // assignment
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; i++) order[i] = i;
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return frames[a].w * frames[a].h > frames[b].w * frames[b].h; });
    std::vector<std::vector<size_t>> assign(k);
    std::vector<double> load(k, 0.0);
    for (size_t i : order) {
        int best = 0;
        for (int j = 1; j < k; j++)
            if (load[j] < load[best]) best = j;
        assign[best].push_back(i);
        load[best] += (double)(frames[i].w * frames[i].h);
    }
    // reserve every workspace the captured stages will ask for
    for (int j = 0; j < k; j++) {
        size_t need[S_COUNT] = {};
        for (size_t i : assign[j]) frame_needs(frames[i], plan_frame(frames[i], i, o), need);
        for (int s = 0; s < S_COUNT; s++) {
            if (!need[s]) continue;
            hipError_t e;
            if (!slot(c->branches[j], (Slot)s, need[s], &e)) return hip_fail(c, e, "batch workspace allocation");
        }
    }
    retire_stale_graphs(c);
    if (o.denoise == DS_DENOISE_LEARNED && !o.net->packed) {
        int rc = run(c, "cnn_pack", [&] { return cnn_pack(o.net->params(), o.net->wpk, c->stream); });
        if (rc) return rc;
        o.net->packed = true;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    std::vector<hipEvent_t> ev(k + 1, nullptr);
    for (auto& e : ev) HIPCHK(c, hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
    auto cleanup = [&] {
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e);
        c->capturing = false;
        for (ds_ctx* b : c->branches) b->capturing = false;
    };
    const bool prof = c->prof;
    c->prof = false;
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed);
    if (e != hipSuccess) {
        c->prof = prof;
        cleanup();
        return hip_fail(c, e, "hipStreamBeginCapture");
    }
    c->capturing = true;
    int rc = DS_OK;
    int forked = 0;                  // branches that waited on the fork event
    std::vector<char> joined(k, 0);  // branches joined back into the context stream
    if (hipEventRecord(ev[k], c->stream) != hipSuccess) rc = fail(c, DS_ERR_DEVICE, "fork event");
    for (int j = 0; j < k && !rc; j++) {
        ds_ctx* b = c->branches[j];
        b->capturing = true;
        if (hipStreamWaitEvent(b->stream, ev[k], 0) != hipSuccess) {
            rc = fail(c, DS_ERR_DEVICE, "fork wait");
            break;
        }
        forked = j + 1;
        for (size_t i : assign[j]) {
            const FramePlan p = plan_frame(frames[i], i, o);
            rc = enqueue_frame(b, frames[i], p, o, p.learned ? counts + i : nullptr);
            if (rc) {
                c->err = "frame " + std::to_string(i) + ": " + b->err;
                break;
            }
        }
        if (!rc && hipEventRecord(ev[j], b->stream) != hipSuccess) rc = fail(c, DS_ERR_DEVICE, "join event");
        if (!rc && hipStreamWaitEvent(c->stream, ev[j], 0) != hipSuccess) rc = fail(c, DS_ERR_DEVICE, "join wait");
        if (!rc) joined[j] = 1;
    }
    // a failed capture still joins every branch it forked, so that ending
    // the capture leaves no stream in capture mode; the graph is discarded
    for (int j = 0; j < forked; j++)
        if (!joined[j] && hipEventRecord(ev[j], c->branches[j]->stream) == hipSuccess) (void)hipStreamWaitEvent(c->stream, ev[j], 0);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(c->stream, &graph);
    c->prof = prof;
    cleanup();
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        return hip_fail(c, e, "hipGraphInstantiate");
    }
    g.graph = graph;
    g.exec = exec;
    g.branches = k;
    g.n_frames = n;
    g.gen = c->slot_gen;
    g.net_uid = o.denoise == DS_DENOISE_LEARNED ? o.net->uid : 0;
    return DS_OK;
}

constexpr size_t kMaxGraphs = 4;

// Waits for the batch and fills the learned path's region counts.
int batch_counts(ds_ctx* c, const ds_frame* frames, size_t n, const ds_pipeline_opts& o, const int* counts, ds_batch_info* info) {
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    if (o.denoise == DS_DENOISE_LEARNED) {
        std::vector<int> cnt(n);
        HIPCHK(c, hipMemcpy(cnt.data(), counts, n * sizeof(int), hipMemcpyDeviceToHost), "D2H");
        for (size_t i = 0; i < n; i++) {
            info->fallback_regions += (size_t)cnt[i];
            const FramePlan p = plan_frame(frames[i], i, o);
            info->regions += nlm_regions(p.w, p.h);
        }
    }
    return DS_OK;
}

// Host-frame batches (raw in, RGB out in pinned host memory) run outside the
// graph, as a pipeline over three streams: uploads (copy_in) and downloads
// (copy_out) overlap the kernels of other frames on the context stream.
// Frames go largest first, so that each download (12 B/px at PCIe rate)
// hides under the next frame's kernels and only the smallest frame's download
// is left after the last kernel.  HOST_STAGES device staging buffers per
// direction rotate; events order each buffer's reuse after the work that read
// it.  Every workspace is reserved before the first frame is enqueued.
constexpr int HOST_STAGES = 3;

// A stream with a CU mask always gets a hardware queue of its own, where a
// plain stream may share an in-order queue with the context stream or with
// the other copy stream once the runtime's queues are all taken; a shared
// queue puts each copy between the kernels of other frames.  The mask names
// every CU, so nothing is withheld from kernels on that stream.
hipError_t create_copy_stream(int device, hipStream_t* s) {
    int cus = 0;
    hipError_t e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e != hipSuccess) return e;
    std::vector<uint32_t> mask((size_t)(cus + 31) / 32, 0u);
    for (int i = 0; i < cus; i++) mask[(size_t)i / 32] |= 1u << (i % 32);
    return hipExtStreamCreateWithCUMask(s, (uint32_t)mask.size(), mask.data());
}

int host_pipeline_batch(ds_ctx* c, const ds_frame* frames, size_t n, const ds_pipeline_opts& o, int* counts, ds_batch_info* info) {
    for (hipStream_t* s : {&c->copy_in, &c->copy_out})
        if (!*s) HIPCHK(c, create_copy_stream(c->device, s), "hipExtStreamCreateWithCUMask (copy stream)");
    size_t raw_max = 0, out_max = 0;
    size_t need[S_COUNT] = {};
    std::vector<FramePlan> plans(n);
    std::vector<size_t> order(n);
    for (size_t i = 0; i < n; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return frames[a].w * frames[a].h > frames[b].w * frames[b].h; });
    for (size_t i = 0; i < n; i++) {
        plans[i] = plan_frame(frames[i], i, o);
        frame_needs(frames[i], plans[i], need);
        raw_max = std::max(raw_max, frames[i].stride * frames[i].h * 2);
        out_max = std::max(out_max, plans[i].w * plans[i].h * 12);
    }
    raw_max = (raw_max + 255) & ~(size_t)255;
    out_max = (out_max + 255) & ~(size_t)255;
    need[S_PIPE_RAW] = std::max(need[S_PIPE_RAW], HOST_STAGES * raw_max);
    need[S_PIPE_OUT] = std::max(need[S_PIPE_OUT], HOST_STAGES * out_max);
    for (int s = 0; s < S_COUNT; s++) {
        if (!need[s]) continue;
        hipError_t e;
        if (!slot(c, (Slot)s, need[s], &e)) return hip_fail(c, e, "batch workspace allocation");
    }
    if (o.denoise == DS_DENOISE_LEARNED && !o.net->packed) {
        int rc = run(c, "cnn_pack", [&] { return cnn_pack(o.net->params(), o.net->wpk, c->stream); });
        if (rc) return rc;
        o.net->packed = true;
    }
    uint8_t* st_raw = static_cast<uint8_t*>(c->slots[S_PIPE_RAW].p);
    uint8_t* st_out = static_cast<uint8_t*>(c->slots[S_PIPE_OUT].p);
    hipEvent_t ev[3][HOST_STAGES] = {};  // uploaded, computed, downloaded; per staging buffer
    auto destroy = [&] {
        for (auto& row : ev)
            for (hipEvent_t& e : row)
                if (e) (void)hipEventDestroy(e);
    };
    for (auto& row : ev)
        for (hipEvent_t& e : row)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
                destroy();
                return fail(c, DS_ERR_DEVICE, "hipEventCreate");
            }
    int rc = DS_OK;
    for (size_t j = 0; j < n && !rc; j++) {
        const size_t i = order[j];
        const int k = (int)(j % HOST_STAGES);
        const bool reuse = j >= HOST_STAGES;  // buffer k held an earlier frame
        const FramePlan& p = plans[i];
        ds_frame f = frames[i];
        f.raw = reinterpret_cast<const uint16_t*>(st_raw + k * raw_max);
        f.rgb_out = reinterpret_cast<float*>(st_out + k * out_max);
        const size_t in_bytes = frames[i].stride * frames[i].h * 2, out_bytes = p.w * p.h * 12;
        bool ok = true;
        // upload into buffer k once the kernels of its previous frame have read it
        if (reuse) ok = ok && hipStreamWaitEvent(c->copy_in, ev[1][k], 0) == hipSuccess;
        ok = ok && hipMemcpyAsync(const_cast<uint16_t*>(f.raw), frames[i].raw, in_bytes, hipMemcpyHostToDevice, c->copy_in) == hipSuccess;
        ok = ok && hipEventRecord(ev[0][k], c->copy_in) == hipSuccess;
        // kernels once the upload has landed and the previous frame's download has read buffer k
        ok = ok && hipStreamWaitEvent(c->stream, ev[0][k], 0) == hipSuccess;
        if (reuse) ok = ok && hipStreamWaitEvent(c->stream, ev[2][k], 0) == hipSuccess;
        if (!ok) {
            rc = fail(c, DS_ERR_DEVICE, "frame %zu: upload", i);
            break;
        }
        if ((rc = enqueue_stages(c, f, p, o, p.learned ? counts + i : nullptr))) {
            c->err = "frame " + std::to_string(i) + ": " + c->err;
            break;
        }
        ok = hipEventRecord(ev[1][k], c->stream) == hipSuccess && hipStreamWaitEvent(c->copy_out, ev[1][k], 0) == hipSuccess &&
             hipMemcpyAsync(frames[i].rgb_out, f.rgb_out, out_bytes, hipMemcpyDeviceToHost, c->copy_out) == hipSuccess &&
             hipEventRecord(ev[2][k], c->copy_out) == hipSuccess;
        if (!ok) rc = fail(c, DS_ERR_DEVICE, "frame %zu: download", i);
    }
    // everything enqueued finishes before the events go and the call returns
    const bool synced = hipStreamSynchronize(c->copy_in) == hipSuccess && hipStreamSynchronize(c->stream) == hipSuccess &&
                        hipStreamSynchronize(c->copy_out) == hipSuccess;
    destroy();
    if (rc) return rc;
    if (!synced) return fail(c, DS_ERR_DEVICE, "host-frame batch did not complete");
    if (!info) return DS_OK;
    info->frames = n;
    info->branches = 1;
    return batch_counts(c, frames, n, o, counts, info);
}

int pipeline_batch(ds_ctx* c, const ds_frame* frames, size_t n, const ds_pipeline_opts& o, ds_batch_info* info) {
    if (info) std::memset(info, 0, sizeof(*info));
    if (int rc = validate_batch(c, frames, n, o)) return rc;
    if (n == 0) return DS_OK;
    if (o.autotune)
        for (size_t i = 0; i < n; i++)
            if (!c->tuning.count({frames[i].w, frames[i].h}) && frames[i].w >= 16 && frames[i].h >= 16)
                if (int rc = autotune_shape(c, frames[i].w, frames[i].h, nullptr)) return rc;
    SLOT(c, counts, int*, S_COUNTS, std::max<size_t>(n, 1024) * sizeof(int));
    if (o.host_frames) return host_pipeline_batch(c, frames, n, o, counts, info);
    retire_stale_graphs(c);
    std::vector<int> candidates;
    if (o.streams > 0)
        candidates.push_back(o.streams);
    else if (o.autotune)
        candidates = {1, 2, 4};
    else
        candidates.push_back(2);
    for (int& k : candidates) k = (int)std::min<size_t>((size_t)k, n);
    std::sort(candidates.begin(), candidates.end());
    candidates.erase(std::unique(candidates.begin(), candidates.end()), candidates.end());
    // a cached graph for this exact batch?
    GraphEntry* hit = nullptr;
    for (int k : candidates) {
        const auto key = batch_key(c, frames, n, o, k);
        for (auto& g : c->graphs)
            if (g.key == key) hit = &g;
        if (hit) break;
    }
    double capture_ms = 0.0;
    if (!hit && candidates.size() > 1) {
        // branch-count autotune: capture each candidate, keep the fastest
        GraphEntry best;
        float best_ms = 1e30f;
        for (int k : candidates) {
            GraphEntry g;
            auto t0 = std::chrono::steady_clock::now();
            if (int rc = capture_batch(c, frames, n, o, k, counts, g)) {
                destroy_graph(best);
                return rc;
            }
            capture_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            const float ms = time_ms(c, 1, [&] { return hipGraphLaunch(g.exec, c->stream); });
            if (ms >= 0.0f && ms < best_ms) {
                destroy_graph(best);
                best = g;
                best.key = batch_key(c, frames, n, o, k);
                best_ms = ms;
            } else {
                destroy_graph(g);
            }
        }
        if (!best.exec) return fail(c, DS_ERR_LAUNCH, "batch graph failed to run");
        if (best.gen != c->slot_gen) {  // a later candidate's reservation replaced a buffer: capture the winner again
            const int k = best.branches;
            destroy_graph(best);
            if (int rc = capture_batch(c, frames, n, o, k, counts, best)) return rc;
            best.key = batch_key(c, frames, n, o, k);
        }
        c->graphs.push_back(best);
        hit = &c->graphs.back();
    } else if (!hit) {
        GraphEntry g;
        auto t0 = std::chrono::steady_clock::now();
        if (int rc = capture_batch(c, frames, n, o, candidates[0], counts, g)) return rc;
        capture_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        g.key = batch_key(c, frames, n, o, candidates[0]);
        c->graphs.push_back(g);
        hit = &c->graphs.back();
    } else if (info) {
        info->graph_reused = 1;
    }
    if (c->graphs.size() > kMaxGraphs) {
        // keep the newest; `hit` is the last element or was found among the newest
        const size_t drop = c->graphs.size() - kMaxGraphs;
        const GraphEntry keep = *hit;
        for (size_t i = 0; i < drop; i++)
            if (c->graphs[i].exec != keep.exec) destroy_graph(c->graphs[i]);
        c->graphs.erase(c->graphs.begin(), c->graphs.begin() + (long)drop);
        hit = nullptr;
        for (auto& g : c->graphs)
            if (g.exec == keep.exec) hit = &g;
        if (!hit) {
            c->graphs.push_back(keep);
            hit = &c->graphs.back();
        }
    }
    if (o.denoise == DS_DENOISE_LEARNED && !o.net->packed) {
        int rc = run(c, "cnn_pack", [&] { return cnn_pack(o.net->params(), o.net->wpk, c->stream); });
        if (rc) return rc;
        o.net->packed = true;
    }
    hipGraphExec_t exec = hit->exec;
    int rc = run(c, "pipeline_batch", [&] { return hipGraphLaunch(exec, c->stream); });
    if (rc) return rc;
    if (!info) return DS_OK;
    info->frames = n;
    info->branches = (size_t)hit->branches;
    info->capture_ms = capture_ms;
    return batch_counts(c, frames, n, o, counts, info);
}

int wb_stats_dev(ds_ctx* c, const float* rgb, size_t w, size_t h, const float* n_as, ds_wb_stats* out) {
    const size_t npx = w * h;
    const size_t cnt = (npx + 2) / 3;
    const int nblk = (int)std::min<size_t>(1024, (cnt + WB_NT - 1) / WB_NT);
    SLOT(c, part, double*, S_STATS, ((size_t)(nblk > 0 ? nblk : 1) * 4 + 4) * sizeof(double));
    double* res = part + (size_t)(nblk > 0 ? nblk : 1) * 4;
    double host[4] = {0.0, 0.0, 0.0, 0.0};
    if (npx) {
        int rc = run(c, "wb_stats", [&] {
            hipLaunchKernelGGL(k_wb_partial, dim3(nblk), dim3(WB_NT), 0, c->stream, rgb, npx, n_as[0], n_as[1], n_as[2], part);
            hipLaunchKernelGGL(k_wb_final, dim3(1), dim3(64), 0, c->stream, part, nblk, res);
            return hipGetLastError();
        });
        if (rc) return rc;
        HIPCHK(c, hipMemcpyAsync(host, res, sizeof host, hipMemcpyDeviceToHost, c->stream), "D2H");
        HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    }
    for (int k = 0; k < 3; k++) out->acc[k] = host[k];
    out->n = host[3];
    out->valid = host[3] > 0.0 ? 1 : 0;
    if (out->valid) {  // pipeline/mod.rs:439-441
        const double g = std::exp(host[1] / host[3]);
        out->neutral[0] = (float)(std::exp(host[0] / host[3]) / g);
        out->neutral[1] = 1.0f;
        out->neutral[2] = (float)(std::exp(host[2] / host[3]) / g);
    } else {
        out->neutral[0] = out->neutral[1] = out->neutral[2] = 1.0f;
    }
    return DS_OK;
}


}  // namespace

namespace {
// Upload in0 (and in1), run f(d_in0, d_in1, d_out), download out.
template <class F>
int host_op(ds_ctx* c, const char* name, const float* in0, const float* in1, size_t in_floats, float* out, size_t out_floats, F&& f) {
    if (!in0 || !out) return fail(c, DS_ERR_INVALID, "null pointer");
    if (in_floats == 0 || out_floats == 0) return DS_OK;
    SLOT(c, d0, float*, S_IN, in_floats * 4);
    float* d1 = nullptr;
    if (in1) {
        SLOT(c, dd, float*, S_IN2, in_floats * 4);
        d1 = dd;
    }
    SLOT(c, dout, float*, S_OUT, out_floats * 4);
    HIPCHK(c, hipMemcpyAsync(d0, in0, in_floats * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    if (in1) HIPCHK(c, hipMemcpyAsync(d1, in1, in_floats * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    if (int rc = run(c, name, [&] { return f(d0, d1, dout); })) return rc;
    HIPCHK(c, hipMemcpyAsync(out, dout, out_floats * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}
}  // namespace

// ---------------------------------------------------------------------------
// exported API

extern "C" {

const char* ds_version(void) { return "darkslide-mi355x 0.1 (gfx950)"; }

// ---- device memory of the allocation-free entry point

namespace {

// The radix select's histograms and sample buffers for
// ds_estimate_noise_noalloc: NOISE_SCRATCH_MAX_U32 words (3.2 MB), enough
// for every plane.  Each GPU the library runs on holds its own copy until the
// process ends.  In host code the name is a host-side stand-in: kernels get
// the device address from noalloc_scratch_ptr instead.
static __device__ __align__(256) uint32_t noalloc_scratch[NOISE_SCRATCH_MAX_U32];

std::mutex g_noalloc_mu;
std::map<int, uint32_t*> g_noalloc_ptr;  // device -> address of noalloc_scratch there

// The address of noalloc_scratch on `device`, which must be current; looked
// up once per device.
hipError_t noalloc_scratch_ptr(int device, uint32_t** out) {
    std::lock_guard<std::mutex> g(g_noalloc_mu);
    auto it = g_noalloc_ptr.find(device);
    if (it != g_noalloc_ptr.end()) {
        *out = it->second;
        return hipSuccess;
    }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch));
    if (e != hipSuccess) return e;
    g_noalloc_ptr[device] = (uint32_t*)p;
    *out = (uint32_t*)p;
    return hipSuccess;
}

}  // namespace

int ds_ctx_create(int device, ds_ctx** out) {
    if (!out) return fail(nullptr, DS_ERR_INVALID, "out is null");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return fail(nullptr, DS_ERR_DEVICE, "no HIP device: %s", hipGetErrorString(e));
    if (device < 0 || device >= n) return fail(nullptr, DS_ERR_INVALID, "device %d out of range (%d devices)", device, n);
    e = hipSetDevice(device);
    if (e != hipSuccess) return fail(nullptr, DS_ERR_DEVICE, "hipSetDevice: %s", hipGetErrorString(e));
    ds_ctx* c = new ds_ctx();
    c->device = device;
    e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return fail(nullptr, DS_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    c->ok = build_oklab();
    // set-up of the allocation-free entry point on this device; a failed
    // lookup is retried (and reported) by that entry point's first call
    uint32_t* scr = nullptr;
    (void)noalloc_scratch_ptr(device, &scr);
    *out = c;
    return DS_OK;
}

void ds_ctx_destroy(ds_ctx* c) {
    if (!c) return;
    {
        std::lock_guard<std::recursive_mutex> g(c->mu);
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        for (auto& p : c->pending) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        for (auto& g : c->graphs) destroy_graph(g);
        c->graphs.clear();
        for (ds_ctx* b : c->branches) {
            (void)hipStreamSynchronize(b->stream);
            for (auto& sb : b->slots)
                if (sb.p) (void)hipFree(sb.p);
            (void)hipStreamDestroy(b->stream);
            delete b;
        }
        c->branches.clear();
        for (hipStream_t s : {c->copy_in, c->copy_out})
            if (s) {
                (void)hipStreamSynchronize(s);
                (void)hipStreamDestroy(s);
            }
        for (auto& b : c->slots)
            if (b.p) (void)hipFree(b.p);
        for (void* p : c->user_allocs) (void)hipFree(p);
        for (void* p : c->host_allocs) (void)hipHostFree(p);
        (void)hipStreamDestroy(c->stream);
    }
    delete c;
}

const char* ds_last_error(const ds_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

#define LOCK(c)                                       \
    if (!c) return DS_ERR_INVALID;                    \
    std::lock_guard<std::recursive_mutex> guard_(c->mu); \
    c->err.clear();                                   \
    {                                                 \
        int rc_ = set_device(c);                      \
        if (rc_) return rc_;                          \
    }

static int export_detail_host(ds_ctx* c, float* rgb, size_t w, size_t h, bool superres, int mode, int amount) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    const size_t n = w * h;
    if (!rgb && n) return fail(c, DS_ERR_INVALID, "null pointer");
    if (n == 0) return export_detail_dev(c, rgb, 0, 0, superres, mode, amount);
    SLOT(c, d_rgb, float*, S_OUT, n * 12);
    HIPCHK(c, hipMemcpyAsync(d_rgb, rgb, n * 12, hipMemcpyHostToDevice, c->stream), "H2D");
    if (int rc = export_detail_dev(c, d_rgb, w, h, superres, mode, amount)) return rc;
    HIPCHK(c, hipMemcpyAsync(rgb, d_rgb, n * 12, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

static int export_detail_device(ds_ctx* c, float* rgb, size_t w, size_t h, bool superres, int mode, int amount) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (!rgb && w * h) return fail(c, DS_ERR_INVALID, "null pointer");
    return export_detail_dev(c, rgb, w, h, superres, mode, amount);
}

int ds_dev_alloc(ds_ctx* c, size_t bytes, void** dptr) {
    LOCK(c);
    if (!dptr) return fail(c, DS_ERR_INVALID, "dptr is null");
    HIPCHK(c, hipMalloc(dptr, bytes ? bytes : 1), "hipMalloc");
    c->user_allocs.push_back(*dptr);
    return DS_OK;
}

int ds_dev_free(ds_ctx* c, void* p) {
    LOCK(c);
    for (size_t i = 0; i < c->user_allocs.size(); i++)
        if (c->user_allocs[i] == p) {
            HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            HIPCHK(c, hipFree(p), "hipFree");
            c->user_allocs.erase(c->user_allocs.begin() + (long)i);
            return DS_OK;
        }
    return fail(c, DS_ERR_INVALID, "pointer was not allocated by ds_dev_alloc on this context");
}

int ds_host_alloc(ds_ctx* c, size_t bytes, void** hptr) {
    LOCK(c);
    if (!hptr) return fail(c, DS_ERR_INVALID, "hptr is null");
    HIPCHK(c, hipHostMalloc(hptr, bytes ? bytes : 1, hipHostMallocDefault), "hipHostMalloc");
    c->host_allocs.push_back(*hptr);
    return DS_OK;
}

int ds_host_free(ds_ctx* c, void* p) {
    LOCK(c);
    for (size_t i = 0; i < c->host_allocs.size(); i++)
        if (c->host_allocs[i] == p) {
            HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
            HIPCHK(c, hipHostFree(p), "hipHostFree");
            c->host_allocs.erase(c->host_allocs.begin() + (long)i);
            return DS_OK;
        }
    return fail(c, DS_ERR_INVALID, "pointer was not allocated by ds_host_alloc on this context");
}

int ds_memcpy_h2d(ds_ctx* c, void* dst, const void* src, size_t bytes) {
    LOCK(c);
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream), "hipMemcpy H2D");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_memcpy_d2h(ds_ctx* c, void* dst, const void* src, size_t bytes) {
    LOCK(c);
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream), "hipMemcpy D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_dev_copy(ds_ctx* c, void* dst, const void* src, size_t bytes) {
    LOCK(c);
    if (!dst || !src || (bytes % 16) || ((uintptr_t)dst % 16) || ((uintptr_t)src % 16))
        return fail(c, DS_ERR_INVALID, "ds_dev_copy needs 16-byte aligned pointers and size");
    const size_t n4 = bytes / 16;
    if (n4 == 0) return DS_OK;
    return run(c, "copy", [&] {
        hipLaunchKernelGGL(k_copy, dim3(4096), dim3(256), 0, c->stream, (const float4*)src, (float4*)dst, n4);
        return hipGetLastError();
    });
}

int ds_sync(ds_ctx* c) {
    LOCK(c);
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_demosaic_u16_dev(ds_ctx* c, const uint16_t* raw, size_t w, size_t h, size_t stride, const uint8_t pattern[4], float black,
                        const float white[3], const float mul[3], int algo, float* rgb_out, float* clip_out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (!white) return fail(c, DS_ERR_INVALID, "white is null");
    return demosaic_dev(c, nullptr, raw, w, h, stride, pattern, black, nullptr, white, mul, algo, rgb_out, clip_out);
}

int ds_demosaic_f32_dev(ds_ctx* c, const float* cfa, size_t w, size_t h, const uint8_t pattern[4], const float mul[3], int algo,
                        float* rgb_out, float* clip_out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    return demosaic_dev(c, cfa, nullptr, w, h, w, pattern, 0.0f, nullptr, nullptr, mul, algo, rgb_out, clip_out);
}

int ds_noise_reduction_dev(ds_ctx* c, float* rgb, size_t w, size_t h, const ds_nr_params* nr, const float* noise_field, float scale) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (!nr || (!rgb && w * h)) return fail(c, DS_ERR_INVALID, "null pointer");
    const size_t n = w * h;
    if (n == 0) return DS_OK;
    float fmax = 0.0f;
    if (noise_field) {
        SLOT(c, red, float*, S_HL, HL_FLOATS * 4);
        HIPCHK(c, launch_field_max(noise_field, n, red + HL_PART_OFF, red, c->stream), "field max");
        HIPCHK(c, hipMemcpyAsync(&fmax, red, 4, hipMemcpyDeviceToHost, c->stream), "D2H");
        HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    }
    return noise_reduction_dev(c, rgb, w, h, nr, noise_field, fmax, scale);
}

int ds_sharpen_dev(ds_ctx* c, float* rgb, size_t w, size_t h, const ds_sharpen_params* sh, const float* sharpness_field, float scale) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (!sh || (!rgb && w * h)) return fail(c, DS_ERR_INVALID, "null pointer");
    if (w * h == 0) return DS_OK;
    return sharpen_dev(c, rgb, w, h, sh, sharpness_field, scale);
}

int ds_fast_guided_filter_dev(ds_ctx* c, const float* I, const float* p, size_t w, size_t h, size_t r, float eps, size_t sub, float* out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if ((!I || !p || !out) && w * h) return fail(c, DS_ERR_INVALID, "null pointer");
    return fgf_dev(c, I, p, w, h, r, eps, sub, out);
}

int ds_plane_resize_dev(ds_ctx* c, const float* in, size_t w, size_t h, size_t w2, size_t h2, float* out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (int rc = check_dims(c, w2, h2)) return rc;
    if ((!in || !out) && w2 * h2) return fail(c, DS_ERR_INVALID, "null pointer");
    if (w2 * h2 && w * h == 0) return fail(c, DS_ERR_INVALID, "resize of an empty plane");
    return run(c, "plane_resize", [&] { return launch_plane_resize(in, w, h, w2, h2, out, c->stream); });
}

int ds_cfa_ingest_dev(ds_ctx* c, const ds_raw_cfa* raw, float* cfa_out, uint8_t pattern_out[4]) {
    LOCK(c);
    RawView v{};
    if (int rc = raw_view(c, raw, v)) return rc;
    if (!cfa_out || !pattern_out) return fail(c, DS_ERR_INVALID, "null pointer");
    std::memcpy(pattern_out, v.pat, 4);
    if (v.w == 0 || v.h == 0) return DS_OK;
    return normalize_dev(c, v.raw, v.w, v.h, v.stride, v.pat, 0.0f, v.black_site, v.white, cfa_out);
}

int ds_demosaic_raw_dev(ds_ctx* c, const ds_raw_cfa* raw, const float mul[3], int algo, int quality, float* rgb_out, float* clip_out) {
    LOCK(c);
    RawView v{};
    if (int rc = raw_view(c, raw, v)) return rc;
    return demosaic_raw_dev(c, v, mul, algo, quality, rgb_out, clip_out);
}

int ds_cfa_clip_levels_raw_dev(ds_ctx* c, const ds_raw_cfa* raw, float* levels_out) {
    LOCK(c);
    RawView v{};
    if (int rc = raw_view(c, raw, v)) return rc;
    return clip_levels_dev(c, nullptr, v.raw, v.w, v.h, v.stride, v.pat, 0.0f, v.black_site, v.white, levels_out);
}

int ds_superpixel_f32_dev(ds_ctx* c, const float* cfa, size_t w, size_t h, const uint8_t pattern[4], const float mul[3], float* rgb_out,
                          float* clip_out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    return superpixel_dev(c, cfa, nullptr, w, h, w, pattern, 0.0f, nullptr, nullptr, mul, rgb_out, clip_out);
}

int ds_cfa_clip_levels_f32_dev(ds_ctx* c, const float* cfa, size_t w, size_t h, const uint8_t pattern[4], float* levels_out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    return clip_levels_dev(c, cfa, nullptr, w, h, w, pattern, 0.0f, nullptr, nullptr, levels_out);
}

int ds_cfa_clip_levels_u16_dev(ds_ctx* c, const uint16_t* raw, size_t w, size_t h, size_t stride, const uint8_t pattern[4], float black,
                               const float white[3], float* levels_out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (!white) return fail(c, DS_ERR_INVALID, "white is null");
    return clip_levels_dev(c, nullptr, raw, w, h, stride, pattern, black, nullptr, white, levels_out);
}

int ds_reconstruct_highlights_dev(ds_ctx* c, float* rgb, const float* clip, size_t w, size_t h, const float mul[3], const float* raw_clip) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    return highlights_dev(c, rgb, clip, w, h, mul, raw_clip);
}

int ds_estimate_noise_dev(ds_ctx* c, const float* plane, size_t w, size_t h, float* sigma_out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (!sigma_out || (!plane && w * h)) return fail(c, DS_ERR_INVALID, "null pointer");
    return noise_dev(c, plane, w, h, sigma_out);
}

int ds_estimate_noise_noalloc(ds_ctx* c, const float* plane, size_t w, size_t h, float* sigma_out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (!sigma_out || (!plane && w * h)) return fail(c, DS_ERR_INVALID, "null pointer");
    const size_t need = noise_scratch_u32(w, h);
    if (need > NOISE_SCRATCH_MAX_U32)
        return fail(c, DS_ERR_INVALID, "plane %zux%zu needs %zu scratch words, the static scratch holds %zu", w, h, need,
                    NOISE_SCRATCH_MAX_U32);
    uint32_t* scr = nullptr;
    const hipError_t e = noalloc_scratch_ptr(c->device, &scr);
    if (e != hipSuccess) return hip_fail(c, e, "hipGetSymbolAddress (noise scratch)");
    const hipError_t le = launch_estimate_noise(plane, w, h, scr, sigma_out, c->stream);
    if (le != hipSuccess) return fail(c, DS_ERR_LAUNCH, "estimate_noise: %s", hipGetErrorString(le));
    return DS_OK;
}

int ds_nlm_dev(ds_ctx* c, const float* in, size_t w, size_t h, float h_param, long search, size_t patch, float* out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if ((!in || !out) && w * h) return fail(c, DS_ERR_INVALID, "null pointer");
    return nlm_dev(c, in, w, h, NlmH{h_param, nullptr, 1.0f}, search, patch, out);
}

int ds_denoise_ex_dev(ds_ctx* c, const float* rgb, size_t w, size_t h, float amount, long search, size_t patch, float* rgb_out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if ((!rgb || !rgb_out) && w * h) return fail(c, DS_ERR_INVALID, "null pointer");
    return denoise_dev(c, rgb, w, h, amount, search, patch, rgb_out);
}

int ds_nlm_reference_order_dev(ds_ctx* c, const float* in, size_t w, size_t h, float h_param, long search, size_t patch, int threads,
                               float* out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if ((!in || !out) && w * h) return fail(c, DS_ERR_INVALID, "null pointer");
    return nlm_ref_dev(c, in, w, h, NlmH{h_param, nullptr, 1.0f}, search, patch, threads, out);
}

int ds_denoise_reference_order_dev(ds_ctx* c, const float* rgb, size_t w, size_t h, float amount, long search, size_t patch, int threads,
                                   float* rgb_out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if ((!rgb || !rgb_out) && w * h) return fail(c, DS_ERR_INVALID, "null pointer");
    if (threads < 1) return fail(c, DS_ERR_INVALID, "thread count %d out of range", threads);
    return denoise_dev(c, rgb, w, h, amount, search, patch, rgb_out, threads);
}

int ds_raw_details_dev(ds_ctx* c, const float* rgb, size_t w, size_t h, float sigma, size_t iterations, float* rgb_out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if ((!rgb || !rgb_out) && w * h) return fail(c, DS_ERR_INVALID, "null pointer");
    return raw_details_dev(c, rgb, w, h, sigma, iterations, rgb_out);
}

// ---- synchronous host-pointer entry points

int ds_demosaic_f32(ds_ctx* c, const float* cfa, size_t w, size_t h, const uint8_t pattern[4], const float mul[3], int algo,
                    float* rgb_out, float* clip_out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (!cfa || !rgb_out || !mul) return fail(c, DS_ERR_INVALID, "null pointer");
    const size_t n = w * h;
    if (n == 0) return DS_OK;
    SLOT(c, d_in, float*, S_IN, n * 4);
    SLOT(c, d_rgb, float*, S_OUT, n * 12);
    float* d_clip = nullptr;
    if (clip_out) {
        SLOT(c, dc, float*, S_CLIP, n * 4);
        d_clip = dc;
    }
    HIPCHK(c, hipMemcpyAsync(d_in, cfa, n * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    if (int rc = demosaic_dev(c, d_in, nullptr, w, h, w, pattern, 0.0f, nullptr, nullptr, mul, algo, d_rgb, d_clip)) return rc;
    HIPCHK(c, hipMemcpyAsync(rgb_out, d_rgb, n * 12, hipMemcpyDeviceToHost, c->stream), "D2H");
    if (clip_out) HIPCHK(c, hipMemcpyAsync(clip_out, d_clip, n * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_demosaic_u16(ds_ctx* c, const uint16_t* raw, size_t w, size_t h, size_t stride, const uint8_t pattern[4], float black,
                    const float white[3], const float mul[3], int algo, float* rgb_out, float* clip_out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (!raw || !rgb_out || !mul || !white) return fail(c, DS_ERR_INVALID, "null pointer");
    if (stride < w) return fail(c, DS_ERR_INVALID, "stride %zu < width %zu", stride, w);
    const size_t n = w * h;
    if (n == 0) return DS_OK;
    SLOT(c, d_raw, uint16_t*, S_IN, h * stride * 2);
    SLOT(c, d_rgb, float*, S_OUT, n * 12);
    float* d_clip = nullptr;
    if (clip_out) {
        SLOT(c, dc, float*, S_CLIP, n * 4);
        d_clip = dc;
    }
    HIPCHK(c, hipMemcpyAsync(d_raw, raw, ((h - 1) * stride + w) * 2, hipMemcpyHostToDevice, c->stream), "H2D");
    if (int rc = demosaic_dev(c, nullptr, d_raw, w, h, stride, pattern, black, nullptr, white, mul, algo, d_rgb, d_clip)) return rc;
    HIPCHK(c, hipMemcpyAsync(rgb_out, d_rgb, n * 12, hipMemcpyDeviceToHost, c->stream), "D2H");
    if (clip_out) HIPCHK(c, hipMemcpyAsync(clip_out, d_clip, n * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_noise_reduction(ds_ctx* c, float* rgb, size_t w, size_t h, const ds_nr_params* nr, const float* noise_field, float scale) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (!nr || (!rgb && w * h)) return fail(c, DS_ERR_INVALID, "null pointer");
    const size_t n = w * h;
    if (n == 0) return DS_OK;
    SLOT(c, d_rgb, float*, S_OUT, n * 12);
    float* d_field = nullptr;
    float fmax = 0.0f;
    if (noise_field) {
        for (size_t i = 0; i < n; i++) fmax = std::fmax(fmax, noise_field[i]);  // the fold at detail.rs:34
        SLOT(c, df, float*, S_CLIP, n * 4);
        HIPCHK(c, hipMemcpyAsync(df, noise_field, n * 4, hipMemcpyHostToDevice, c->stream), "H2D");
        d_field = df;
    }
    HIPCHK(c, hipMemcpyAsync(d_rgb, rgb, n * 12, hipMemcpyHostToDevice, c->stream), "H2D");
    if (int rc = noise_reduction_dev(c, d_rgb, w, h, nr, d_field, fmax, scale)) return rc;
    HIPCHK(c, hipMemcpyAsync(rgb, d_rgb, n * 12, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_sharpen(ds_ctx* c, float* rgb, size_t w, size_t h, const ds_sharpen_params* sh, const float* sharpness_field, float scale) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (!sh || (!rgb && w * h)) return fail(c, DS_ERR_INVALID, "null pointer");
    const size_t n = w * h;
    if (n == 0) return DS_OK;
    SLOT(c, d_rgb, float*, S_OUT, n * 12);
    float* d_field = nullptr;
    if (sharpness_field) {
        SLOT(c, df, float*, S_CLIP, n * 4);
        HIPCHK(c, hipMemcpyAsync(df, sharpness_field, n * 4, hipMemcpyHostToDevice, c->stream), "H2D");
        d_field = df;
    }
    HIPCHK(c, hipMemcpyAsync(d_rgb, rgb, n * 12, hipMemcpyHostToDevice, c->stream), "H2D");
    if (int rc = sharpen_dev(c, d_rgb, w, h, sh, d_field, scale)) return rc;
    HIPCHK(c, hipMemcpyAsync(rgb, d_rgb, n * 12, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_fast_guided_filter(ds_ctx* c, const float* I, const float* p, size_t w, size_t h, size_t r, float eps, size_t sub, float* out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if ((!I || !p || !out) && w * h) return fail(c, DS_ERR_INVALID, "null pointer");
    const size_t n = w * h;
    if (n == 0) return DS_OK;
    SLOT(c, dI, float*, S_IN, n * 4);
    SLOT(c, dP, float*, S_IN2, n * 4);
    SLOT(c, dO, float*, S_OUT, n * 4);
    HIPCHK(c, hipMemcpyAsync(dI, I, n * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    HIPCHK(c, hipMemcpyAsync(dP, p, n * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    if (int rc = fgf_dev(c, dI, dP, w, h, r, eps, sub, dO)) return rc;
    HIPCHK(c, hipMemcpyAsync(out, dO, n * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_plane_resize(ds_ctx* c, const float* in, size_t w, size_t h, size_t w2, size_t h2, float* out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (int rc = check_dims(c, w2, h2)) return rc;
    if ((!in || !out) && w2 * h2) return fail(c, DS_ERR_INVALID, "null pointer");
    if (w2 * h2 == 0) return DS_OK;
    if (w * h == 0) return fail(c, DS_ERR_INVALID, "resize of an empty plane");
    SLOT(c, dI, float*, S_IN, w * h * 4);
    SLOT(c, dO, float*, S_OUT, w2 * h2 * 4);
    HIPCHK(c, hipMemcpyAsync(dI, in, w * h * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    HIPCHK(c, launch_plane_resize(dI, w, h, w2, h2, dO, c->stream), "plane_resize");
    HIPCHK(c, hipMemcpyAsync(out, dO, w2 * h2 * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_super_res_refine(ds_ctx* c, float* rgb, size_t w, size_t h) { return export_detail_host(c, rgb, w, h, true, 0, 0); }

int ds_output_sharpen(ds_ctx* c, float* rgb, size_t w, size_t h, int mode, int amount) {
    return export_detail_host(c, rgb, w, h, false, mode, amount);
}

int ds_super_res_refine_dev(ds_ctx* c, float* rgb, size_t w, size_t h) { return export_detail_device(c, rgb, w, h, true, 0, 0); }

int ds_output_sharpen_dev(ds_ctx* c, float* rgb, size_t w, size_t h, int mode, int amount) {
    return export_detail_device(c, rgb, w, h, false, mode, amount);
}

int ds_cfa_ingest(ds_ctx* c, const ds_raw_cfa* raw, float* cfa_out, uint8_t pattern_out[4]) {
    LOCK(c);
    RawView v{};
    if (int rc = upload_raw(c, raw, v)) return rc;
    if (!cfa_out || !pattern_out) return fail(c, DS_ERR_INVALID, "null pointer");
    std::memcpy(pattern_out, v.pat, 4);
    const size_t n = v.w * v.h;
    if (n == 0) return DS_OK;
    SLOT(c, d_out, float*, S_OUT, n * 4);
    if (int rc = normalize_dev(c, v.raw, v.w, v.h, v.stride, v.pat, 0.0f, v.black_site, v.white, d_out)) return rc;
    HIPCHK(c, hipMemcpyAsync(cfa_out, d_out, n * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_demosaic_raw(ds_ctx* c, const ds_raw_cfa* raw, const float mul[3], int algo, int quality, float* rgb_out, float* clip_out) {
    LOCK(c);
    RawView v{};
    if (int rc = upload_raw(c, raw, v)) return rc;
    if (!rgb_out || !mul) return fail(c, DS_ERR_INVALID, "null pointer");
    const bool draft = quality == DS_QUALITY_DRAFT;
    const size_t n = draft ? (v.w / 2) * (v.h / 2) : v.w * v.h;
    if (n == 0) return DS_OK;
    SLOT(c, d_rgb, float*, S_OUT, n * 12);
    float* d_clip = nullptr;
    if (clip_out) {
        SLOT(c, dc, float*, S_CLIP, n * 4);
        d_clip = dc;
    }
    if (int rc = demosaic_raw_dev(c, v, mul, algo, quality, d_rgb, d_clip)) return rc;
    HIPCHK(c, hipMemcpyAsync(rgb_out, d_rgb, n * 12, hipMemcpyDeviceToHost, c->stream), "D2H");
    if (clip_out) HIPCHK(c, hipMemcpyAsync(clip_out, d_clip, n * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_cfa_clip_levels_raw(ds_ctx* c, const ds_raw_cfa* raw, float levels_out[3]) {
    LOCK(c);
    RawView v{};
    if (int rc = upload_raw(c, raw, v)) return rc;
    if (!levels_out) return fail(c, DS_ERR_INVALID, "null pointer");
    SLOT(c, d_lv, float*, S_SIGMA, 64);
    if (int rc = clip_levels_dev(c, nullptr, v.raw, v.w, v.h, v.stride, v.pat, 0.0f, v.black_site, v.white, d_lv)) return rc;
    HIPCHK(c, hipMemcpyAsync(levels_out, d_lv, 12, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_superpixel_f32(ds_ctx* c, const float* cfa, size_t w, size_t h, const uint8_t pattern[4], const float mul[3], float* rgb_out,
                      float* clip_out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (!cfa || !rgb_out || !mul) return fail(c, DS_ERR_INVALID, "null pointer");
    const size_t n = w * h, no = (w / 2) * (h / 2);
    if (no == 0) return DS_OK;
    SLOT(c, d_in, float*, S_IN, n * 4);
    SLOT(c, d_rgb, float*, S_OUT, no * 12);
    float* d_clip = nullptr;
    if (clip_out) {
        SLOT(c, dc, float*, S_CLIP, no * 4);
        d_clip = dc;
    }
    HIPCHK(c, hipMemcpyAsync(d_in, cfa, n * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    if (int rc = superpixel_dev(c, d_in, nullptr, w, h, w, pattern, 0.0f, nullptr, nullptr, mul, d_rgb, d_clip)) return rc;
    HIPCHK(c, hipMemcpyAsync(rgb_out, d_rgb, no * 12, hipMemcpyDeviceToHost, c->stream), "D2H");
    if (clip_out) HIPCHK(c, hipMemcpyAsync(clip_out, d_clip, no * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_cfa_clip_levels_f32(ds_ctx* c, const float* cfa, size_t w, size_t h, const uint8_t pattern[4], float levels_out[3]) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (!cfa || !levels_out) return fail(c, DS_ERR_INVALID, "null pointer");
    const size_t n = w * h;
    SLOT(c, d_in, float*, S_IN, (n ? n : 1) * 4);
    SLOT(c, d_lv, float*, S_SIGMA, 64);
    if (n) HIPCHK(c, hipMemcpyAsync(d_in, cfa, n * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    if (int rc = clip_levels_dev(c, d_in, nullptr, w, h, w, pattern, 0.0f, nullptr, nullptr, d_lv)) return rc;
    HIPCHK(c, hipMemcpyAsync(levels_out, d_lv, 12, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_cfa_clip_levels_u16(ds_ctx* c, const uint16_t* raw, size_t w, size_t h, size_t stride, const uint8_t pattern[4], float black,
                           const float white[3], float levels_out[3]) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (!raw || !levels_out || !white) return fail(c, DS_ERR_INVALID, "null pointer");
    if (stride < w) return fail(c, DS_ERR_INVALID, "stride %zu < width %zu", stride, w);
    const size_t n = h ? (h - 1) * stride + w : 0;
    SLOT(c, d_raw, uint16_t*, S_IN, (n ? n : 1) * 2);
    SLOT(c, d_lv, float*, S_SIGMA, 64);
    if (n) HIPCHK(c, hipMemcpyAsync(d_raw, raw, n * 2, hipMemcpyHostToDevice, c->stream), "H2D");
    if (int rc = clip_levels_dev(c, nullptr, d_raw, w, h, stride, pattern, black, nullptr, white, d_lv)) return rc;
    HIPCHK(c, hipMemcpyAsync(levels_out, d_lv, 12, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_reconstruct_highlights(ds_ctx* c, float* rgb, const float* clip, size_t w, size_t h, const float mul[3], const float raw_clip[3]) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (!rgb || !clip || !mul || !raw_clip) return fail(c, DS_ERR_INVALID, "null pointer");
    const size_t n = w * h;
    if (n == 0) return DS_OK;
    SLOT(c, d_rgb, float*, S_OUT, n * 12);
    SLOT(c, d_clip, float*, S_CLIP, n * 4);
    SLOT(c, d_rc, float*, S_SIGMA, 64);
    HIPCHK(c, hipMemcpyAsync(d_rgb, rgb, n * 12, hipMemcpyHostToDevice, c->stream), "H2D");
    HIPCHK(c, hipMemcpyAsync(d_clip, clip, n * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    HIPCHK(c, hipMemcpyAsync(d_rc, raw_clip, 12, hipMemcpyHostToDevice, c->stream), "H2D");
    if (int rc = highlights_dev(c, d_rgb, d_clip, w, h, mul, d_rc)) return rc;
    HIPCHK(c, hipMemcpyAsync(rgb, d_rgb, n * 12, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_estimate_noise(ds_ctx* c, const float* plane, size_t w, size_t h, float* sigma_out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (!sigma_out || (!plane && w * h)) return fail(c, DS_ERR_INVALID, "null pointer");
    const size_t n = w * h;
    SLOT(c, d_in, float*, S_IN, (n ? n : 1) * 4);
    SLOT(c, d_sig, float*, S_SIGMA, 64);
    if (n) HIPCHK(c, hipMemcpyAsync(d_in, plane, n * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    if (int rc = noise_dev(c, d_in, w, h, d_sig)) return rc;
    HIPCHK(c, hipMemcpyAsync(sigma_out, d_sig, 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_nlm(ds_ctx* c, const float* in, size_t w, size_t h, float h_param, long search, size_t patch, float* out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    const size_t n = w * h;
    if (n == 0) return DS_OK;
    if (!in || !out) return fail(c, DS_ERR_INVALID, "null pointer");
    SLOT(c, d_in, float*, S_IN, n * 4);
    SLOT(c, d_out, float*, S_OUT, n * 4);
    HIPCHK(c, hipMemcpyAsync(d_in, in, n * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    if (int rc = nlm_dev(c, d_in, w, h, NlmH{h_param, nullptr, 1.0f}, search, patch, d_out)) return rc;
    HIPCHK(c, hipMemcpyAsync(out, d_out, n * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_nlm_reference_order(ds_ctx* c, const float* in, size_t w, size_t h, float h_param, long search, size_t patch, int threads,
                           float* out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    const size_t n = w * h;
    if (n == 0) return DS_OK;
    if (!in || !out) return fail(c, DS_ERR_INVALID, "null pointer");
    SLOT(c, d_in, float*, S_IN, n * 4);
    SLOT(c, d_out, float*, S_OUT, n * 4);
    HIPCHK(c, hipMemcpyAsync(d_in, in, n * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    if (int rc = nlm_ref_dev(c, d_in, w, h, NlmH{h_param, nullptr, 1.0f}, search, patch, threads, d_out)) return rc;
    HIPCHK(c, hipMemcpyAsync(out, d_out, n * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_denoise_reference_order(ds_ctx* c, const float* rgb, size_t w, size_t h, float amount, long search, size_t patch, int threads,
                               float* rgb_out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    const size_t n = w * h;
    if (n == 0) return DS_OK;
    if (!rgb || !rgb_out) return fail(c, DS_ERR_INVALID, "null pointer");
    if (threads < 1) return fail(c, DS_ERR_INVALID, "thread count %d out of range", threads);
    SLOT(c, d_in, float*, S_IN, n * 12);
    SLOT(c, d_out, float*, S_OUT, n * 12);
    HIPCHK(c, hipMemcpyAsync(d_in, rgb, n * 12, hipMemcpyHostToDevice, c->stream), "H2D");
    if (int rc = denoise_dev(c, d_in, w, h, amount, search, patch, d_out, threads)) return rc;
    HIPCHK(c, hipMemcpyAsync(rgb_out, d_out, n * 12, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_denoise_ex(ds_ctx* c, const float* rgb, size_t w, size_t h, float amount, long search, size_t patch, float* rgb_out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    const size_t n = w * h;
    if (n == 0) return DS_OK;
    if (!rgb || !rgb_out) return fail(c, DS_ERR_INVALID, "null pointer");
    SLOT(c, d_in, float*, S_IN, n * 12);
    SLOT(c, d_out, float*, S_OUT, n * 12);
    HIPCHK(c, hipMemcpyAsync(d_in, rgb, n * 12, hipMemcpyHostToDevice, c->stream), "H2D");
    if (int rc = denoise_dev(c, d_in, w, h, amount, search, patch, d_out)) return rc;
    HIPCHK(c, hipMemcpyAsync(rgb_out, d_out, n * 12, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_denoise(ds_ctx* c, const float* rgb, size_t w, size_t h, float amount, float* rgb_out) {
    return ds_denoise_ex(c, rgb, w, h, amount, 3, 1, rgb_out);  // enhance.rs:84
}

int ds_raw_details(ds_ctx* c, const float* rgb, size_t w, size_t h, float sigma, size_t iterations, float* rgb_out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    const size_t n = w * h;
    if (n == 0) return DS_OK;
    if (!rgb || !rgb_out) return fail(c, DS_ERR_INVALID, "null pointer");
    SLOT(c, d_in, float*, S_IN, n * 12);
    SLOT(c, d_out, float*, S_OUT, n * 12);
    HIPCHK(c, hipMemcpyAsync(d_in, rgb, n * 12, hipMemcpyHostToDevice, c->stream), "H2D");
    if (int rc = raw_details_dev(c, d_in, w, h, sigma, iterations, d_out)) return rc;
    HIPCHK(c, hipMemcpyAsync(rgb_out, d_out, n * 12, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

// ---- plane primitives (synchronous)


int ds_box_blur(ds_ctx* c, const float* in, size_t w, size_t h, size_t r, float* out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    SLOT(c, ws, float*, S_WS, (w * h ? w * h : 1) * 4);
    return host_op(c, "box_blur", in, nullptr, w * h, out, w * h,
                   [&](float* a, float*, float* o) { return launch_box_blur(a, ws, o, w, h, r, c->stream); });
}

int ds_gaussian_blur(ds_ctx* c, const float* in, size_t w, size_t h, float sigma, float* out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (!(sigma >= 0.0f) || sigma > 1000.0f) return fail(c, DS_ERR_INVALID, "sigma %g out of range", (double)sigma);
    SLOT(c, ws, float*, S_WS, (w * h ? 2 * w * h : 1) * 4);
    return host_op(c, "gaussian_blur", in, nullptr, w * h, out, w * h,
                   [&](float* a, float*, float* o) { return launch_gaussian_blur(a, ws, o, w, h, sigma, c->stream); });
}

int ds_guided_filter(ds_ctx* c, const float* guide, const float* src, size_t w, size_t h, size_t r, float eps, float* out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (!src) return fail(c, DS_ERR_INVALID, "null pointer");
    SLOT(c, ws, float*, S_WS, (w * h ? 8 * w * h : 1) * 4);
    return host_op(c, "guided_filter", guide, src, w * h, out, w * h,
                   [&](float* a, float* b, float* o) { return launch_guided_filter(a, b, w, h, r, eps, ws, o, c->stream); });
}

int ds_oklab_from_prophoto(ds_ctx* c, const float* rgb, size_t n, float* lab_out) {
    LOCK(c);
    return host_op(c, "oklab_fwd", rgb, nullptr, 3 * n, lab_out, 3 * n,
                   [&](float* a, float*, float* o) { return launch_oklab_fwd_interleaved(a, n, c->ok, o, c->stream); });
}

int ds_oklab_to_prophoto(ds_ctx* c, const float* lab, size_t n, float* rgb_out) {
    LOCK(c);
    return host_op(c, "oklab_inv", lab, nullptr, 3 * n, rgb_out, 3 * n,
                   [&](float* a, float*, float* o) { return launch_oklab_inv(a, n, c->ok, o, c->stream); });
}

int ds_luminance(ds_ctx* c, const float* rgb, size_t n, float* out) {
    LOCK(c);
    return host_op(c, "luminance", rgb, nullptr, 3 * n, out, n,
                   [&](float* a, float*, float* o) { return launch_luminance(a, n, o, c->stream); });
}

int ds_libm_hash(ds_ctx* c, int fn, uint64_t* hash_out) {
    LOCK(c);
    if (fn != DS_LIBM_CBRTF && fn != DS_LIBM_EXPF) return fail(c, DS_ERR_INVALID, "unknown function %d", fn);
    if (!hash_out) return fail(c, DS_ERR_INVALID, "null pointer");
    SLOT(c, d_acc, unsigned long long*, S_STATS, 64);
    if (int rc = run(c, "libm_hash", [&] { return launch_libm_hash(fn, d_acc, c->stream); })) return rc;
    unsigned long long v = 0;
    HIPCHK(c, hipMemcpyAsync(&v, d_acc, 8, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    *hash_out = (uint64_t)v;
    return DS_OK;
}

// ---- profiling

int ds_prof_enable(ds_ctx* c, int on) {
    LOCK(c);
    collect(c);
    c->prof = on != 0;
    return DS_OK;
}

int ds_prof_reset(ds_ctx* c) {
    LOCK(c);
    collect(c);
    c->totals.clear();
    return DS_OK;
}

int ds_prof_get(ds_ctx* c, ds_prof_entry* out, int cap) {
    LOCK(c);
    collect(c);
    int i = 0;
    for (auto& kv : c->totals) {
        if (out && i < cap) {
            std::memset(&out[i], 0, sizeof(ds_prof_entry));
            std::strncpy(out[i].name, kv.first.c_str(), sizeof(out[i].name) - 1);
            out[i].total_ms = kv.second.first;
            out[i].launches = kv.second.second;
        }
        i++;
    }
    return i;
}

// ---- synthetic frames

int ds_synth_bayer_host(uint64_t seed, size_t w, size_t h, int bits, float iso, const uint8_t pattern[4], uint16_t* raw_out) {
    if (!raw_out || !valid_bayer(pattern) || (bits != 12 && bits != 14) || !(iso >= 1.0f || iso == 0.0f)) return DS_ERR_INVALID;
    if (w > (1u << 20) || h > (1u << 20)) return DS_ERR_INVALID;
    SynthScene* s = new SynthScene;
    synth_scene_init(s, seed, (int)w, (int)h, bits, iso, pattern);
    for (size_t y = 0; y < h; y++)
        for (size_t x = 0; x < w; x++) raw_out[y * w + x] = synth_dn(*s, (int)x, (int)y);
    delete s;
    return DS_OK;
}

int ds_synth_bayer_dev(ds_ctx* c, uint64_t seed, size_t w, size_t h, int bits, float iso, const uint8_t pattern[4], uint16_t* raw_out) {
    LOCK(c);
    return synth_dev(c, seed, w, h, bits, iso, pattern, raw_out);
}

// ---- learned denoiser

size_t ds_cnn_param_count(void) { return CNN_NPARAM; }

int ds_cnn_create(ds_ctx* c, uint64_t seed, ds_cnn** out) {
    LOCK(c);
    if (!out) return fail(c, DS_ERR_INVALID, "out is null");
    *out = nullptr;
    static std::atomic<uint64_t> next_uid{0};
    ds_cnn* net = new ds_cnn();
    net->ctx = c;
    net->uid = ++next_uid;
    net->seed = seed;
    hipError_t e = hipMalloc(&net->state, 4 * CNN_NPARAM * sizeof(float));
    if (e == hipSuccess) e = hipMalloc(&net->wpk, CNN_WPK * sizeof(uint16_t));
    if (e != hipSuccess) {
        if (net->state) (void)hipFree(net->state);
        delete net;
        return hip_fail(c, e, "hipMalloc (CNN state)");
    }
    std::vector<float> p;
    cnn_init_params(seed, p);
    std::vector<float> zero(3 * CNN_NPARAM, 0.0f);
    e = hipMemcpyAsync(net->state, p.data(), CNN_NPARAM * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(net->state + CNN_NPARAM, zero.data(), 3 * CNN_NPARAM * 4, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
        (void)hipFree(net->state);
        (void)hipFree(net->wpk);
        delete net;
        return hip_fail(c, e, "CNN initialisation");
    }
    *out = net;
    return DS_OK;
}

void ds_cnn_destroy(ds_cnn* net) {
    if (!net) return;
    ds_ctx* c = net->ctx;
    {
        std::lock_guard<std::recursive_mutex> g(c->mu);
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        // batch graphs captured with this network read its buffers: retire them
        std::vector<GraphEntry> keep;
        for (auto& ge : c->graphs) {
            if (ge.net_uid == net->uid)
                destroy_graph(ge);
            else
                keep.push_back(ge);
        }
        c->graphs.swap(keep);
        (void)hipFree(net->state);
        (void)hipFree(net->wpk);
        if (net->tr_act) (void)hipFree(net->tr_act);
        if (net->tr_f) (void)hipFree(net->tr_f);
    }
    delete net;
}

int ds_cnn_get_params(ds_cnn* net, float* params, size_t n) {
    if (!net) return DS_ERR_INVALID;
    ds_ctx* c = net->ctx;
    LOCK(c);
    if (!params || n != CNN_NPARAM) return fail(c, DS_ERR_INVALID, "expected %zu parameters", CNN_NPARAM);
    HIPCHK(c, hipMemcpyAsync(params, net->params(), n * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_cnn_set_params(ds_cnn* net, const float* params, size_t n) {
    if (!net) return DS_ERR_INVALID;
    ds_ctx* c = net->ctx;
    LOCK(c);
    if (!params || n != CNN_NPARAM) return fail(c, DS_ERR_INVALID, "expected %zu parameters", CNN_NPARAM);
    std::vector<float> zero(3 * CNN_NPARAM, 0.0f);
    HIPCHK(c, hipMemcpyAsync(net->params(), params, n * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    HIPCHK(c, hipMemcpyAsync(net->grad(), zero.data(), 3 * CNN_NPARAM * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    net->step = 0;
    net->packed = false;
    return DS_OK;
}

long ds_cnn_step(const ds_cnn* net) { return net ? net->step : -1; }

// Checkpoint file: "DSCNN001", u32 nparam, u32 0, i64 step, u64 seed,
// then params, adam m, adam v (nparam f32 each), little-endian.
int ds_cnn_save(ds_cnn* net, const char* path) {
    if (!net) return DS_ERR_INVALID;
    ds_ctx* c = net->ctx;
    LOCK(c);
    if (!path) return fail(c, DS_ERR_INVALID, "path is null");
    std::vector<float> buf(4 * CNN_NPARAM);
    HIPCHK(c, hipMemcpyAsync(buf.data(), net->state, buf.size() * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    FILE* f = std::fopen(path, "wb");
    if (!f) return fail(c, DS_ERR_INVALID, "cannot open %s for writing", path);
    const char magic[8] = {'D', 'S', 'C', 'N', 'N', '0', '0', '1'};
    const uint32_t hdr[2] = {(uint32_t)CNN_NPARAM, 0u};
    const int64_t step = net->step;
    const uint64_t seed = net->seed;
    bool ok = std::fwrite(magic, 1, 8, f) == 8 && std::fwrite(hdr, 4, 2, f) == 2 && std::fwrite(&step, 8, 1, f) == 1 &&
              std::fwrite(&seed, 8, 1, f) == 1 && std::fwrite(buf.data(), 4, CNN_NPARAM, f) == CNN_NPARAM &&  // params
              std::fwrite(buf.data() + 2 * CNN_NPARAM, 4, 2 * CNN_NPARAM, f) == 2 * CNN_NPARAM;              // m, v
    ok = (std::fclose(f) == 0) && ok;
    return ok ? DS_OK : fail(c, DS_ERR_INVALID, "write to %s failed", path);
}

int ds_cnn_load(ds_cnn* net, const char* path) {
    if (!net) return DS_ERR_INVALID;
    ds_ctx* c = net->ctx;
    LOCK(c);
    if (!path) return fail(c, DS_ERR_INVALID, "path is null");
    FILE* f = std::fopen(path, "rb");
    if (!f) return fail(c, DS_ERR_INVALID, "cannot open %s", path);
    char magic[8];
    uint32_t hdr[2];
    int64_t step = 0;
    uint64_t seed = 0;
    std::vector<float> buf(4 * CNN_NPARAM, 0.0f);
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, "DSCNN001", 8) == 0 && std::fread(hdr, 4, 2, f) == 2 &&
              hdr[0] == CNN_NPARAM && std::fread(&step, 8, 1, f) == 1 && std::fread(&seed, 8, 1, f) == 1 &&
              std::fread(buf.data(), 4, CNN_NPARAM, f) == CNN_NPARAM &&
              std::fread(buf.data() + 2 * CNN_NPARAM, 4, 2 * CNN_NPARAM, f) == 2 * CNN_NPARAM && step >= 0;
    std::fclose(f);
    if (!ok) return fail(c, DS_ERR_INVALID, "%s is not a checkpoint of this network", path);
    HIPCHK(c, hipMemcpyAsync(net->state, buf.data(), buf.size() * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    net->step = (long)step;
    net->seed = seed;
    net->packed = false;
    return DS_OK;
}

int ds_cnn_forward(ds_cnn* net, const float* x, size_t n_img, size_t h, size_t w, float* out) {
    if (!net) return DS_ERR_INVALID;
    ds_ctx* c = net->ctx;
    LOCK(c);
    if (!x || !out) return fail(c, DS_ERR_INVALID, "null pointer");
    if (int rc = cnn_check_batch(net, n_img, h, w, 0)) return rc;
    const size_t npx = n_img * h * w;
    float* dx = net->tr_f;
    float* dout = net->tr_f + 2 * npx;
    HIPCHK(c, hipMemcpyAsync(dx, x, npx * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    if (int rc = cnn_forward_train(net, dx, (int)n_img, (int)h, (int)w, dout)) return rc;
    HIPCHK(c, hipMemcpyAsync(out, dout, npx * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_cnn_infer(ds_cnn* net, const float* x, size_t n_img, size_t h, size_t w, float* out) {
    if (!net) return DS_ERR_INVALID;
    ds_ctx* c = net->ctx;
    LOCK(c);
    if (!x || !out) return fail(c, DS_ERR_INVALID, "null pointer");
    if (int rc = cnn_check_shape(c, n_img, h, w)) return rc;
    const size_t npx = n_img * h * w;
    SLOT(c, dx, float*, S_X, npx * 4);
    SLOT(c, dy, float*, S_T, npx * 4);
    if (int rc = cnn_pack_now(net)) return rc;
    HIPCHK(c, hipMemcpyAsync(dx, x, npx * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    SLOT(c, act, uint16_t*, S_ACT, cnn_fused_act_bytes((int)n_img, (int)h, (int)w));
    if (int rc = run(c, "cnn_forward",
                     [&] { return cnn_forward_fused(net->params(), net->wpk, dx, (int)n_img, (int)h, (int)w, act, dy, c->stream); }))
        return rc;
    HIPCHK(c, hipMemcpyAsync(out, dy, npx * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_cnn_loss_grad(ds_cnn* net, const float* x, const float* tgt, size_t n_img, size_t h, size_t w, size_t margin, double* loss,
                     float* grad) {
    if (!net) return DS_ERR_INVALID;
    ds_ctx* c = net->ctx;
    LOCK(c);
    if (!x || !tgt) return fail(c, DS_ERR_INVALID, "null pointer");
    if (int rc = cnn_check_batch(net, n_img, h, w, margin)) return rc;
    const size_t npx = n_img * h * w;
    HIPCHK(c, hipMemcpyAsync(net->tr_f, x, npx * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    HIPCHK(c, hipMemcpyAsync(net->tr_f + npx, tgt, npx * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    if (int rc = cnn_grad_dev(net, (int)n_img, (int)h, (int)w, (int)margin)) return rc;
    if (grad) HIPCHK(c, hipMemcpyAsync(grad, net->grad(), CNN_NPARAM * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    double l = 0.0;
    if (int rc = cnn_loss_host(net, npx, (int)n_img, (int)h, (int)w, (int)margin, &l)) return rc;
    if (loss) *loss = l;
    return DS_OK;
}

int ds_cnn_train_step_dev(ds_cnn* net, const float* x, const float* tgt, size_t n_img, size_t h, size_t w, size_t margin, float lr,
                          double* loss_out) {
    if (!net) return DS_ERR_INVALID;
    ds_ctx* c = net->ctx;
    LOCK(c);
    if (!x || !tgt) return fail(c, DS_ERR_INVALID, "null pointer");
    if (!(lr > 0.0f) || lr > 1.0f) return fail(c, DS_ERR_INVALID, "learning rate %g out of range", (double)lr);
    if (int rc = cnn_check_batch(net, n_img, h, w, margin)) return rc;
    const size_t npx = n_img * h * w;
    HIPCHK(c, hipMemcpyAsync(net->tr_f, x, npx * 4, hipMemcpyDeviceToDevice, c->stream), "D2D");
    HIPCHK(c, hipMemcpyAsync(net->tr_f + npx, tgt, npx * 4, hipMemcpyDeviceToDevice, c->stream), "D2D");
    if (int rc = cnn_grad_dev(net, (int)n_img, (int)h, (int)w, (int)margin)) return rc;
    if (int rc = cnn_adam_step(net, lr)) return rc;
    if (loss_out) return cnn_loss_host(net, npx, (int)n_img, (int)h, (int)w, (int)margin, loss_out);
    return DS_OK;
}

int ds_cnn_make_pairs_dev(ds_ctx* c, uint64_t seed, size_t frame_w, size_t frame_h, float iso, long search, size_t patch, size_t crop,
                          size_t n_crops, float* x_out, float* t_out) {
    LOCK(c);
    if (int rc = check_dims(c, frame_w, frame_h)) return rc;
    if (!x_out || !t_out) return fail(c, DS_ERR_INVALID, "null pointer");
    if (search < 0 || search > 1024 || patch > 64) return fail(c, DS_ERR_INVALID, "search %ld / patch %zu out of range", search, patch);
    return make_pairs_dev(c, seed, frame_w, frame_h, iso, search, patch, crop, n_crops, x_out, t_out);
}

int ds_cnn_denoise_plane(ds_cnn* net, const float* L, size_t w, size_t h, float* lf_out) {
    if (!net) return DS_ERR_INVALID;
    ds_ctx* c = net->ctx;
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    const size_t n = w * h;
    if (n == 0) return DS_OK;
    if (!L || !lf_out) return fail(c, DS_ERR_INVALID, "null pointer");
    SLOT(c, dL, float*, S_IN, n * 4);
    SLOT(c, dLF, float*, S_OUT, n * 4);
    SLOT(c, sig, float*, S_SIGMA, 64);
    HIPCHK(c, hipMemcpyAsync(dL, L, n * 4, hipMemcpyHostToDevice, c->stream), "H2D");
    if (int rc = noise_dev(c, dL, w, h, sig)) return rc;
    if (int rc = cnn_plane_dev(c, net, dL, w, h, sig, dLF)) return rc;
    HIPCHK(c, hipMemcpyAsync(lf_out, dLF, n * 4, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}

int ds_denoise_learned_dev(ds_ctx* c, ds_cnn* net, const float* rgb, size_t w, size_t h, float amount, long search, size_t patch,
                           float threshold_db, float* rgb_out, ds_learned_stats* stats) {
    LOCK(c);
    if (!net || net->ctx != c) return fail(c, DS_ERR_INVALID, "network does not belong to this context");
    if (int rc = check_dims(c, w, h)) return rc;
    if ((!rgb || !rgb_out) && w * h) return fail(c, DS_ERR_INVALID, "null pointer");
    return denoise_learned_dev(c, net, rgb, w, h, amount, search, patch, threshold_db, rgb_out, stats);
}

int ds_denoise_learned(ds_ctx* c, ds_cnn* net, const float* rgb, size_t w, size_t h, float amount, long search, size_t patch,
                       float threshold_db, float* rgb_out, ds_learned_stats* stats) {
    LOCK(c);
    if (!net || net->ctx != c) return fail(c, DS_ERR_INVALID, "network does not belong to this context");
    if (int rc = check_dims(c, w, h)) return rc;
    const size_t n = w * h;
    if (stats) std::memset(stats, 0, sizeof(*stats));
    if (n == 0) return DS_OK;
    if (!rgb || !rgb_out) return fail(c, DS_ERR_INVALID, "null pointer");
    SLOT(c, d_in, float*, S_IN, n * 12);
    SLOT(c, d_out, float*, S_OUT, n * 12);
    HIPCHK(c, hipMemcpyAsync(d_in, rgb, n * 12, hipMemcpyHostToDevice, c->stream), "H2D");
    if (int rc = denoise_learned_dev(c, net, d_in, w, h, amount, search, patch, threshold_db, d_out, stats)) return rc;
    HIPCHK(c, hipMemcpyAsync(rgb_out, d_out, n * 12, hipMemcpyDeviceToHost, c->stream), "D2H");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return DS_OK;
}


// ---- batched export, autotuner, statistics

int ds_pipeline_batch(ds_ctx* c, const ds_frame* frames, size_t n, const ds_pipeline_opts* opts, ds_batch_info* info) {
    LOCK(c);
    if (!opts) return fail(c, DS_ERR_INVALID, "opts is null");
    return pipeline_batch(c, frames, n, *opts, info);
}

int ds_autotune(ds_ctx* c, size_t w, size_t h, ds_tuning* out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    return autotune_shape(c, w, h, out);
}

int ds_set_tuning(ds_ctx* c, size_t w, size_t h, const ds_tuning* t) {
    LOCK(c);
    if (!t) return fail(c, DS_ERR_INVALID, "tuning is null");
    if (t->rcd_variant < 0 || t->rcd_variant >= demosaic_variants(DS_ALGO_RCD) || t->mhc_variant < 0 ||
        t->mhc_variant >= demosaic_variants(DS_ALGO_MHC) || t->rl_variant < 0 || t->rl_variant >= RL_VARIANTS)
        return fail(c, DS_ERR_INVALID, "tile variant out of range");
    c->tuning[{w, h}] = Tuned{t->rcd_variant, t->mhc_variant, t->rl_variant};
    return DS_OK;
}

int ds_get_tuning(ds_ctx* c, size_t w, size_t h, ds_tuning* out) {
    LOCK(c);
    if (!out) return fail(c, DS_ERR_INVALID, "out is null");
    std::memset(out, 0, sizeof(*out));
    const Tuned t = tuned_for(c, w, h);
    out->rcd_variant = t.rcd;
    out->mhc_variant = t.mhc;
    out->rl_variant = t.rl;
    return DS_OK;
}

int ds_auto_wb_stats_dev(ds_ctx* c, const float* rgb, size_t w, size_t h, const float as_shot_neutral[3], ds_wb_stats* out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (!out || !as_shot_neutral || (!rgb && w * h)) return fail(c, DS_ERR_INVALID, "null pointer");
    return wb_stats_dev(c, rgb, w, h, as_shot_neutral, out);
}

int ds_auto_wb_stats(ds_ctx* c, const float* rgb, size_t w, size_t h, const float as_shot_neutral[3], ds_wb_stats* out) {
    LOCK(c);
    if (int rc = check_dims(c, w, h)) return rc;
    if (!out || !as_shot_neutral || (!rgb && w * h)) return fail(c, DS_ERR_INVALID, "null pointer");
    const size_t n = w * h;
    float* d = nullptr;
    if (n) {
        SLOT(c, dd, float*, S_IN, n * 12);
        d = dd;
        HIPCHK(c, hipMemcpyAsync(d, rgb, n * 12, hipMemcpyHostToDevice, c->stream), "H2D");
    }
    return wb_stats_dev(c, d, w, h, as_shot_neutral, out);
}

}  // extern "C"
