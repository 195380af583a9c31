// Bayer demosaic for gfx950: RCD (demosaic.rs:145-344) with MHC frame
// borders (demosaic.rs:89-124), the clip plane (demosaic.rs:29-42) and the
// CFA normalisation of decode.rs:291-306 fused into one tile kernel.
//
// One workgroup owns a TX x TY output tile.  It loads cfa*mul for the tile
// plus a 10-pixel halo (the dependency radius of RCD's last step on its
// input) into LDS with clamped coordinates, then runs every RCD stage on
// LDS, each over exactly the halo the later stages read.  Every stage keeps
// the reference's per-stage frame border, so a pixel gets the same
// computation (RCD or MHC) it gets in the reference.  Arithmetic keeps the
// reference's operation order and is compiled without FMA contraction.
//
// Tile-local coordinates (tx, ty) run from -halo to T+halo; the global
// pixel is (ox + tx, oy + ty).  LDS slot (tx, ty) of the input holds the
// value at the clamped global coordinate, which is what MHC's clamp-to-edge
// addressing reads.

#include <type_traits>

#include "common.h"
#include "demosaic.h"

#pragma clang fp contract(off)

namespace ds {

namespace {

constexpr float kClip = 0.995f;  // demosaic.rs:17
constexpr float kEps = 1e-5f;    // demosaic.rs:155
constexpr int kBorder = 4;       // demosaic.rs:158

// LDS pointer whose value the optimiser cannot see through (an empty asm
// statement claims to change it).  Reads at non-negative constant offsets
// from it compile to that one base register plus the LDS instruction's
// unsigned offset field, where a base of the optimiser's own choosing can sit
// inside a stencil and cost an address add per read on its other side.
template <class T>
using lds_cptr = const __attribute__((address_space(3))) T*;

template <class T>
__device__ __forceinline__ lds_cptr<T> lds_opaque(const T* p) {
    uint32_t a = (uint32_t)(uintptr_t)(lds_cptr<T>)p;
    asm("" : "+v"(a));
    return (lds_cptr<T>)(uintptr_t)a;
}

template <int TX, int TY, int H>
struct Tile {
    static constexpr int W = TX + 2 * H;
    static constexpr int R = TY + 2 * H;
    static constexpr int N = W * R;
    float* p;
    __device__ __forceinline__ float& operator()(int x, int y) const { return p[(y + H) * W + (x + H)]; }
    // stencil at (x, y): s(dx, dy) reads (x + dx, y + dy), dx, dy >= 0
    struct Sten {
        lds_cptr<float> b;
        __device__ __forceinline__ float operator()(int dx, int dy) const { return b[dy * W + dx]; }
    };
    __device__ __forceinline__ Sten sten(int x, int y) const { return Sten{lds_opaque(&(*this)(x, y))}; }
};

struct CfaSrc {
    // exactly one of f32 / u16 is set
    const float* f32;
    const uint16_t* u16;
    size_t stride;  // elements per row
    float blk[4];   // black level per CFA site, (y & 1) * 2 + (x & 1)
    float white[3];
    float mul[3];
    Bayer pat;
    int w, h;

    // white[c], mul[c] by selects: a runtime index into the arrays would put
    // this struct in scratch memory
    __device__ __forceinline__ float white_of(int c) const { return c == 0 ? white[0] : (c == 1 ? white[1] : white[2]); }
    __device__ __forceinline__ float mul_of(int c) const { return c == 0 ? mul[0] : (c == 1 ? mul[1] : mul[2]); }

    // normalised CFA value at an in-image pixel: decode.rs:300-303, with the
    // black level black_at(x, y) of a table of period <= 2
    __device__ __forceinline__ float cfa(int x, int y, int c) const {
        if (f32) return f32[(size_t)y * stride + x];
        const float raw = (float)u16[(size_t)y * stride + x];
        const float b = (y & 1) ? ((x & 1) ? blk[3] : blk[2]) : ((x & 1) ? blk[1] : blk[0]);
        return (raw - b) / fmaxr(white_of(c) - b, 1.0f);
    }
};

// Malvar-He-Cutler at tile position (x, y) for output channel ch, reading
// cm = cfa*mul (not floored) with clamp-to-edge via the clamped LDS tile.
// demosaic.rs:99-123, same operand order.
template <class T>
__device__ __forceinline__ float mhc_at(const T& cm, int x, int y, int col, int row_c, int ch) {
    auto p = [&](int dx, int dy) { return cm(x + dx, y + dy); };
    float v = cm(x, y);
    float out;
    if (col == ch) {
        out = v;
    } else if (col == 1) {
        // R/B at G: horizontal neighbours carry row_c
        bool want_row = (ch == row_c);
        if (want_row) {
            out = (5.0f * v + 4.0f * (p(-1, 0) + p(1, 0)) - (p(-2, 0) + p(2, 0) + p(-1, -1) + p(1, -1) + p(-1, 1) + p(1, 1)) +
                   0.5f * (p(0, -2) + p(0, 2))) /
                  8.0f;
        } else {
            out = (5.0f * v + 4.0f * (p(0, -1) + p(0, 1)) - (p(0, -2) + p(0, 2) + p(-1, -1) + p(1, -1) + p(-1, 1) + p(1, 1)) +
                   0.5f * (p(-2, 0) + p(2, 0))) /
                  8.0f;
        }
    } else if (ch == 1) {
        out = (4.0f * v + 2.0f * (p(0, -1) + p(-1, 0) + p(1, 0) + p(0, 1)) - (p(0, -2) + p(-2, 0) + p(2, 0) + p(0, 2))) / 8.0f;
    } else {
        out = (6.0f * v + 2.0f * (p(-1, -1) + p(1, -1) + p(-1, 1) + p(1, 1)) - 1.5f * (p(0, -2) + p(-2, 0) + p(2, 0) + p(0, 2))) / 8.0f;
    }
    return fmaxr(out, 0.0f);
}

__device__ __forceinline__ float sq(float v) { return v * v; }

// disc at (x, y) of a plane, through its stencil s based at (x - 1, y - 1)
template <class S>
__device__ __forceinline__ float disc(const S& s) {
    float central = s(1, 1);
    float neigh = 0.25f * (s(0, 0) + s(2, 0) + s(0, 2) + s(2, 2));
    return fabsf(0.5f - central) < fabsf(0.5f - neigh) ? neigh : central;
}

// Clip flags of a TX x TY tile (demosaic.rs:29-42): byte (ty + 1) * FW +
// (tx + 1) holds cfa >= 0.995 at the clamped pixel (tx, ty), tx in [-1, TX],
// ty in [-1, TY].  Rows are padded to TX + 4 bytes, so the 6 flags a run of
// 4 outputs needs from one row are two aligned 32-bit words.
template <int TX, int TY>
struct ClipFlags {
    static constexpr int FW = TX + 4, N = (TY + 2) * FW;
    static_assert(TX % 4 == 0, "runs of 4 outputs");
    __device__ static __forceinline__ int at(int tx, int ty) { return (ty + 1) * FW + (tx + 1); }
};

// The clip plane of a tile: clip = 1 where a flag in the clamped 3 x 3
// neighbourhood is set, else 0.  A thread takes 4 consecutive outputs of a
// row: the three flag rows' bytes are ORed as one 48-bit word, the 3-wide
// horizontal OR is two shifts, and the row goes out as one 16-byte store
// where the plane allows.
template <int TX, int TY, int NT>
__device__ __forceinline__ void store_clip(const uint8_t* __restrict__ flag, float* __restrict__ clip, int ox, int oy, int w, int h) {
    constexpr int FW = ClipFlags<TX, TY>::FW, NQ = TX / 4;
    const bool vec = (w & 3) == 0 && (reinterpret_cast<uintptr_t>(clip) & 15) == 0;
    for (int t = threadIdx.x; t < TY * NQ; t += NT) {
        const int y = t / NQ, x0 = 4 * (t % NQ);
        const int gy = oy + y, gx = ox + x0;
        if (gy >= h || gx >= w) continue;
        uint64_t v = 0;
#pragma unroll
        for (int dy = 0; dy < 3; dy++) {
            // flags of columns x0 - 1 .. x0 + 4 (bytes 0..5; 6, 7 are the next two)
            const uint32_t* r = reinterpret_cast<const uint32_t*>(flag + (y + dy) * FW + x0);
            v |= (uint64_t)r[0] | ((uint64_t)r[1] << 32);
        }
        v &= 0xffffffffffffull;
        const uint32_t m = (uint32_t)(v | (v >> 8) | (v >> 16));  // byte i: output x0 + i, 0 or 1
        const float o0 = (float)(m & 0xffu), o1 = (float)((m >> 8) & 0xffu), o2 = (float)((m >> 16) & 0xffu), o3 = (float)(m >> 24);
        float* dst = clip + (size_t)gy * (size_t)w + (size_t)gx;
        if (vec && gx + 4 <= w) {
            *reinterpret_cast<float4*>(dst) = make_float4(o0, o1, o2, o3);
        } else {
            dst[0] = o0;
            if (gx + 1 < w) dst[1] = o1;
            if (gx + 2 < w) dst[2] = o2;
            if (gx + 3 < w) dst[3] = o3;
        }
    }
}

// Halo of each stage, in pixels around the output tile.
constexpr int H_CM = 10;   // cfa*mul
constexpr int H_BVH = 7;   // bv, bh   (later bp, bq at 5)
constexpr int H_VH = 6;    // vh
constexpr int H_LPF = 7;   // lpf
constexpr int H_G = 5;     // green
constexpr int H_PQ = 4;    // pq
constexpr int H_RB = 3;    // r, b after step 4.3
constexpr int H_FLAG = 1;  // clip flags

// LDS of one RCD tile, in three regions reused as planes die (stage numbers
// as in rcd_body):
//   cm, flag       stages 0-6
//   b: bv | bh     stages 1-2, then g | vd | pq (g, vd 3-6; pq 4-5)
//   c: lpf | vh    stages 1-3, then bp | bq (3b-4), then r43 | b43 (5-6)
// vd holds disc(vh) at the tile's G sites, so vh can die after stage 3.
// Two 64 x 32 tiles fit in one CU's LDS.
template <int TX, int TY>
struct RcdLds {
    static constexpr int mx(int a, int b) { return a > b ? a : b; }
    static constexpr int NB = mx(2 * Tile<TX, TY, H_BVH>::N, Tile<TX, TY, H_G>::N + TX * TY + Tile<TX, TY, H_PQ>::N);
    static constexpr int NC = mx(Tile<TX, TY, H_LPF>::N + Tile<TX, TY, H_VH>::N, mx(2 * Tile<TX, TY, H_G>::N, 2 * Tile<TX, TY, H_RB>::N));
    float cm[Tile<TX, TY, H_CM>::N];
    float b[NB];
    float c[NC];
    alignas(4) uint8_t flag[ClipFlags<TX, TY>::N];
};

// Tiles whose 10-pixel halo lies inside [4, w-4) x [4, h-4) (INTERIOR) take
// every stage's RCD branch, so their body drops the per-pixel border checks.
//
// Every stage walks only the sites whose values a later stage reads: lpf,
// bp / bq and pq are read at R/B sites alone (lpf two pixels away along a
// row or column, bp / bq and pq diagonally, all R/B sites again), so they
// are computed there alone; the output stage runs the G sites (R and B
// estimated) and the R/B sites (values copied) as two loops, so no wave
// carries both paths.  Sites of a plane that no stage computes are never
// read.  Stencil reads go through an opaque base at the stencil's top-left
// corner (Tile::sten).
template <int TX, int TY, int NT, bool INTERIOR>
__device__ __forceinline__ void rcd_body(RcdLds<TX, TY>& L, const CfaSrc& src, float* __restrict__ rgb, float* __restrict__ clip) {
    const int tid = threadIdx.x;
    const int ox = blockIdx.x * TX, oy = blockIdx.y * TY;
    const int w = src.w, h = src.h;
    const Bayer pat = src.pat;
    // R/B sites: ((ox + tx) + (oy + ty)) & 1 == rbp, and ox, oy are even
    const int rbp = pat.rb_parity();

    Tile<TX, TY, H_CM> cm{L.cm};
    Tile<TX, TY, H_BVH> bv{L.b}, bh{L.b + Tile<TX, TY, H_BVH>::N};
    Tile<TX, TY, H_G> gg{L.b};
    float* const vdt = L.b + Tile<TX, TY, H_G>::N;  // TX x TY, G sites only
    Tile<TX, TY, H_PQ> pq{L.b + Tile<TX, TY, H_G>::N + TX * TY};
    Tile<TX, TY, H_LPF> lpf{L.c};
    Tile<TX, TY, H_VH> vh{L.c + Tile<TX, TY, H_LPF>::N};
    Tile<TX, TY, H_G> bp{L.c}, bq{L.c + Tile<TX, TY, H_G>::N};
    Tile<TX, TY, H_RB> r43{L.c}, b43{L.c + Tile<TX, TY, H_RB>::N};

    auto in_img = [&](int gx, int gy) { return INTERIOR || (gx >= 0 && gx < w && gy >= 0 && gy < h); };
    auto in_box = [&](int gx, int gy, int b) { return INTERIOR || (gx >= b && gx < w - b && gy >= b && gy < h - b); };
    // RCD reads c = max(cm, 0); MHC (frame-border pixels only) reads cm.  An
    // interior tile never runs MHC, so it stores c itself.
    auto cv = [&](float v) { return INTERIOR ? v : fmaxr(v, 0.0f); };
    auto cval = [&](int x, int y) { return cv(cm(x, y)); };
    // c around (x, y) through a base at (x - R, y - R): c(dx, dy), |dx|, |dy| <= R
    auto c_around = [&](int x, int y, auto r_tag) {
        constexpr int R = decltype(r_tag)::value;
        const auto s = cm.sten(x - R, y - R);
        return [=](int dx, int dy) { return cv(s(dx + R, dy + R)); };
    };
    // the i-th site of a plane with halo HH walked on the R/B checkerboard
    // (G = false) or the G checkerboard (G = true), half_w sites per row.
    // Consecutive items alternate between two rows at the same column pair:
    // the rows' sites then sit an odd number of words apart (the row length
    // is even), so a half-wave's 32 lanes hit 32 distinct LDS banks, where
    // 32 sites of one row, two words apart, would share 16.
    auto site = [&](int i, int HH, int half_w, bool G, int& tx, int& ty) {
        const int j = i >> 1;
        ty = 2 * (j / half_w) + (i & 1) - HH;
        tx = -HH + 2 * (j % half_w) + ((rbp + HH + ty + (G ? 1 : 0)) & 1);
    };

    // ---- stage 0: load cfa*mul (and clip flags)
    if constexpr (INTERIOR) {
        // No clamps.  A thread takes a pair of columns, an even and an odd
        // one (ox and H_CM are even), so the pair's two CFA sites and their
        // constants follow from the row's parity alone; the normalisation is
        // CfaSrc::cfa's expression with those constants.
        constexpr int CW = Tile<TX, TY, H_CM>::W, PW = CW / 2, NP = PW * Tile<TX, TY, H_CM>::R, K = (NP + NT - 1) / NT;
        static_assert(CW % 2 == 0 && H_CM % 2 == 0, "column pairs start at even columns");
        const size_t stride = src.stride;
        const bool f32 = src.f32 != nullptr;
        float r0[K], r1[K];
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int j = tid + k * NT;
            r0[k] = r1[k] = 0.0f;
            if (j < NP) {
                const size_t at = (size_t)(oy - H_CM + j / PW) * stride + (size_t)(ox - H_CM + 2 * (j % PW));
                if (f32) {
                    r0[k] = src.f32[at];
                    r1[k] = src.f32[at + 1];
                } else {
                    r0[k] = (float)src.u16[at];
                    r1[k] = (float)src.u16[at + 1];
                }
            }
        }
        // site s = (y & 1) * 2 + (x & 1): colour, black, denominator, multiplier
        const int c0 = pat.at(0, 0), c1 = pat.at(1, 0), c2 = pat.at(0, 1), c3 = pat.at(1, 1);
        const float b0 = src.blk[0], b1 = src.blk[1], b2 = src.blk[2], b3 = src.blk[3];
        const float d0 = fmaxr(src.white_of(c0) - b0, 1.0f), d1 = fmaxr(src.white_of(c1) - b1, 1.0f);
        const float d2 = fmaxr(src.white_of(c2) - b2, 1.0f), d3 = fmaxr(src.white_of(c3) - b3, 1.0f);
        const float m0 = src.mul_of(c0), m1 = src.mul_of(c1), m2 = src.mul_of(c2), m3 = src.mul_of(c3);
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int j = tid + k * NT;
            if (j >= NP) continue;
            const int ty = j / PW - H_CM, tx = 2 * (j % PW) - H_CM;
            const bool odd = (ty & 1) != 0;
            float v0 = r0[k], v1 = r1[k];
            if (!f32) {
                v0 = (v0 - (odd ? b2 : b0)) / (odd ? d2 : d0);
                v1 = (v1 - (odd ? b3 : b1)) / (odd ? d3 : d1);
            }
            cm(tx, ty) = fmaxr(v0 * (odd ? m2 : m0), 0.0f);
            cm(tx + 1, ty) = fmaxr(v1 * (odd ? m3 : m1), 0.0f);
            if (clip && ty >= -H_FLAG && ty < TY + H_FLAG) {
                uint8_t* f = L.flag + ClipFlags<TX, TY>::at(tx, ty);
                if (tx >= -H_FLAG && tx < TX + H_FLAG) f[0] = v0 >= kClip ? 1 : 0;
                if (tx + 1 >= -H_FLAG && tx + 1 < TX + H_FLAG) f[1] = v1 >= kClip ? 1 : 0;
            }
        }
    } else {
        // clamped coordinates
        constexpr int CW = Tile<TX, TY, H_CM>::W;
        stage_tile<Tile<TX, TY, H_CM>::N, NT>(
            [&](int i) {
                int gx = clampi(ox + i % CW - H_CM, 0, w - 1), gy = clampi(oy + i / CW - H_CM, 0, h - 1);
                return src.cfa(gx, gy, pat.at(gx, gy));
            },
            [&](int i, float v) {
                int tx = i % CW - H_CM, ty = i / CW - H_CM;
                int gx = clampi(ox + tx, 0, w - 1), gy = clampi(oy + ty, 0, h - 1);
                const float m = v * src.mul_of(pat.at(gx, gy));
                cm(tx, ty) = INTERIOR ? fmaxr(m, 0.0f) : m;
                if (clip && tx >= -H_FLAG && tx < TX + H_FLAG && ty >= -H_FLAG && ty < TY + H_FLAG)
                    L.flag[ClipFlags<TX, TY>::at(tx, ty)] = v >= kClip ? 1 : 0;
            });
    }
    __syncthreads();

    // ---- stage 1: bv / bh (demosaic.rs:160-172) over halo 7, every site
    for (int i = tid; i < Tile<TX, TY, H_BVH>::N; i += NT) {
        int tx = i % Tile<TX, TY, H_BVH>::W - H_BVH, ty = i / Tile<TX, TY, H_BVH>::W - H_BVH;
        int gx = ox + tx, gy = oy + ty;
        float v = 0.0f, hz = 0.0f;
        if (in_box(gx, gy, 3)) {
            const auto c = c_around(tx, ty, std::integral_constant<int, 3>{});
            v = sq((c(0, -3) - c(0, -1) - c(0, 1) + c(0, 3)) - 3.0f * (c(0, -2) + c(0, 2)) + 6.0f * c(0, 0));
            hz = sq((c(-3, 0) - c(-1, 0) - c(1, 0) + c(3, 0)) - 3.0f * (c(-2, 0) + c(2, 0)) + 6.0f * c(0, 0));
        }
        bv(tx, ty) = v;
        bh(tx, ty) = hz;
    }
    // lpf (demosaic.rs:188-201) over halo 7, R/B sites
    {
        constexpr int HW = Tile<TX, TY, H_LPF>::W / 2;
        for (int i = tid; i < HW * Tile<TX, TY, H_LPF>::R; i += NT) {
            int tx, ty;
            site(i, H_LPF, HW, false, tx, ty);
            float v = 0.0f;
            if (in_box(ox + tx, oy + ty, 2)) {
                const auto c = c_around(tx, ty, std::integral_constant<int, 1>{});
                v = c(0, 0) + 0.5f * (c(0, -1) + c(0, 1) + c(-1, 0) + c(1, 0)) + 0.25f * (c(-1, -1) + c(1, -1) + c(-1, 1) + c(1, 1));
            }
            lpf(tx, ty) = v;
        }
    }
    __syncthreads();

    // ---- stage 2: vh (demosaic.rs:173-184) over halo 6, every site; 0.5 off the border box
    for (int i = tid; i < Tile<TX, TY, H_VH>::N; i += NT) {
        int tx = i % Tile<TX, TY, H_VH>::W - H_VH, ty = i / Tile<TX, TY, H_VH>::W - H_VH;
        int gx = ox + tx, gy = oy + ty;
        float v = 0.5f;
        if (in_box(gx, gy, kBorder)) {
            const auto sv = bv.sten(tx, ty - 1);
            const auto sh = bh.sten(tx - 1, ty);
            float vs = fmaxr(sv(0, 0) + sv(0, 1) + sv(0, 2), kEps);
            float hs = fmaxr(sh(0, 0) + sh(1, 0) + sh(2, 0), kEps);
            v = vs / (vs + hs);
        }
        vh(tx, ty) = v;
    }
    __syncthreads();

    // ---- stage 3: green (demosaic.rs:222-245).  G sites copy c; R/B sites
    // in the border box get the RCD estimate, outside it the MHC value.
    constexpr int GW = Tile<TX, TY, H_G>::W, GH = Tile<TX, TY, H_G>::R;
    for (int i = tid; i < (GW / 2) * GH; i += NT) {
        int tx, ty, tx_g, ty_g;
        site(i, H_G, GW / 2, false, tx, ty);
        site(i, H_G, GW / 2, true, tx_g, ty_g);
        const int gy = oy + ty;
        // green site
        gg(tx_g, ty) = in_img(ox + tx_g, gy) ? cval(tx_g, ty) : 0.0f;
        // R/B site
        const int gx = ox + tx;
        float gv = 0.0f;
        if (in_img(gx, gy)) {
            if (in_box(gx, gy, kBorder)) {
                float vd = disc(vh.sten(tx - 1, ty - 1));
                const auto c = c_around(tx, ty, std::integral_constant<int, 4>{});
                float n_grad = kEps + fabsf(c(0, -1) - c(0, 1)) + fabsf(c(0, 0) - c(0, -2)) + fabsf(c(0, -1) - c(0, -3)) + fabsf(c(0, -2) - c(0, -4));
                float s_grad = kEps + fabsf(c(0, -1) - c(0, 1)) + fabsf(c(0, 0) - c(0, 2)) + fabsf(c(0, 1) - c(0, 3)) + fabsf(c(0, 2) - c(0, 4));
                float w_grad = kEps + fabsf(c(-1, 0) - c(1, 0)) + fabsf(c(0, 0) - c(-2, 0)) + fabsf(c(-1, 0) - c(-3, 0)) + fabsf(c(-2, 0) - c(-4, 0));
                float e_grad = kEps + fabsf(c(-1, 0) - c(1, 0)) + fabsf(c(0, 0) - c(2, 0)) + fabsf(c(1, 0) - c(3, 0)) + fabsf(c(2, 0) - c(4, 0));
                const auto lp = lpf.sten(tx - 2, ty - 2);
                float l0 = lp(2, 2);
                float n_est = c(0, -1) * 2.0f * l0 / (kEps + l0 + lp(2, 0));
                float s_est = c(0, 1) * 2.0f * l0 / (kEps + l0 + lp(2, 4));
                float w_est = c(-1, 0) * 2.0f * l0 / (kEps + l0 + lp(0, 2));
                float e_est = c(1, 0) * 2.0f * l0 / (kEps + l0 + lp(4, 2));
                float v_est = (s_grad * n_est + n_grad * s_est) / (n_grad + s_grad);
                float h_est = (w_grad * e_est + e_grad * w_est) / (e_grad + w_grad);
                gv = fmaxr(vd * h_est + (1.0f - vd) * v_est, 0.0f);
            } else {
                gv = mhc_at(cm, tx, ty, pat.at(gx, gy), pat.at(gx + 1, gy), 1);
            }
        }
        gg(tx, ty) = gv;
    }
    // disc(vh) at the tile's G sites, for stage 6
    for (int i = tid; i < TX * TY / 2; i += NT) {
        int tx, ty;
        site(i, 0, TX / 2, true, tx, ty);
        vdt[ty * TX + tx] = disc(vh.sten(tx - 1, ty - 1));
    }
    __syncthreads();
    // bp / bq (demosaic.rs:248-260) over halo 5, R/B sites, over lpf / vh
    // (free since the barrier above)
    for (int i = tid; i < (GW / 2) * GH; i += NT) {
        int tx, ty;
        site(i, H_G, GW / 2, false, tx, ty);
        float p = 0.0f, q = 0.0f;
        if (in_box(ox + tx, oy + ty, 3)) {
            const auto c = c_around(tx, ty, std::integral_constant<int, 3>{});
            p = sq((c(-3, -3) - c(-1, -1) - c(1, 1) + c(3, 3)) - 3.0f * (c(-2, -2) + c(2, 2)) + 6.0f * c(0, 0));
            q = sq((c(3, -3) - c(1, -1) - c(-1, 1) + c(-3, 3)) - 3.0f * (c(2, -2) + c(-2, 2)) + 6.0f * c(0, 0));
        }
        bp(tx, ty) = p;
        bq(tx, ty) = q;
    }
    __syncthreads();

    // ---- stage 4: pq (demosaic.rs:261-272) over halo 4, R/B sites
    {
        constexpr int HW = Tile<TX, TY, H_PQ>::W / 2;
        for (int i = tid; i < HW * Tile<TX, TY, H_PQ>::R; i += NT) {
            int tx, ty;
            site(i, H_PQ, HW, false, tx, ty);
            float v = 0.5f;
            if (in_box(ox + tx, oy + ty, kBorder)) {
                const auto sp = bp.sten(tx - 1, ty - 1);
                const auto sq_ = bq.sten(tx - 1, ty - 1);
                float ps = fmaxr(sp(0, 0) + sp(1, 1) + sp(2, 2), kEps);
                float qs = fmaxr(sq_(2, 0) + sq_(1, 1) + sq_(0, 2), kEps);
                v = ps / (ps + qs);
            }
            pq(tx, ty) = v;
        }
    }
    __syncthreads();

    // ---- stage 5: r and b at R/B sites after step 4.3 (demosaic.rs:276-308)
    // over halo 3.  R sites keep c for r; B sites in the border box get the
    // diagonal estimate, outside it the MHC value; symmetrically for b.  The
    // pre-pass channel at the diagonal neighbours of a B (R) site is an R (B)
    // site, i.e. c itself.
    constexpr int QW = Tile<TX, TY, H_RB>::W, QR = Tile<TX, TY, H_RB>::R;
    for (int i = tid; i < (QW / 2) * QR; i += NT) {
        int tx, ty;
        site(i, H_RB, QW / 2, false, tx, ty);
        const int gx = ox + tx, gy = oy + ty;
        float rv = 0.0f, bvv = 0.0f;
        if (in_img(gx, gy)) {
            const int col = pat.at(gx, gy);
            float own = cval(tx, ty);
            float other;
            if (in_box(gx, gy, kBorder)) {
                float pd = disc(pq.sten(tx - 1, ty - 1));
                const auto cc = c_around(tx, ty, std::integral_constant<int, 3>{});
                const auto gs = gg.sten(tx - 2, ty - 2);
                auto g = [&](int dx, int dy) { return gs(dx + 2, dy + 2); };
                float nw_g = kEps + fabsf(cc(-1, -1) - cc(1, 1)) + fabsf(cc(-1, -1) - cc(-3, -3)) + fabsf(g(0, 0) - g(-2, -2));
                float ne_g = kEps + fabsf(cc(1, -1) - cc(-1, 1)) + fabsf(cc(1, -1) - cc(3, -3)) + fabsf(g(0, 0) - g(2, -2));
                float sw_g = kEps + fabsf(cc(-1, 1) - cc(1, -1)) + fabsf(cc(-1, 1) - cc(-3, 3)) + fabsf(g(0, 0) - g(-2, 2));
                float se_g = kEps + fabsf(cc(1, 1) - cc(-1, -1)) + fabsf(cc(1, 1) - cc(3, 3)) + fabsf(g(0, 0) - g(2, 2));
                float nw_e = cc(-1, -1) - g(-1, -1);
                float ne_e = cc(1, -1) - g(1, -1);
                float sw_e = cc(-1, 1) - g(-1, 1);
                float se_e = cc(1, 1) - g(1, 1);
                float p_est = (nw_g * se_e + se_g * nw_e) / (nw_g + se_g);
                float q_est = (ne_g * sw_e + sw_g * ne_e) / (ne_g + sw_g);
                other = fmaxr(g(0, 0) + (1.0f - pd) * p_est + pd * q_est, 0.0f);
            } else {
                other = mhc_at(cm, tx, ty, col, pat.at(gx + 1, gy), 2 - col);
            }
            if (col == 0) {
                rv = own;
                bvv = other;
            } else {
                rv = other;
                bvv = own;
            }
        }
        r43(tx, ty) = rv;
        b43(tx, ty) = bvv;
    }
    __syncthreads();

    // ---- stage 6: outputs
    auto store = [&](int tx, int ty, float r, float g, float b) {
        const int gx = ox + tx, gy = oy + ty;
        size_t gi = (size_t)gy * (size_t)w + (size_t)gx;
        rgb[3 * gi + 0] = r;
        rgb[3 * gi + 1] = g;
        rgb[3 * gi + 2] = b;
    };
    // R and B at G sites (demosaic.rs:310-341)
    for (int i = tid; i < (TX / 2) * TY; i += NT) {
        int tx, ty;
        site(i, 0, TX / 2, true, tx, ty);
        const int gx = ox + tx, gy = oy + ty;
        if (gx >= w || gy >= h) continue;
        float o[3];
        o[1] = cval(tx, ty);
        if (in_box(gx, gy, kBorder)) {
            float vd = vdt[ty * TX + tx];
            const auto gs = gg.sten(tx - 2, ty - 2);
            auto g = [&](int dx, int dy) { return gs(dx + 2, dy + 2); };
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const auto cs = (k == 0 ? r43 : b43).sten(tx - 3, ty - 3);
                auto cc = [&](int dx, int dy) { return cs(dx + 3, dy + 3); };
                float n_g = kEps + fabsf(g(0, 0) - g(0, -2)) + fabsf(cc(0, -1) - cc(0, 1)) + fabsf(cc(0, -1) - cc(0, -3));
                float s_g = kEps + fabsf(g(0, 0) - g(0, 2)) + fabsf(cc(0, 1) - cc(0, -1)) + fabsf(cc(0, 1) - cc(0, 3));
                float w_g = kEps + fabsf(g(0, 0) - g(-2, 0)) + fabsf(cc(-1, 0) - cc(1, 0)) + fabsf(cc(-1, 0) - cc(-3, 0));
                float e_g = kEps + fabsf(g(0, 0) - g(2, 0)) + fabsf(cc(1, 0) - cc(-1, 0)) + fabsf(cc(1, 0) - cc(3, 0));
                float n_e = cc(0, -1) - g(0, -1);
                float s_e = cc(0, 1) - g(0, 1);
                float w_e = cc(-1, 0) - g(-1, 0);
                float e_e = cc(1, 0) - g(1, 0);
                float v_est = (n_g * s_e + s_g * n_e) / (n_g + s_g);
                float h_est = (e_g * w_e + w_g * e_e) / (e_g + w_g);
                o[k == 0 ? 0 : 2] = fmaxr(g(0, 0) + (1.0f - vd) * v_est + vd * h_est, 0.0f);
            }
        } else {
            int row_c = pat.at(gx + 1, gy);
            o[0] = mhc_at(cm, tx, ty, 1, row_c, 0);
            o[2] = mhc_at(cm, tx, ty, 1, row_c, 2);
        }
        store(tx, ty, o[0], o[1], o[2]);
    }
    // R/B sites: the values of stages 3 and 5
    for (int i = tid; i < (TX / 2) * TY; i += NT) {
        int tx, ty;
        site(i, 0, TX / 2, false, tx, ty);
        if (ox + tx >= w || oy + ty >= h) continue;
        store(tx, ty, r43(tx, ty), gg(tx, ty), b43(tx, ty));
    }
    if (clip) store_clip<TX, TY, NT>(L.flag, clip, ox, oy, w, h);
}

template <int TX, int TY, int NT>
__global__ __launch_bounds__(NT) void k_rcd(CfaSrc src, float* __restrict__ rgb, float* __restrict__ clip) {
    static_assert((TX % 2) == 0 && (TY % 2) == 0, "tile origin must keep Bayer parity");
    __shared__ RcdLds<TX, TY> L;
    const int ox = blockIdx.x * TX, oy = blockIdx.y * TY;
    const bool interior = ox - H_CM >= kBorder && oy - H_CM >= kBorder && ox + TX + H_CM <= src.w - kBorder &&
                          oy + TY + H_CM <= src.h - kBorder;
    if (interior)
        rcd_body<TX, TY, NT, true>(L, src, rgb, clip);
    else
        rcd_body<TX, TY, NT, false>(L, src, rgb, clip);
}

// Plain MHC over the whole frame (demosaic.rs:89-124) plus the clip plane:
// a TX x TY tile with a 2-pixel clamped halo of cfa*mul in LDS.
template <int TX, int TY, int NT>
__global__ __launch_bounds__(NT) void k_mhc(CfaSrc src, float* __restrict__ rgb, float* __restrict__ clip) {
    constexpr int H = 2;
    __shared__ float s_cm[Tile<TX, TY, H>::N];
    __shared__ __attribute__((aligned(4))) uint8_t s_flag[ClipFlags<TX, TY>::N];
    Tile<TX, TY, H> cm{s_cm};
    const int tid = threadIdx.x;
    const int ox = blockIdx.x * TX, oy = blockIdx.y * TY;
    const int w = src.w, h = src.h;
    const Bayer pat = src.pat;
    {
        constexpr int CW = Tile<TX, TY, H>::W;
        stage_tile<Tile<TX, TY, H>::N, NT>(
            [&](int i) {
                int gx = clampi(ox + i % CW - H, 0, w - 1), gy = clampi(oy + i / CW - H, 0, h - 1);
                return src.cfa(gx, gy, pat.at(gx, gy));
            },
            [&](int i, float v) {
                int tx = i % CW - H, ty = i / CW - H;
                int gx = clampi(ox + tx, 0, w - 1), gy = clampi(oy + ty, 0, h - 1);
                cm(tx, ty) = v * src.mul_of(pat.at(gx, gy));
                if (clip && tx >= -1 && tx <= TX && ty >= -1 && ty <= TY) s_flag[ClipFlags<TX, TY>::at(tx, ty)] = v >= kClip ? 1 : 0;
            });
    }
    __syncthreads();
    for (int i = tid; i < TX * TY; i += NT) {
        int tx = i % TX, ty = i / TX;
        int gx = ox + tx, gy = oy + ty;
        if (gx >= w || gy >= h) continue;
        int col = pat.at(gx, gy);
        int row_c = pat.at(gx + 1, gy);
        size_t gi = (size_t)gy * (size_t)w + (size_t)gx;
        rgb[3 * gi + 0] = mhc_at(cm, tx, ty, col, row_c, 0);
        rgb[3 * gi + 1] = mhc_at(cm, tx, ty, col, row_c, 1);
        rgb[3 * gi + 2] = mhc_at(cm, tx, ty, col, row_c, 2);
    }
    if (clip) store_clip<TX, TY, NT>(s_flag, clip, ox, oy, w, h);
}

template <int TX, int TY, int NT>
hipError_t launch_rcd(const CfaSrc& src, float* rgb, float* clip, hipStream_t stream) {
    dim3 grid((unsigned)((src.w + TX - 1) / TX), (unsigned)((src.h + TY - 1) / TY));
    hipLaunchKernelGGL((k_rcd<TX, TY, NT>), grid, dim3(NT), 0, stream, src, rgb, clip);
    return hipGetLastError();
}

template <int TX, int TY, int NT>
hipError_t launch_mhc(const CfaSrc& src, float* rgb, float* clip, hipStream_t stream) {
    dim3 grid((unsigned)((src.w + TX - 1) / TX), (unsigned)((src.h + TY - 1) / TY));
    hipLaunchKernelGGL((k_mhc<TX, TY, NT>), grid, dim3(NT), 0, stream, src, rgb, clip);
    return hipGetLastError();
}

CfaSrc make_src(const DemosaicArgs& a) {
    CfaSrc src{};
    src.f32 = a.cfa_f32;
    src.u16 = a.raw_u16;
    src.stride = a.stride;
    for (int i = 0; i < 4; i++) src.blk[i] = a.per_site ? a.black_site[i] : a.black;
    for (int c = 0; c < 3; c++) {
        src.white[c] = a.white[c];
        src.mul[c] = a.mul[c];
    }
    for (int i = 0; i < 4; i++) src.pat.c[i] = a.pattern[i];
    src.w = (int)a.w;
    src.h = (int)a.h;
    return src;
}

// cfa_clip_levels, first stage (demosaic.rs:419-427): block b takes sampled
// rows 3*(b + k*gridDim.x) and leaves the per-colour max of the normalised
// CFA in partial[3*b + c].  Starting from 0, fmaxf skips NaN exactly like
// f32::max; max is order-free, so the two stages equal the serial loop.
__global__ __launch_bounds__(256) void k_clip_partial(CfaSrc src, float* __restrict__ partial) {
    float mx[3] = {0.0f, 0.0f, 0.0f};
    const int rows = (src.h + 2) / 3;
    for (int r = blockIdx.x; r < rows; r += gridDim.x) {
        const int y = 3 * r;
        for (int x = threadIdx.x; x < src.w; x += 256) {
            const int c = src.pat.at(x, y);
            const float v = src.cfa(x, y, c);
            if (c == 0) mx[0] = fmaxf(mx[0], v);
            else if (c == 1) mx[1] = fmaxf(mx[1], v);
            else mx[2] = fmaxf(mx[2], v);
        }
    }
    __shared__ float red[3][4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int c = 0; c < 3; c++) {
        float v = mx[c];
        for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
        if (lane == 0) red[c][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        const int c = threadIdx.x;
        partial[3 * blockIdx.x + c] = fmaxf(fmaxf(red[c][0], red[c][1]), fmaxf(red[c][2], red[c][3]));
    }
}

// Second stage: a max is trusted as a clip level only at >= 0.95 of the
// nominal white, else f32::MAX (demosaic.rs:428-429).
__global__ __launch_bounds__(256) void k_clip_final(const float* __restrict__ partial, int parts, float* __restrict__ out3) {
    __shared__ float red[3][256];
    for (int c = 0; c < 3; c++) {
        float v = 0.0f;
        for (int i = threadIdx.x; i < parts; i += 256) v = fmaxf(v, partial[3 * i + c]);
        red[c][threadIdx.x] = v;
    }
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int c = 0; c < 3; c++) red[c][threadIdx.x] = fmaxf(red[c][threadIdx.x], red[c][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 3) {
        const float m = red[threadIdx.x][0];
        out3[threadIdx.x] = m >= 0.95f ? m : 3.40282347e+38f;
    }
}

// superpixel (demosaic.rs:44-87), Bayer: output pixel (x, y) from the 2x2
// tile at (2x, 2y), samples taken in the reference's order (dy outer).
__global__ __launch_bounds__(256) void k_superpixel(CfaSrc src, float* __restrict__ rgb, float* __restrict__ clip) {
    const int ow = src.w / 2, oh = src.h / 2;
    const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= ow || y >= oh) return;
    float acc[3] = {0.0f, 0.0f, 0.0f}, mx[3] = {0.0f, 0.0f, 0.0f};
    int n[3] = {0, 0, 0}, clipped = 0;
#pragma unroll
    for (int dy = 0; dy < 2; dy++)
#pragma unroll
        for (int dx = 0; dx < 2; dx++) {
            const int sx = 2 * x + dx, sy = 2 * y + dy;
            const int c = src.pat.at(sx, sy);
            const float v = src.cfa(sx, sy, c);
            clipped += v >= kClip ? 1 : 0;
#pragma unroll
            for (int k = 0; k < 3; k++)
                if (k == c) {
                    mx[k] = fmaxr(mx[k], v);
                    acc[k] += v;
                    n[k]++;
                }
        }
    const size_t o = (size_t)y * ow + x;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const float v = mx[k] >= kClip ? mx[k] : (n[k] > 0 ? acc[k] / (float)n[k] : 0.0f);
        rgb[3 * o + k] = v * src.mul[k];
    }
    if (clip) clip[o] = fminf((float)clipped / 4.0f * 2.0f, 1.0f);
}

// normalised CFA (decode.rs:300-303)
__global__ __launch_bounds__(256) void k_cfa_normalize(CfaSrc src, float* __restrict__ out) {
    const int x = blockIdx.x * 256 + threadIdx.x, y = blockIdx.y;
    if (x >= src.w) return;
    out[(size_t)y * src.w + x] = src.cfa(x, y, src.pat.at(x, y));
}

}  // namespace

hipError_t launch_superpixel(const DemosaicArgs& a, hipStream_t stream) {
    const CfaSrc src = make_src(a);
    const int ow = (int)(a.w / 2), oh = (int)(a.h / 2);
    if (ow == 0 || oh == 0) return hipSuccess;
    hipLaunchKernelGGL(k_superpixel, dim3((unsigned)((ow + 63) / 64), (unsigned)((oh + 3) / 4)), dim3(256), 0, stream, src, a.rgb, a.clip);
    return hipGetLastError();
}

hipError_t launch_cfa_normalize(const DemosaicArgs& a, float* out, hipStream_t stream) {
    const CfaSrc src = make_src(a);
    if (a.w == 0 || a.h == 0) return hipSuccess;
    hipLaunchKernelGGL(k_cfa_normalize, dim3((unsigned)((a.w + 255) / 256), (unsigned)a.h), dim3(256), 0, stream, src, out);
    return hipGetLastError();
}

hipError_t launch_cfa_clip_levels(const DemosaicArgs& a, float* partial, float* out3, hipStream_t stream) {
    const CfaSrc src = make_src(a);
    hipLaunchKernelGGL(k_clip_partial, dim3(CLIP_PARTS), dim3(256), 0, stream, src, partial);
    hipLaunchKernelGGL(k_clip_final, dim3(1), dim3(256), 0, stream, partial, CLIP_PARTS, out3);
    return hipGetLastError();
}

// Tile-shape variants (TX x TY, threads), all with the same per-pixel
// arithmetic, so every variant gives the same result bit for bit.
//   RCD: 0 = 64x32/1024, 1 = 32x32/256, 2 = 64x16/256
//   MHC: 0 = 128x32/1024, 1 = 128x8/256, 2 = 32x32/256
int demosaic_variants(int algo) { return algo == DS_ALGO_RCD ? 3 : 3; }

hipError_t launch_demosaic(const DemosaicArgs& a, hipStream_t stream, int variant) {
    const CfaSrc src = make_src(a);
    // rcd_bayer falls back to MHC below 16 pixels (demosaic.rs:149-151)
    bool rcd = a.algo == DS_ALGO_RCD && a.w >= 16 && a.h >= 16;
    if (rcd) {
        switch (variant) {
            case 1: return launch_rcd<32, 32, 256>(src, a.rgb, a.clip, stream);
            case 2: return launch_rcd<64, 16, 256>(src, a.rgb, a.clip, stream);
            default: return launch_rcd<64, 32, 1024>(src, a.rgb, a.clip, stream);
        }
    }
    switch (a.algo == DS_ALGO_MHC ? variant : 0) {
        case 1: return launch_mhc<128, 8, 256>(src, a.rgb, a.clip, stream);
        case 2: return launch_mhc<32, 32, 256>(src, a.rgb, a.clip, stream);
        default: return launch_mhc<128, 32, 1024>(src, a.rgb, a.clip, stream);
    }
}

}  // namespace ds
