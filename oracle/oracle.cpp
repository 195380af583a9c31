// CPU oracle for the raw detail path: demosaic -> NLM denoise -> RL sharpening.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
// calls this file.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
// baseline leg use it, as the checker and as the timed CPU restatement.
//
// Every function restates the reference Rust (Syha-01/darkslide,
// crates/core/src) operation by operation.  The FP32 instantiation is meant to
// be bit-faithful to the Rust: build with -ffp-contract=off and without
// fast-math, keep the left-to-right evaluation order of each expression,
// and use glibc expf/cbrtf (what Rust's f32::exp / f32::cbrt call on
// x86_64-linux-gnu).  The FP64 instantiation (T = double) is the same
// algorithm in double precision: it is the "FP64 oracle" that per-stage error
// figures are reported against.
//
// Parallelism mirrors rayon's decomposition: row loops run under OpenMP, and
// NLM splits its search offsets into ceil(n / threads) chunks whose partial
// sums are added in chunk order (enhance.rs:43-70).  Only NLM's result depends
// on the thread count.
//
// C ABI: or_<function>_f32 / or_<function>_f64.  All images are row-major,
// tightly packed; RGB images are interleaved [r,g,b] triples (image.rs:7-13).
// A Bayer pattern is 4 bytes, colours (0=R,1=G,2=B) at (x%2, y%2) as
// pat[(y%2)*2 + x%2] (decode.rs:16-28).

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <vector>

#include <omp.h>

namespace {

using std::size_t;
typedef std::ptrdiff_t isz;

template <class T> inline T t_exp(T v);
template <> inline float t_exp<float>(float v) { return expf(v); }
template <> inline double t_exp<double>(double v) { return exp(v); }
template <class T> inline T t_cbrt(T v);
template <> inline float t_cbrt<float>(float v) { return cbrtf(v); }
template <> inline double t_cbrt<double>(double v) { return cbrt(v); }
template <class T> inline T t_sqrt(T v);
template <> inline float t_sqrt<float>(float v) { return sqrtf(v); }
template <> inline double t_sqrt<double>(double v) { return sqrt(v); }

// f32::max / f64::max: returns the non-NaN operand.
template <class T> inline T vmax(T a, T b) { return std::fmax(a, b); }
template <class T> inline T vmin(T a, T b) { return std::fmin(a, b); }
template <class T> inline T vclamp(T v, T lo, T hi) {
    // Rust's f32::clamp: min / max checks in that order, NaN passes through.
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    return v;
}
inline isz iclamp(isz v, isz lo, isz hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Decimal constant as the reference spells it: parsed straight to f32 for
// the FP32 oracle (as rustc does for an f32 literal), to f64 for the FP64 one.
#define F(x) ((T)(sizeof(T) == 4 ? (double)(x##f) : (double)(x)))

inline int bayer_color(const uint8_t* pat, size_t x, size_t y) { return pat[(y % 2) * 2 + (x % 2)]; }

// ---------------------------------------------------------------------------
// Planes and blurs: image.rs

// image.rs:211-215 Plane::get_clamped
template <class T> inline T plane_at(const T* p, size_t w, size_t h, isz x, isz y) {
    return p[(size_t)iclamp(y, 0, (isz)h - 1) * w + (size_t)iclamp(x, 0, (isz)w - 1)];
}

// image.rs:268-313 box_blur: separable running-sum window mean with clamped
// edges.  The horizontal pass works row by row; the vertical pass keeps a
// running sum per column.  Both add (new - old) to the running sum, and scale
// by 1/(2r+1) on output.
template <class T> void box_blur(const T* p, size_t w, size_t h, size_t r, T* out) {
    if (r == 0 || w == 0 || h == 0) {
        std::memcpy(out, p, w * h * sizeof(T));
        return;
    }
    const T norm = (T)1 / (T)(2 * r + 1);
    const isz R = (isz)r;
    std::vector<T> tmp(w * h);
#pragma omp parallel for schedule(static)
    for (isz y = 0; y < (isz)h; y++) {
        const T* row = p + (size_t)y * w;
        T* o = tmp.data() + (size_t)y * w;
        auto at = [&](isz i) { return row[iclamp(i, 0, (isz)w - 1)]; };
        T acc = 0;
        for (isz i = -R; i <= R; i++) acc += at(i);
        for (isz x = 0; x < (isz)w; x++) {
            o[x] = acc * norm;
            acc += at(x + R + 1) - at(x - R);
        }
    }
    // vertical pass; the reference walks 64-column strips, which only
    // changes the schedule, not any sum.
#pragma omp parallel for schedule(static)
    for (isz x = 0; x < (isz)w; x++) {
        auto at = [&](isz y) { return tmp[(size_t)iclamp(y, 0, (isz)h - 1) * w + (size_t)x]; };
        T acc = 0;
        for (isz i = -R; i <= R; i++) acc += at(i);
        for (isz y = 0; y < (isz)h; y++) {
            out[(size_t)y * w + (size_t)x] = acc * norm;
            acc += at(y + R + 1) - at(y - R);
        }
    }
}

// image.rs:341-361 gaussian_blur_exact: r = ceil(3 sigma) taps each side,
// k_i = exp(-(i*i) / (2 sigma sigma)) normalised by their sequential sum,
// clamped edges, taps accumulated in order from -r to +r.
template <class T> void gaussian_taps(T sigma, std::vector<T>& k, isz& r) {
    r = (isz)std::ceil(sigma * (T)3);
    k.clear();
    for (isz i = -r; i <= r; i++) k.push_back(t_exp<T>(-(T)(i * i) / ((T)2 * sigma * sigma)));
    T s = 0;
    for (T v : k) s += v;
    for (T& v : k) v /= s;
}

template <class T> void gaussian_blur_exact(const T* p, size_t w, size_t h, T sigma, T* out) {
    std::vector<T> k;
    isz r;
    gaussian_taps(sigma, k, r);
    std::vector<T> tmp(w * h);
#pragma omp parallel for schedule(static)
    for (isz y = 0; y < (isz)h; y++)
        for (isz x = 0; x < (isz)w; x++) {
            T acc = 0;
            for (size_t i = 0; i < k.size(); i++) acc += k[i] * plane_at(p, w, h, x + (isz)i - r, y);
            tmp[(size_t)y * w + (size_t)x] = acc;
        }
#pragma omp parallel for schedule(static)
    for (isz y = 0; y < (isz)h; y++)
        for (isz x = 0; x < (isz)w; x++) {
            T acc = 0;
            for (size_t i = 0; i < k.size(); i++) acc += k[i] * plane_at(tmp.data(), w, h, x, y + (isz)i - r);
            out[(size_t)y * w + (size_t)x] = acc;
        }
}

// image.rs:363-373 box_radii_for_gauss (Wells' three-box approximation).
// The f32 arithmetic here is always single precision, as in the reference.
static void box_radii_for_gauss(float sigma, int n, size_t* radii) {
    float w_ideal = sqrtf(12.0f * sigma * sigma / (float)n + 1.0f);
    int wl = (int)floorf(w_ideal);
    if (wl % 2 == 0) wl -= 1;
    int wu = wl + 2;
    float m_ideal = (12.0f * sigma * sigma - (float)(n * wl * wl) - 4.0f * (float)n * (float)wl - 3.0f * (float)n) / (-4.0f * (float)wl - 4.0f);
    int m = (int)roundf(m_ideal);
    for (int i = 0; i < n; i++) {
        int v = ((i < m ? wl : wu) - 1) / 2;
        radii[i] = (size_t)(v > 0 ? v : 0);
    }
}

// image.rs:326-338 gaussian_blur: copy below 0.3, exact below 2, else 3 boxes.
template <class T> void gaussian_blur(const T* p, size_t w, size_t h, T sigma, T* out) {
    if (sigma < F(0.3)) {
        std::memcpy(out, p, w * h * sizeof(T));
        return;
    }
    if (sigma < (T)2.0) {
        gaussian_blur_exact(p, w, h, sigma, out);
        return;
    }
    size_t radii[3];
    box_radii_for_gauss((float)sigma, 3, radii);
    std::vector<T> a(w * h), b(w * h);
    box_blur(p, w, h, radii[0], a.data());
    box_blur(a.data(), w, h, radii[1], b.data());
    box_blur(b.data(), w, h, radii[2], out);
}

// image.rs:384-405 guided_filter (He et al.).
template <class T> void guided_filter(const T* I, const T* p, size_t w, size_t h, size_t r, T eps, T* out) {
    size_t n = w * h;
    std::vector<T> mean_i(n), mean_p(n), ip(n), ii(n), mean_ip(n), mean_ii(n), a(n), b(n), ma(n), mb(n);
    box_blur(I, w, h, r, mean_i.data());
    box_blur(p, w, h, r, mean_p.data());
#pragma omp parallel for schedule(static)
    for (isz k = 0; k < (isz)n; k++) {
        ip[k] = I[k] * p[k];
        ii[k] = I[k] * I[k];
    }
    box_blur(ip.data(), w, h, r, mean_ip.data());
    box_blur(ii.data(), w, h, r, mean_ii.data());
#pragma omp parallel for schedule(static)
    for (isz k = 0; k < (isz)n; k++) {
        T cov = mean_ip[k] - mean_i[k] * mean_p[k];
        T var = mean_ii[k] - mean_i[k] * mean_i[k];
        a[k] = cov / (var + eps);
        b[k] = mean_p[k] - a[k] * mean_i[k];
    }
    box_blur(a.data(), w, h, r, ma.data());
    box_blur(b.data(), w, h, r, mb.data());
#pragma omp parallel for schedule(static)
    for (isz k = 0; k < (isz)n; k++) out[k] = ma[k] * I[k] + mb[k];
}

// ---------------------------------------------------------------------------
// Colour: color.rs

typedef float M3f[3][3];

// color.rs:10-13 mul_mv (f32, left to right)
template <class T> inline void mul_mv(const T m[3][3], const T v[3], T o[3]) {
    for (int i = 0; i < 3; i++) o[i] = m[i][0] * v[0] + m[i][1] * v[1] + m[i][2] * v[2];
}

// color.rs:15-23 mul_mm: each entry is an f64 sum of f64 products, cast back.
template <class T> void mul_mm(const T a[3][3], const T b[3][3], T o[3][3]) {
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.0;
            for (int k = 0; k < 3; k++) s += (double)a[i][k] * (double)b[k][j];
            o[i][j] = (T)s;
        }
}

// color.rs:25-39 invert: f64 cofactor inverse, identity when singular.
template <class T> void invert(const T mi[3][3], T o[3][3]) {
    double m[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) m[i][j] = (double)mi[i][j];
    double det = m[0][0] * (m[1][1] * m[2][2] - m[2][1] * m[1][2]) - m[0][1] * (m[1][0] * m[2][2] - m[1][2] * m[2][0]) +
                 m[0][2] * (m[1][0] * m[2][1] - m[1][1] * m[2][0]);
    if (std::fabs(det) < 1e-12) {
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) o[i][j] = (T)(i == j ? 1 : 0);
        return;
    }
    double inv = 1.0 / det;
    double r[3][3] = {
        {(m[1][1] * m[2][2] - m[2][1] * m[1][2]) * inv, (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * inv, (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * inv},
        {(m[1][2] * m[2][0] - m[1][0] * m[2][2]) * inv, (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * inv, (m[1][0] * m[0][2] - m[0][0] * m[1][2]) * inv},
        {(m[1][0] * m[2][1] - m[2][0] * m[1][1]) * inv, (m[2][0] * m[0][1] - m[0][0] * m[2][1]) * inv, (m[0][0] * m[1][1] - m[1][0] * m[0][1]) * inv},
    };
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) o[i][j] = (T)r[i][j];
}

// color.rs:65-69 xy_to_xyz
template <class T> void xy_to_xyz(const T xy[2], T o[3]) {
    T x = xy[0];
    T y = vmax(xy[1], F(1e-6));
    o[0] = x / y;
    o[1] = (T)1;
    o[2] = ((T)1 - x - y) / y;
}

// Constants, color.rs:62-63, :79, :129, :383, :387.  In the FP32 oracle they
// are the f32 roundings of the decimal literals, as in the Rust; in the FP64
// oracle they are the decimal literals themselves.
template <class T> struct Consts {
    T d50[2] = {F(0.34567), F(0.35850)};
    T d65[2] = {F(0.31271), F(0.32902)};
    T bradford[3][3] = {{F(0.8951), F(0.2664), -F(0.1614)}, {-F(0.7502), F(1.7135), F(0.0367)}, {F(0.0389), -F(0.0685), F(1.0296)}};
    T pp_r[2] = {F(0.734699), F(0.265301)};
    T pp_g[2] = {F(0.159597), F(0.840403)};
    T pp_b[2] = {F(0.036598), F(0.000105)};
    T m1[3][3] = {{F(0.818933), F(0.36186674), -F(0.12885971)}, {F(0.032984544), F(0.9293119), F(0.03614564)}, {F(0.0482003), F(0.26436627), F(0.6338517)}};
    T m2[3][3] = {{F(0.21045426), F(0.7936178), -F(0.004072047)}, {F(1.9779985), -F(2.4285922), F(0.4505937)}, {F(0.025904037), F(0.78277177), -F(0.80867577)}};
};

// color.rs:82-87 bradford(src, dst)
template <class T> void bradford(const Consts<T>& K, const T src[2], const T dst[2], T o[3][3]) {
    T xs[3], xd[3], s[3], d[3];
    xy_to_xyz(src, xs);
    xy_to_xyz(dst, xd);
    mul_mv(K.bradford, xs, s);
    mul_mv(K.bradford, xd, d);
    T scale[3][3] = {{d[0] / s[0], 0, 0}, {0, d[1] / s[1], 0}, {0, 0, d[2] / s[2]}};
    T binv[3][3], sb[3][3];
    invert(K.bradford, binv);
    mul_mm(scale, K.bradford, sb);
    mul_mm(binv, sb, o);
}

// color.rs:90-98 rgb_to_xyz_d50 for ProPhoto (white = D50, so no adaptation).
template <class T> void prophoto_to_xyz(const Consts<T>& K, T o[3][3]) {
    T pr[3], pg[3], pb[3], wx[3], s[3];
    xy_to_xyz(K.pp_r, pr);
    xy_to_xyz(K.pp_g, pg);
    xy_to_xyz(K.pp_b, pb);
    T m[3][3] = {{pr[0], pg[0], pb[0]}, {pr[1], pg[1], pb[1]}, {pr[2], pg[2], pb[2]}};
    T minv[3][3];
    invert(m, minv);
    xy_to_xyz(K.d50, wx);
    mul_mv(minv, wx, s);
    T dg[3][3] = {{s[0], 0, 0}, {0, s[1], 0}, {0, 0, s[2]}};
    mul_mm(m, dg, o);
}

// color.rs:382-418 Oklab from / to linear ProPhoto.
template <class T> struct Oklab {
    T to_lms[3][3], from_lms[3][3], m2[3][3], m2_inv[3][3];
    Oklab() {
        Consts<T> K;
        T brad[3][3], xyz[3][3], bx[3][3];
        bradford(K, K.d50, K.d65, brad);
        prophoto_to_xyz(K, xyz);
        mul_mm(brad, xyz, bx);
        mul_mm(K.m1, bx, to_lms);
        invert(to_lms, from_lms);
        std::memcpy(m2, K.m2, sizeof(m2));
        invert(K.m2, m2_inv);
    }
    void from_prophoto(const T p[3], T o[3]) const {
        T lms[3];
        mul_mv(to_lms, p, lms);
        T c[3] = {t_cbrt(lms[0]), t_cbrt(lms[1]), t_cbrt(lms[2])};
        mul_mv(m2, c, o);
    }
    void to_prophoto(const T lab[3], T o[3]) const {
        T l[3];
        mul_mv(m2_inv, lab, l);
        T cube[3] = {l[0] * l[0] * l[0], l[1] * l[1] * l[1], l[2] * l[2] * l[2]};
        mul_mv(from_lms, cube, o);
    }
};

// color.rs:194-198 luma_prophoto
template <class T> inline T luma_prophoto(const T* p) { return F(0.2880402) * p[0] + F(0.7118741) * p[1] + F(0.0000857) * p[2]; }

// ---------------------------------------------------------------------------
// Demosaic: demosaic.rs

const float CLIP = 0.995f;  // demosaic.rs:17

// demosaic.rs:29-42 clip_plane: 1 where the CFA value reaches CLIP, then a
// clamped 3x3 maximum.
template <class T> void clip_plane(const T* cfa, size_t w, size_t h, T* out) {
    std::vector<T> raw(w * h);
    for (size_t i = 0; i < w * h; i++) raw[i] = cfa[i] >= (T)CLIP ? (T)1 : (T)0;
#pragma omp parallel for schedule(static)
    for (isz y = 0; y < (isz)h; y++)
        for (isz x = 0; x < (isz)w; x++) {
            T m = 0;
            for (isz dy = -1; dy <= 1; dy++)
                for (isz dx = -1; dx <= 1; dx++) m = vmax(m, plane_at(raw.data(), w, h, x + dx, y + dy));
            out[(size_t)y * w + (size_t)x] = m;
        }
}

// demosaic.rs:89-124 mhc_bayer: Malvar-He-Cutler on cfa*mul (not clamped),
// clamp-to-edge addressing, outputs floored at 0.
template <class T> void mhc_bayer(const T* cfa, size_t w, size_t h, const uint8_t* pat, const T* mul, T* rgb) {
    auto at = [&](isz x, isz y) -> T {
        size_t cx = (size_t)iclamp(x, 0, (isz)w - 1), cy = (size_t)iclamp(y, 0, (isz)h - 1);
        return cfa[cy * w + cx] * mul[bayer_color(pat, cx, cy)];
    };
#pragma omp parallel for schedule(static)
    for (isz yi = 0; yi < (isz)h; yi++)
        for (isz xi = 0; xi < (isz)w; xi++) {
            int c = bayer_color(pat, (size_t)xi, (size_t)yi);
            T v = at(xi, yi);
            auto p = [&](isz dx, isz dy) { return at(xi + dx, yi + dy); };
            T g_at_rb = ((T)4 * v + (T)2 * (p(0, -1) + p(-1, 0) + p(1, 0) + p(0, 1)) - (p(0, -2) + p(-2, 0) + p(2, 0) + p(0, 2))) / (T)8;
            T rb_g_row = ((T)5 * v + (T)4 * (p(-1, 0) + p(1, 0)) - (p(-2, 0) + p(2, 0) + p(-1, -1) + p(1, -1) + p(-1, 1) + p(1, 1)) +
                          (T)0.5 * (p(0, -2) + p(0, 2))) /
                         (T)8;
            T rb_g_col = ((T)5 * v + (T)4 * (p(0, -1) + p(0, 1)) - (p(0, -2) + p(0, 2) + p(-1, -1) + p(1, -1) + p(-1, 1) + p(1, 1)) +
                          (T)0.5 * (p(-2, 0) + p(2, 0))) /
                         (T)8;
            T rb_diag = ((T)6 * v + (T)2 * (p(-1, -1) + p(1, -1) + p(-1, 1) + p(1, 1)) - (T)1.5 * (p(0, -2) + p(-2, 0) + p(2, 0) + p(0, 2))) / (T)8;
            T o[3];
            if (c == 1) {
                int row_c = bayer_color(pat, (size_t)xi + 1, (size_t)yi);
                if (row_c == 0) {
                    o[0] = rb_g_row; o[1] = v; o[2] = rb_g_col;
                } else {
                    o[0] = rb_g_col; o[1] = v; o[2] = rb_g_row;
                }
            } else if (c == 0) {
                o[0] = v; o[1] = g_at_rb; o[2] = rb_diag;
            } else {
                o[0] = rb_diag; o[1] = g_at_rb; o[2] = v;
            }
            T* dst = rgb + 3 * ((size_t)yi * w + (size_t)xi);
            for (int k = 0; k < 3; k++) dst[k] = vmax(o[k], (T)0);
        }
}

// demosaic.rs:145-344 rcd_bayer.  Stage order and the border of each stage
// follow the reference exactly; frame borders keep the MHC values.
template <class T> void rcd_bayer(const T* cfa, size_t w, size_t h, const uint8_t* pat, const T* mul, T* rgb) {
    if (w < 16 || h < 16) {
        mhc_bayer(cfa, w, h, pat, mul, rgb);
        return;
    }
    const size_t n = w * h;
    auto fc = [&](size_t x, size_t y) { return bayer_color(pat, x, y); };
    std::vector<T> c(n);
    for (size_t i = 0; i < n; i++) c[i] = vmax(cfa[i] * mul[fc(i % w, i / w)], (T)0);
    const T eps = F(1e-5);
    auto sq = [](T v) { return v * v; };
    const size_t w1 = w, w2 = 2 * w, w3 = 3 * w, w4 = 4 * w;
    const size_t border = 4;

    // step 1 (demosaic.rs:160-186): vertical / horizontal discrimination
    std::vector<T> bv(n, 0), bh(n, 0);
#pragma omp parallel for schedule(static)
    for (isz yy = 3; yy < (isz)h - 3; yy++)
        for (size_t x = 3; x < w - 3; x++) {
            size_t i = (size_t)yy * w + x;
            bv[i] = sq((c[i - w3] - c[i - w1] - c[i + w1] + c[i + w3]) - (T)3 * (c[i - w2] + c[i + w2]) + (T)6 * c[i]);
            bh[i] = sq((c[i - 3] - c[i - 1] - c[i + 1] + c[i + 3]) - (T)3 * (c[i - 2] + c[i + 2]) + (T)6 * c[i]);
        }
    std::vector<T> vh(n, (T)0.5);
#pragma omp parallel for schedule(static)
    for (isz yy = (isz)border; yy < (isz)(h - border); yy++)
        for (size_t x = border; x < w - border; x++) {
            size_t i = (size_t)yy * w + x;
            T vs = vmax(bv[i - w1] + bv[i] + bv[i + w1], eps);
            T hs = vmax(bh[i - 1] + bh[i] + bh[i + 1], eps);
            vh[i] = vs / (vs + hs);
        }

    // step 2 (demosaic.rs:188-201): low pass at R/B sites
    std::vector<T> lpf(n, 0);
#pragma omp parallel for schedule(static)
    for (isz yy = 2; yy < (isz)h - 2; yy++)
        for (size_t x = 2; x < w - 2; x++) {
            if (fc(x, (size_t)yy) == 1) continue;
            size_t i = (size_t)yy * w + x;
            lpf[i] = c[i] + (T)0.5 * (c[i - w1] + c[i + w1] + c[i - 1] + c[i + 1]) +
                     (T)0.25 * (c[i - w1 - 1] + c[i - w1 + 1] + c[i + w1 - 1] + c[i + w1 + 1]);
        }

    // demosaic.rs:203-214: MHC base, channel split
    std::vector<T> base(3 * n);
    mhc_bayer(cfa, w, h, pat, mul, base.data());
    std::vector<T> r(n), g(n), b(n);
    for (size_t i = 0; i < n; i++) {
        int col = fc(i % w, i / w);
        r[i] = col == 0 ? c[i] : base[3 * i + 0];
        g[i] = col == 1 ? c[i] : base[3 * i + 1];
        b[i] = col == 2 ? c[i] : base[3 * i + 2];
    }

    // demosaic.rs:216-220 disc
    auto disc = [&](const std::vector<T>& d, size_t i) -> T {
        T central = d[i];
        T neigh = (T)0.25 * (d[i - w1 - 1] + d[i - w1 + 1] + d[i + w1 - 1] + d[i + w1 + 1]);
        return std::fabs((T)0.5 - central) < std::fabs((T)0.5 - neigh) ? neigh : central;
    };

    // step 3 (demosaic.rs:222-245): green at R/B
#pragma omp parallel for schedule(static)
    for (isz yy = (isz)border; yy < (isz)(h - border); yy++)
        for (size_t x = border; x < w - border; x++) {
            if (fc(x, (size_t)yy) == 1) continue;
            size_t i = (size_t)yy * w + x;
            T vd = disc(vh, i);
            T n_grad = eps + std::fabs(c[i - w1] - c[i + w1]) + std::fabs(c[i] - c[i - w2]) + std::fabs(c[i - w1] - c[i - w3]) + std::fabs(c[i - w2] - c[i - w4]);
            T s_grad = eps + std::fabs(c[i - w1] - c[i + w1]) + std::fabs(c[i] - c[i + w2]) + std::fabs(c[i + w1] - c[i + w3]) + std::fabs(c[i + w2] - c[i + w4]);
            T w_grad = eps + std::fabs(c[i - 1] - c[i + 1]) + std::fabs(c[i] - c[i - 2]) + std::fabs(c[i - 1] - c[i - 3]) + std::fabs(c[i - 2] - c[i - 4]);
            T e_grad = eps + std::fabs(c[i - 1] - c[i + 1]) + std::fabs(c[i] - c[i + 2]) + std::fabs(c[i + 1] - c[i + 3]) + std::fabs(c[i + 2] - c[i + 4]);
            T n_est = c[i - w1] * (T)2 * lpf[i] / (eps + lpf[i] + lpf[i - w2]);
            T s_est = c[i + w1] * (T)2 * lpf[i] / (eps + lpf[i] + lpf[i + w2]);
            T w_est = c[i - 1] * (T)2 * lpf[i] / (eps + lpf[i] + lpf[i - 2]);
            T e_est = c[i + 1] * (T)2 * lpf[i] / (eps + lpf[i] + lpf[i + 2]);
            T v_est = (s_grad * n_est + n_grad * s_est) / (n_grad + s_grad);
            T h_est = (w_grad * e_est + e_grad * w_est) / (e_grad + w_grad);
            g[i] = vmax(vd * h_est + ((T)1 - vd) * v_est, (T)0);
        }

    // step 4.1/4.2 (demosaic.rs:248-274): diagonal discrimination
    std::vector<T> bp(n, 0), bq(n, 0);
#pragma omp parallel for schedule(static)
    for (isz yy = 3; yy < (isz)h - 3; yy++)
        for (size_t x = 3; x < w - 3; x++) {
            size_t i = (size_t)yy * w + x;
            bp[i] = sq((c[i - w3 - 3] - c[i - w1 - 1] - c[i + w1 + 1] + c[i + w3 + 3]) - (T)3 * (c[i - w2 - 2] + c[i + w2 + 2]) + (T)6 * c[i]);
            bq[i] = sq((c[i - w3 + 3] - c[i - w1 + 1] - c[i + w1 - 1] + c[i + w3 - 3]) - (T)3 * (c[i - w2 + 2] + c[i + w2 - 2]) + (T)6 * c[i]);
        }
    std::vector<T> pq(n, (T)0.5);
#pragma omp parallel for schedule(static)
    for (isz yy = (isz)border; yy < (isz)(h - border); yy++)
        for (size_t x = border; x < w - border; x++) {
            size_t i = (size_t)yy * w + x;
            T ps = vmax(bp[i - w1 - 1] + bp[i] + bp[i + w1 + 1], eps);
            T qs = vmax(bq[i - w1 + 1] + bq[i] + bq[i + w1 - 1], eps);
            pq[i] = ps / (ps + qs);
        }

    // step 4.3 (demosaic.rs:276-308): R at B, then B at R, from the diagonal
    // colour differences; chan is a copy taken before the pass.
    auto pass43 = [&](std::vector<T>& target, int at_color) {
        std::vector<T> chan = target;
#pragma omp parallel for schedule(static)
        for (isz yy = (isz)border; yy < (isz)(h - border); yy++)
            for (size_t x = border; x < w - border; x++) {
                if (fc(x, (size_t)yy) != at_color) continue;
                size_t i = (size_t)yy * w + x;
                T pd = disc(pq, i);
                const std::vector<T>& cc = chan;
                T nw_g = eps + std::fabs(cc[i - w1 - 1] - cc[i + w1 + 1]) + std::fabs(cc[i - w1 - 1] - cc[i - w3 - 3]) + std::fabs(g[i] - g[i - w2 - 2]);
                T ne_g = eps + std::fabs(cc[i - w1 + 1] - cc[i + w1 - 1]) + std::fabs(cc[i - w1 + 1] - cc[i - w3 + 3]) + std::fabs(g[i] - g[i - w2 + 2]);
                T sw_g = eps + std::fabs(cc[i + w1 - 1] - cc[i - w1 + 1]) + std::fabs(cc[i + w1 - 1] - cc[i + w3 - 3]) + std::fabs(g[i] - g[i + w2 - 2]);
                T se_g = eps + std::fabs(cc[i + w1 + 1] - cc[i - w1 - 1]) + std::fabs(cc[i + w1 + 1] - cc[i + w3 + 3]) + std::fabs(g[i] - g[i + w2 + 2]);
                T nw_e = cc[i - w1 - 1] - g[i - w1 - 1];
                T ne_e = cc[i - w1 + 1] - g[i - w1 + 1];
                T sw_e = cc[i + w1 - 1] - g[i + w1 - 1];
                T se_e = cc[i + w1 + 1] - g[i + w1 + 1];
                T p_est = (nw_g * se_e + se_g * nw_e) / (nw_g + se_g);
                T q_est = (ne_g * sw_e + sw_g * ne_e) / (ne_g + sw_g);
                target[i] = vmax(g[i] + ((T)1 - pd) * p_est + pd * q_est, (T)0);
            }
    };
    pass43(r, 2);
    pass43(b, 0);

    // step 4.4 (demosaic.rs:310-341): R and B at G
    auto pass44 = [&](std::vector<T>& target) {
        std::vector<T> chan = target;
#pragma omp parallel for schedule(static)
        for (isz yy = (isz)border; yy < (isz)(h - border); yy++)
            for (size_t x = border; x < w - border; x++) {
                if (fc(x, (size_t)yy) != 1) continue;
                size_t i = (size_t)yy * w + x;
                T vd = disc(vh, i);
                const std::vector<T>& cc = chan;
                T n_g = eps + std::fabs(g[i] - g[i - w2]) + std::fabs(cc[i - w1] - cc[i + w1]) + std::fabs(cc[i - w1] - cc[i - w3]);
                T s_g = eps + std::fabs(g[i] - g[i + w2]) + std::fabs(cc[i + w1] - cc[i - w1]) + std::fabs(cc[i + w1] - cc[i + w3]);
                T w_g = eps + std::fabs(g[i] - g[i - 2]) + std::fabs(cc[i - 1] - cc[i + 1]) + std::fabs(cc[i - 1] - cc[i - 3]);
                T e_g = eps + std::fabs(g[i] - g[i + 2]) + std::fabs(cc[i + 1] - cc[i - 1]) + std::fabs(cc[i + 1] - cc[i + 3]);
                T n_e = cc[i - w1] - g[i - w1];
                T s_e = cc[i + w1] - g[i + w1];
                T w_e = cc[i - 1] - g[i - 1];
                T e_e = cc[i + 1] - g[i + 1];
                T v_est = (n_g * s_e + s_g * n_e) / (n_g + s_g);
                T h_est = (e_g * w_e + w_g * e_e) / (e_g + w_g);
                target[i] = vmax(g[i] + ((T)1 - vd) * v_est + vd * h_est, (T)0);
            }
    };
    pass44(r);
    pass44(b);

    for (size_t i = 0; i < n; i++) {
        rgb[3 * i + 0] = r[i];
        rgb[3 * i + 1] = g[i];
        rgb[3 * i + 2] = b[i];
    }
}

// decode.rs:291-306: v = (raw - black) / max(white[c] - black, 1), no clamp.
template <class T> void cfa_normalize_u16(const uint16_t* raw, size_t w, size_t h, size_t stride, const uint8_t* pat, T black, const T* white, T* out) {
#pragma omp parallel for schedule(static)
    for (isz y = 0; y < (isz)h; y++)
        for (size_t x = 0; x < w; x++) {
            int c = bayer_color(pat, x, (size_t)y);
            T v = (T)raw[(size_t)y * stride + x];
            out[(size_t)y * w + x] = (v - black) / vmax(white[c] - black, (T)1);
        }
}

// load_raw's CFA branch (decode.rs:276-307, the non-float case) on a raw
// buffer of `width` samples per row: the crop rectangle, the colour of the
// buffer position (pattern given at the buffer origin), the black level
// black_at(sx, sy, 0) from a bw x bh table of bcpp values per entry (or one
// value), white_at(c) from one or >= 3 values, and the pattern of the
// cropped image (CfaPattern::shifted, decode.rs:38-46).
template <class T>
void cfa_ingest_u16(const uint16_t* raw, size_t width, size_t crop_x, size_t crop_y, size_t cw, size_t ch, const uint8_t* pat, const T* black,
                    size_t nblack, size_t bw, size_t bh, size_t bcpp, const T* white, size_t nwhite, T* out, uint8_t* pat_out) {
    bw = bw ? bw : 1;
    bh = bh ? bh : 1;
    bcpp = bcpp ? bcpp : 1;
    auto black_at = [&](size_t x, size_t y, size_t c) -> T {
        if (nblack == 1) return black[0];
        return black[((y % bh) * bw + (x % bw)) * bcpp + (c < bcpp - 1 ? c : bcpp - 1)];
    };
    auto white_at = [&](int c) -> T { return nwhite >= 3 ? white[(size_t)c < nwhite - 1 ? (size_t)c : nwhite - 1] : white[0]; };
#pragma omp parallel for schedule(static)
    for (isz yy = 0; yy < (isz)ch; yy++) {
        const size_t sy = (size_t)yy + crop_y;
        for (size_t x = 0; x < cw; x++) {
            const size_t sx = x + crop_x;
            const int c = bayer_color(pat, sx, sy);
            const T raw_v = (T)raw[sy * width + sx];
            const T b = black_at(sx, sy, 0);
            out[(size_t)yy * cw + x] = (raw_v - b) / vmax(white_at(c) - b, (T)1);
        }
    }
    for (size_t y = 0; y < 2; y++)
        for (size_t x = 0; x < 2; x++) pat_out[y * 2 + x] = (uint8_t)bayer_color(pat, x + crop_x, y + crop_y);
}

// superpixel (demosaic.rs:44-87) for a Bayer CFA: each 2x2 tile becomes one
// pixel, w/2 x h/2.  A colour with a sample >= 0.995 keeps its max, else the
// mean of its samples; the clip value is the clipped fraction times 2, at
// most 1.
template <class T> void superpixel(const T* cfa, size_t w, size_t h, const uint8_t* pat, const T* mul, T* rgb, T* clip) {
    const size_t ow = w / 2, oh = h / 2;
#pragma omp parallel for schedule(static)
    for (isz yy = 0; yy < (isz)oh; yy++) {
        const size_t y = (size_t)yy;
        for (size_t x = 0; x < ow; x++) {
            T acc[3] = {0, 0, 0}, mx[3] = {0, 0, 0};
            unsigned n[3] = {0, 0, 0}, clipped = 0;
            for (size_t dy = 0; dy < 2; dy++)
                for (size_t dx = 0; dx < 2; dx++) {
                    const size_t sx = x * 2 + dx, sy = y * 2 + dy;
                    const int c = bayer_color(pat, sx, sy);
                    const T v = cfa[sy * w + sx];
                    if (v >= F(0.995)) clipped++;
                    mx[c] = vmax(mx[c], v);
                    acc[c] += v;
                    n[c]++;
                }
            for (int c = 0; c < 3; c++) {
                const T v = mx[c] >= F(0.995) ? mx[c] : (n[c] > 0 ? acc[c] / (T)n[c] : (T)0);
                rgb[3 * (y * ow + x) + c] = v * mul[c];
            }
            clip[y * ow + x] = vmin((T)clipped / F(4.0) * F(2.0), (T)1);
        }
    }
}

// ---------------------------------------------------------------------------
// Enhance: enhance.rs

// enhance.rs:9-32 estimate_noise: every step-th pixel off the 1-px border,
// |4c - l - r - u - d|, upper median, / 0.6745 / sqrt(20).
template <class T> T estimate_noise(const T* p, size_t w, size_t h) {
    size_t step = std::max<size_t>((w * h) / 200000, 1);
    std::vector<T> v;
    for (size_t i = 0; i < w * h; i += step) {
        size_t x = i % w, y = i / w;
        if (x < 1 || y < 1 || x + 1 >= w || y + 1 >= h) continue;
        T c = p[i];
        T lap = (T)4 * c - p[i - 1] - p[i + 1] - p[i - w] - p[i + w];
        v.push_back(std::fabs(lap));
    }
    if (v.empty()) return (T)0;
    size_t mid = v.size() / 2;
    std::nth_element(v.begin(), v.begin() + (isz)mid, v.end());
    return v[mid] / F(0.6745) / t_sqrt((T)20);
}

// enhance.rs:34-72 nlm.  Offsets run dy-major over [-S,S]^2; each chunk of
// ceil(n / threads) offsets keeps its own (acc, wsum) planes, which are then
// added into zero-initialised totals in chunk order.
template <class T> void nlm(const T* p, size_t w, size_t h, T h_param, isz search, size_t patch, int threads, T* out) {
    const size_t n = w * h;
    const T h2 = vmax(h_param * h_param, F(1e-12));
    std::vector<std::pair<isz, isz>> offsets;
    for (isz dy = -search; dy <= search; dy++)
        for (isz dx = -search; dx <= search; dx++) offsets.push_back({dx, dy});
    if (threads < 1) threads = 1;
    size_t chunk = std::max<size_t>((offsets.size() + (size_t)threads - 1) / (size_t)threads, 1);
    size_t nchunks = (offsets.size() + chunk - 1) / chunk;
    std::vector<std::vector<T>> pa(nchunks), ps(nchunks);
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (isz ci = 0; ci < (isz)nchunks; ci++) {
        std::vector<T>& a = pa[(size_t)ci];
        std::vector<T>& s = ps[(size_t)ci];
        a.assign(n, (T)0);
        s.assign(n, (T)0);
        std::vector<T> diff(n), dist(n);
        size_t o_end = std::min(offsets.size(), ((size_t)ci + 1) * chunk);
        for (size_t o = (size_t)ci * chunk; o < o_end; o++) {
            isz dx = offsets[o].first, dy = offsets[o].second;
            for (isz y = 0; y < (isz)h; y++)
                for (isz x = 0; x < (isz)w; x++) {
                    T q = plane_at(p, w, h, x + dx, y + dy);
                    T d = p[(size_t)y * w + (size_t)x] - q;
                    diff[(size_t)y * w + (size_t)x] = d * d;
                }
            // nested parallelism is off inside a chunk (rayon steals, which
            // does not change any sum)
            if (patch == 0) {
                std::memcpy(dist.data(), diff.data(), n * sizeof(T));
            } else {
                const T norm = (T)1 / (T)(2 * patch + 1);
                const isz R = (isz)patch;
                std::vector<T> tmp(n);
                for (isz y = 0; y < (isz)h; y++) {
                    const T* row = diff.data() + (size_t)y * w;
                    auto at = [&](isz i) { return row[iclamp(i, 0, (isz)w - 1)]; };
                    T acc = 0;
                    for (isz i = -R; i <= R; i++) acc += at(i);
                    for (isz x = 0; x < (isz)w; x++) {
                        tmp[(size_t)y * w + (size_t)x] = acc * norm;
                        acc += at(x + R + 1) - at(x - R);
                    }
                }
                for (isz x = 0; x < (isz)w; x++) {
                    auto at = [&](isz y) { return tmp[(size_t)iclamp(y, 0, (isz)h - 1) * w + (size_t)x]; };
                    T acc = 0;
                    for (isz i = -R; i <= R; i++) acc += at(i);
                    for (isz y = 0; y < (isz)h; y++) {
                        dist[(size_t)y * w + (size_t)x] = acc * norm;
                        acc += at(y + R + 1) - at(y - R);
                    }
                }
            }
            for (isz y = 0; y < (isz)h; y++)
                for (isz x = 0; x < (isz)w; x++) {
                    size_t i = (size_t)y * w + (size_t)x;
                    T wt = t_exp<T>(-(dist[i] / h2));
                    a[i] += wt * plane_at(p, w, h, x + dx, y + dy);
                    s[i] += wt;
                }
        }
    }
    std::vector<T> acc(n, (T)0), wsum(n, (T)0);
    for (size_t ci = 0; ci < nchunks; ci++)
        for (size_t i = 0; i < n; i++) {
            acc[i] += pa[ci][i];
            wsum[i] += ps[ci][i];
        }
    for (size_t i = 0; i < n; i++) out[i] = acc[i] / vmax(wsum[i], F(1e-12));
}

// enhance.rs:74-92 denoise, with the NLM search / patch radii as arguments
// (the reference passes 3, 1 at enhance.rs:84).
template <class T> void denoise(const T* img, size_t w, size_t h, T amount, isz search, size_t patch, int threads, T* out) {
    const size_t n = w * h;
    Oklab<T> ok;
    std::vector<T> L(n), A(n), B(n);
#pragma omp parallel for schedule(static)
    for (isz i = 0; i < (isz)n; i++) {
        T lab[3];
        ok.from_prophoto(img + 3 * (size_t)i, lab);
        L[(size_t)i] = lab[0];
        A[(size_t)i] = lab[1];
        B[(size_t)i] = lab[2];
    }
    T sigma = vmax(estimate_noise(L.data(), w, h), F(1e-4));
    T k = amount / (T)100;
    std::vector<T> lf(n), l2(n), a2(n), b2(n);
    nlm(L.data(), w, h, sigma * (F(0.6) + F(1.6) * k), search, patch, threads, lf.data());
    for (size_t i = 0; i < n; i++) l2[i] = L[i] + (lf[i] - L[i]) * ((T)0.5 + (T)0.5 * k);
    T rf = vmax(((T)std::max(w, h) / (T)600) * ((T)1 + (T)2 * k), (T)2);
    size_t r = (size_t)rf;
    T e2 = sigma * (T)2;
    T eps = e2 * e2;
    guided_filter(l2.data(), A.data(), w, h, r, eps, a2.data());
    guided_filter(l2.data(), B.data(), w, h, r, eps, b2.data());
#pragma omp parallel for schedule(static)
    for (isz i = 0; i < (isz)n; i++) {
        T lab[3] = {l2[(size_t)i], a2[(size_t)i], b2[(size_t)i]};
        T o[3];
        ok.to_prophoto(lab, o);
        for (int c = 0; c < 3; c++) out[3 * (size_t)i + c] = vmax(o[c], (T)0);
    }
}

// enhance.rs:94-119 raw_details: Richardson-Lucy on luminance with a
// gaussian PSF, applied as a clamped luminance ratio.
template <class T> void raw_details(const T* img, size_t w, size_t h, T sigma, size_t iters, T* out) {
    const size_t n = w * h;
    std::vector<T> y(n), est(n), conv(n), ratio(n), corr(n);
    for (size_t i = 0; i < n; i++) y[i] = vmax(luma_prophoto(img + 3 * i), F(1e-6));
    est = y;
    for (size_t it = 0; it < iters; it++) {
        gaussian_blur(est.data(), w, h, sigma, conv.data());
        for (size_t i = 0; i < n; i++) ratio[i] = y[i] / vmax(conv[i], F(1e-6));
        gaussian_blur(ratio.data(), w, h, sigma, corr.data());
        for (size_t i = 0; i < n; i++) est[i] = vclamp(est[i] * corr[i], (T)0, (T)64);
    }
    for (size_t i = 0; i < n; i++) {
        T r = vclamp(est[i] / y[i], F(0.6), F(1.6));
        for (int c = 0; c < 3; c++) out[3 * i + c] = img[3 * i + c] * r;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// ---- highlight reconstruction (demosaic.rs:418-502)

template <class T> inline T f32max() { return (T)3.40282346638528859812e+38; }  // f32::MAX in either type

// cfa_clip_levels (demosaic.rs:419-430): per-colour max over every third row,
// trusted as a clip level only when >= 0.95.
template <class T> void cfa_clip_levels(const T* cfa, size_t w, size_t h, const uint8_t* pat, T* out3) {
    T mx[3] = {0, 0, 0};
    for (size_t y = 0; y < h; y += 3)
        for (size_t x = 0; x < w; x++) {
            const int c = pat[((y & 1) << 1) | (x & 1)];
            mx[c] = vmax(mx[c], cfa[y * w + x]);
        }
    for (int c = 0; c < 3; c++) out3[c] = mx[c] >= F(0.95) ? mx[c] : f32max<T>();
}

// reconstruct_highlights (demosaic.rs:435-502), in place on interleaved RGB.
template <class T> void reconstruct_highlights(T* img, const T* clip, size_t w, size_t h, const T* mul, const T* raw_clip) {
    const size_t n = w * h;
    bool any = false;
    for (size_t i = 0; i < n && !any; i++) any = clip[i] > (T)0;
    if (!any) return;
    T lvl[3];
    for (int c = 0; c < 3; c++) lvl[c] = raw_clip[c] == f32max<T>() ? f32max<T>() : raw_clip[c] * mul[c] * F(0.97);
    auto is_clipped = [&](const T* p, int c, size_t i) { return p[c] >= lvl[c] || (clip[i] > (T)0 && p[c] >= lvl[c] * F(0.8)); };
    const T sigma = vmax((T)(w > h ? w : h) / F(120.0), F(2.0));
    const T lmin = vmin(vmin(vmin(lvl[0], lvl[1]), lvl[2]), F(4.0));
    std::vector<T> near_clip(n), wgt(n, (T)0), rr(n, (T)0), br(n, (T)0);
    box_blur<T>(clip, w, h, 3, near_clip.data());
#pragma omp parallel for schedule(static)
    for (isz ii = 0; ii < (isz)n; ii++) {
        const size_t i = (size_t)ii;
        const T* p = img + 3 * i;
        if (near_clip[i] > (T)0 || is_clipped(p, 0, i) || is_clipped(p, 1, i) || is_clipped(p, 2, i) || p[1] <= F(1e-4)) continue;
        const T y = (p[0] + p[1] + p[2]) / F(3.0) / lmin;
        const T k = vmin(y * y, F(1.0));
        wgt[i] = k;
        rr[i] = vclamp(p[0] / p[1], F(0.1), F(10.0)) * k;
        br[i] = vclamp(p[2] / p[1], F(0.1), F(10.0)) * k;
    }
    std::vector<T> wb(n), rrb(n), brb(n);
    gaussian_blur<T>(wgt.data(), w, h, sigma, wb.data());
    gaussian_blur<T>(rr.data(), w, h, sigma, rrb.data());
    gaussian_blur<T>(br.data(), w, h, sigma, brb.data());
#pragma omp parallel for schedule(static)
    for (isz ii = 0; ii < (isz)n; ii++) {
        const size_t i = (size_t)ii;
        T* p = img + 3 * i;
        const bool cl[3] = {is_clipped(p, 0, i), is_clipped(p, 1, i), is_clipped(p, 2, i)};
        const int nc = (int)cl[0] + (int)cl[1] + (int)cl[2];
        if (nc == 0) continue;
        const T m = vmax(vmax(p[0], p[1]), p[2]);
        if (nc == 3) {
            p[0] = p[1] = p[2] = m;
            continue;
        }
        const T wv = wb[i];
        T r = F(1.0), b = F(1.0);
        if (wv > F(1e-4)) {
            r = rrb[i] / wv;
            b = brb[i] / wv;
        }
        const T trust = vmin(wv * F(3.0), F(1.0));
        r = F(1.0) + (r - F(1.0)) * trust;
        b = F(1.0) + (b - F(1.0)) * trust;
        const T ratio[3] = {r, F(1.0), b};
        T g = p[1];
        for (int c = 0; c < 3; c++)
            if (!cl[c]) g = vmax(g, p[c] / ratio[c]);
        for (int c = 0; c < 3; c++)
            if (cl[c]) p[c] = vmax(p[c], g * ratio[c]);
        const T top = vmax(vmax(p[0], p[1]), p[2]);
        const T q = (T)nc / F(3.0);
        const T t = q * q;
        for (int c = 0; c < 3; c++) p[c] += (top - p[c]) * t * F(0.5);
    }
}

// ---------------------------------------------------------------------------
// Detail: pipeline/detail.rs (noise_reduction, sharpen) and the helpers they
// use: Plane::resize / sample_bilinear (image.rs:217-262), fast_guided_filter
// (image.rs:408-431), srgb_encode / srgb_decode (color.rs:185-192), enc / dec
// and smoothstep (pipeline/tone.rs:16-19, :328-335).

template <class T> inline T t_pow(T a, T b);
template <> inline float t_pow<float>(float a, float b) { return powf(a, b); }
template <> inline double t_pow<double>(double a, double b) { return pow(a, b); }
template <class T> inline T t_round(T v);
template <> inline float t_round<float>(float v) { return roundf(v); }
template <> inline double t_round<double>(double v) { return round(v); }

template <class T> inline T srgb_encode(T v) { return v <= F(0.0031308) ? F(12.92) * vmax(v, (T)0) : F(1.055) * t_pow(v, F(1.0) / F(2.4)) - F(0.055); }
template <class T> inline T srgb_decode(T v) { return v <= F(0.04045) ? v / F(12.92) : t_pow((v + F(0.055)) / F(1.055), F(2.4)); }
template <class T> inline T tone_enc(T v) { return v <= (T)1 ? srgb_encode(v) : (T)1 + (v - (T)1) / F(2.4); }
template <class T> inline T tone_dec(T e) { return e <= (T)1 ? srgb_decode(e) : (T)1 + (e - (T)1) * F(2.4); }
template <class T> inline T smoothstep(T a, T b, T x) {
    const T t = vclamp((x - a) / (b - a), (T)0, (T)1);
    return t * t * (F(3.0) - F(2.0) * t);
}

// Plane::resize (image.rs:238-262): box average when shrinking by more than
// 1.5x on an axis, else bilinear at pixel centres.
template <class T> void plane_resize(const T* p, size_t w, size_t h, size_t w2, size_t h2, T* out) {
    if (w2 == w && h2 == h) {
        std::memcpy(out, p, w * h * sizeof(T));
        return;
    }
    // Plane::resize computes the scale in f32 whatever the element type
    const float sx = (float)w / (float)w2, sy = (float)h / (float)h2;
    auto at = [&](isz x, isz y) { return p[(size_t)iclamp(y, 0, (isz)h - 1) * w + (size_t)iclamp(x, 0, (isz)w - 1)]; };
    if (sx > 1.5f || sy > 1.5f) {
#pragma omp parallel for schedule(static)
        for (isz yy = 0; yy < (isz)h2; yy++) {
            const size_t y = (size_t)yy;
            for (size_t x = 0; x < w2; x++) {
                const size_t x0 = (size_t)((float)x * sx);
                size_t x1 = (size_t)((float)(x + 1) * sx);
                x1 = x1 < x0 + 1 ? x0 + 1 : (x1 > w ? w : x1);
                const size_t y0 = (size_t)((float)y * sy);
                size_t y1 = (size_t)((float)(y + 1) * sy);
                y1 = y1 < y0 + 1 ? y0 + 1 : (y1 > h ? h : y1);
                T acc = 0;
                for (size_t v = y0; v < y1; v++)
                    for (size_t u = x0; u < x1; u++) acc += p[v * w + u];
                out[y * w2 + x] = acc / (T)((x1 - x0) * (y1 - y0));
            }
        }
    } else {
#pragma omp parallel for schedule(static)
        for (isz yy = 0; yy < (isz)h2; yy++) {
            const size_t y = (size_t)yy;
            for (size_t x = 0; x < w2; x++) {
                const float fx = ((float)x + 0.5f) * sx - 0.5f, fy = ((float)y + 0.5f) * sy - 0.5f;
                const float x0 = std::floor(fx), y0 = std::floor(fy);
                const T ax = (T)(fx - x0), ay = (T)(fy - y0);
                const isz ix = (isz)x0, iy = (isz)y0;
                const T a = at(ix, iy) * ((T)1 - ax) + at(ix + 1, iy) * ax;
                const T b = at(ix, iy + 1) * ((T)1 - ax) + at(ix + 1, iy + 1) * ax;
                out[y * w2 + x] = a + (b - a) * ay;
            }
        }
    }
}

// fast_guided_filter (image.rs:408-431): coefficients on a 1/sub grid.
template <class T> void fast_guided_filter(const T* I, const T* p, size_t w, size_t h, size_t r, T eps, size_t sub, T* out) {
    if (sub <= 1 || w / sub < 8 || h / sub < 8) {
        guided_filter(I, p, w, h, r, eps, out);
        return;
    }
    const size_t sw = w / sub, sh = h / sub, n = sw * sh, rs = r / sub > 1 ? r / sub : 1;
    std::vector<T> is(n), ps(n), ip(n), ii(n), mi(n), mp(n), mip(n), mii(n), a(n), b(n), ba(n), bb(n), ma(w * h), mb(w * h);
    plane_resize(I, w, h, sw, sh, is.data());
    plane_resize(p, w, h, sw, sh, ps.data());
    for (size_t k = 0; k < n; k++) {
        ip[k] = is[k] * ps[k];
        ii[k] = is[k] * is[k];
    }
    box_blur(is.data(), sw, sh, rs, mi.data());
    box_blur(ps.data(), sw, sh, rs, mp.data());
    box_blur(ip.data(), sw, sh, rs, mip.data());
    box_blur(ii.data(), sw, sh, rs, mii.data());
    for (size_t k = 0; k < n; k++) {
        const T cov = mip[k] - mi[k] * mp[k];
        const T var = mii[k] - mi[k] * mi[k];
        a[k] = cov / (var + eps);
        b[k] = mp[k] - a[k] * mi[k];
    }
    box_blur(a.data(), sw, sh, rs, ba.data());
    box_blur(b.data(), sw, sh, rs, bb.data());
    plane_resize(ba.data(), sw, sh, w, h, ma.data());
    plane_resize(bb.data(), sw, sh, w, h, mb.data());
    for (size_t k = 0; k < w * h; k++) out[k] = ma[k] * I[k] + mb[k];
}

// detail::noise_reduction (pipeline/detail.rs:13-92) in place on interleaved
// ProPhoto.  nr = {luminance, luminance_detail, luminance_contrast, color,
// color_detail, color_smoothness} (params.rs:319-326); field: the
// LocalFields noise plane or null.
template <class T> void noise_reduction(T* img, size_t w, size_t h, const T* nr, const T* field, T scale) {
    const bool local = field != nullptr;
    const T lum = nr[0] / F(100.0), col = nr[3] / F(100.0);
    if (lum <= (T)0 && col <= (T)0 && !local) return;
    if (scale < F(0.34)) return;
    const size_t n = w * h;
    const T eff = vclamp(scale, F(0.05), F(1.0));
    Oklab<T> ok;
    std::vector<T> l(n), a(n), b(n);
    for (size_t i = 0; i < n; i++) {
        T lab[3];
        ok.from_prophoto(img + 3 * i, lab);
        l[i] = lab[0];
        a[i] = lab[1];
        b[i] = lab[2];
    }
    bool changed = false;
    if (lum > (T)0 || local) {
        T fmax = 0;
        if (local)
            for (size_t i = 0; i < n; i++) fmax = vmax(fmax, field[i]);
        const T max_strength = vclamp(lum + (local ? fmax / F(100.0) : (T)0), (T)0, (T)1);
        if (max_strength > (T)0) {
            const T detail = nr[1] / F(100.0);
            const T sigma = F(0.004) + max_strength * F(0.045) * t_sqrt(eff);
            const T q = sigma * (F(1.4) - detail);
            const T eps = q * q;
            const T rr1 = t_round(F(1.5) * eff), rr2 = t_round(F(4.0) * eff);
            const size_t r1 = rr1 > (T)1 ? (size_t)rr1 : 1, r2 = rr2 > (T)1 ? (size_t)rr2 : 1;
            std::vector<T> f1(n), f2(n);
            guided_filter(l.data(), l.data(), w, h, r1, eps, f1.data());
            guided_filter(f1.data(), f1.data(), w, h, r2, eps * F(0.5), f2.data());
            const T contrast = nr[2] / F(100.0);
            for (size_t i = 0; i < n; i++) {
                T s = lum + (local ? field[i] : (T)0) / F(100.0);
                s = vclamp(s, (T)0, (T)1);
                const T filt = f2[i] + (l[i] - f2[i]) * contrast * F(0.35);
                l[i] = l[i] + (filt - l[i]) * vmin(s / vmax(max_strength, F(1e-6)), (T)1);
            }
            changed = true;
        }
    }
    if (col > (T)0) {
        const T detail = nr[4] / F(100.0), smooth = nr[5] / F(100.0);
        const T rf = vmax(t_round((F(2.0) + col * F(10.0)) * (F(1.0) + smooth * F(1.5)) * eff), F(1.0));
        const size_t r = (size_t)rf;
        const T q = F(0.002) + (F(1.0) - detail) * F(0.02);
        const T eps = q * q;
        const size_t sub = r >= 8 ? 4 : (r >= 3 ? 2 : 1);
        std::vector<T> fa(n), fb(n);
        fast_guided_filter(l.data(), a.data(), w, h, r, eps, sub, fa.data());
        fast_guided_filter(l.data(), b.data(), w, h, r, eps, sub, fb.data());
        const T k = vmin(col, F(1.0));
        for (size_t i = 0; i < n; i++) {
            a[i] = a[i] + (fa[i] - a[i]) * k;
            b[i] = b[i] + (fb[i] - b[i]) * k;
        }
        changed = true;
    }
    if (changed)
        for (size_t i = 0; i < n; i++) {
            const T lab[3] = {l[i], a[i], b[i]};
            T o[3];
            ok.to_prophoto(lab, o);
            for (int c = 0; c < 3; c++) img[3 * i + c] = vmax(o[c], (T)0);
        }
}

// detail::sharpen (pipeline/detail.rs:96-139) in place.  sh = {amount,
// radius, detail, masking} (params.rs:304-309); field: the LocalFields
// sharpness plane or null.
template <class T> void sharpen(T* img, size_t w, size_t h, const T* sh, const T* field, T scale) {
    const bool local = field != nullptr;
    if (sh[0] <= (T)0 && !local) return;
    const T sigma = sh[1] * vmin(scale, F(1.0));
    if (sigma < F(0.25)) return;
    const size_t n = w * h;
    std::vector<T> e(n), blur(n), s, m;
    for (size_t i = 0; i < n; i++) e[i] = tone_enc(vmax(luma_prophoto(img + 3 * i), (T)0));
    gaussian_blur(e.data(), w, h, sigma, blur.data());
    const T detail = sh[2] / F(100.0);
    const bool mask = sh[3] > (T)0;
    if (mask) {
        s.resize(n);
        m.resize(n);
        gaussian_blur(e.data(), w, h, F(1.0) * vmax(scale, F(0.5)), s.data());
        const T t = sh[3] / F(100.0) * F(0.06);
        auto at = [&](isz x, isz y) { return s[(size_t)iclamp(y, 0, (isz)h - 1) * w + (size_t)iclamp(x, 0, (isz)w - 1)]; };
        for (size_t y = 0; y < h; y++)
            for (size_t x = 0; x < w; x++) {
                const T gx = at((isz)x + 1, (isz)y) - at((isz)x - 1, (isz)y);
                const T gy = at((isz)x, (isz)y + 1) - at((isz)x, (isz)y - 1);
                const T g = t_sqrt(gx * gx + gy * gy);
                m[y * w + x] = smoothstep(t * F(0.5), t * F(1.5) + F(1e-4), g);
            }
    }
    for (size_t i = 0; i < n; i++) {
        const T amt = (sh[0] + (local ? field[i] : (T)0) * F(1.5)) / F(100.0) * F(1.6);
        if (amt == (T)0) continue;
        const T d = e[i] - blur[i];
        const T damp = detail + (F(1.0) - detail) * t_exp(-std::fabs(d) / F(0.04));
        T dv = d * damp;
        if (mask) dv *= m[i];
        const T e2 = vmax(e[i] + amt * dv, (T)0);
        const T y0 = vmax(tone_dec(e[i]), F(1e-6));
        const T y1 = tone_dec(e2);
        const T r = y1 / y0;
        for (int c = 0; c < 3; c++) img[3 * i + c] = img[3 * i + c] * r;
    }
}

// enhance.rs:122-131 super_res_refine in place: square-root luminance, its
// gaussian (sigma 1.2), a clamped detail boost, applied as a clamped ratio.
template <class T> void super_res_refine(T* img, size_t w, size_t h) {
    const size_t n = w * h;
    std::vector<T> y(n), b(n);
    for (size_t i = 0; i < n; i++) y[i] = t_sqrt(vmax(luma_prophoto(img + 3 * i), (T)0));
    gaussian_blur(y.data(), w, h, F(1.2), b.data());
    for (size_t i = 0; i < n; i++) {
        const T d = vclamp(y[i] - b[i], F(-0.05), F(0.05));
        const T nv = vmax(y[i] + d * F(0.9), (T)0);
        const T r = (nv * nv) / vmax(y[i] * y[i], F(1e-8));
        const T rc = vclamp(r, F(0.5), F(2.0));
        for (int c = 0; c < 3; c++) img[3 * i + c] = img[3 * i + c] * rc;
    }
}

// export.rs:199-223 output_sharpen in place.  mode: 0 None, 1 Screen,
// 2 Matte, 3 Glossy (SharpenFor); amt: 0 Low, 1 Standard, 2 High.
template <class T> void output_sharpen(T* img, size_t w, size_t h, int mode, int amt) {
    T radius, base;
    switch (mode) {
        case 1: radius = F(0.6), base = F(0.45); break;
        case 2: radius = F(1.0), base = F(0.8); break;
        case 3: radius = F(0.8), base = F(0.6); break;
        default: return;
    }
    const T k = base * (amt == 0 ? F(0.6) : amt == 2 ? F(1.5) : F(1.0));
    const size_t n = w * h;
    std::vector<T> l(n), b(n);
    for (size_t i = 0; i < n; i++) l[i] = tone_enc(vmax(luma_prophoto(img + 3 * i), (T)0));
    gaussian_blur(l.data(), w, h, radius, b.data());
    for (size_t i = 0; i < n; i++) {
        const T d = l[i] - b[i];
        const T e2 = vmax(l[i] + k * vclamp(d, F(-0.1), F(0.1)), (T)0);
        const T y0 = vmax(tone_dec(l[i]), F(1e-6));
        const T r = tone_dec(e2) / y0;
        for (int c = 0; c < 3; c++) img[3 * i + c] = img[3 * i + c] * r;
    }
}

// C ABI

#define OR_BOTH(X) X(float, f32) X(double, f64)

extern "C" {

#define DEF(T, S)                                                                                                                  \
    void or_cfa_normalize_u16_##S(const uint16_t* raw, size_t w, size_t h, size_t stride, const uint8_t* pat, T black,              \
                                  const T* white, T* out) {                                                                        \
        cfa_normalize_u16<T>(raw, w, h, stride, pat, black, white, out);                                                            \
    }                                                                                                                              \
    void or_cfa_ingest_u16_##S(const uint16_t* raw, size_t width, size_t cx, size_t cy, size_t cw, size_t ch, const uint8_t* pat,     \
                               const T* black, size_t nblack, size_t bw, size_t bh, size_t bcpp, const T* white, size_t nwhite, T* out,   \
                               uint8_t* pat_out) {                                                                                   \
        cfa_ingest_u16<T>(raw, width, cx, cy, cw, ch, pat, black, nblack, bw, bh, bcpp, white, nwhite, out, pat_out);                 \
    }                                                                                                                              \
    void or_superpixel_##S(const T* cfa, size_t w, size_t h, const uint8_t* pat, const T* mul, T* rgb, T* clip) {                  \
        superpixel<T>(cfa, w, h, pat, mul, rgb, clip);                                                                             \
    }                                                                                                                              \
    void or_plane_resize_##S(const T* p, size_t w, size_t h, size_t w2, size_t h2, T* out) { plane_resize<T>(p, w, h, w2, h2, out); } \
    void or_fast_guided_filter_##S(const T* I, const T* p, size_t w, size_t h, size_t r, T eps, size_t sub, T* out) {             \
        fast_guided_filter<T>(I, p, w, h, r, eps, sub, out);                                                                       \
    }                                                                                                                              \
    void or_noise_reduction_##S(T* img, size_t w, size_t h, const T* nr, const T* field, T scale) {                                \
        noise_reduction<T>(img, w, h, nr, field, scale);                                                                           \
    }                                                                                                                              \
    void or_sharpen_##S(T* img, size_t w, size_t h, const T* sh, const T* field, T scale) { sharpen<T>(img, w, h, sh, field, scale); } \
    void or_super_res_refine_##S(T* img, size_t w, size_t h) { super_res_refine<T>(img, w, h); }                                   \
    void or_output_sharpen_##S(T* img, size_t w, size_t h, int mode, int amt) { output_sharpen<T>(img, w, h, mode, amt); }         \
    T or_tone_enc_##S(T v) { return tone_enc<T>(v); }                                                                              \
    T or_tone_dec_##S(T v) { return tone_dec<T>(v); }                                                                              \
    void or_mhc_bayer_##S(const T* cfa, size_t w, size_t h, const uint8_t* pat, const T* mul, T* rgb) {                            \
        mhc_bayer<T>(cfa, w, h, pat, mul, rgb);                                                                                    \
    }                                                                                                                              \
    void or_rcd_bayer_##S(const T* cfa, size_t w, size_t h, const uint8_t* pat, const T* mul, T* rgb) {                            \
        rcd_bayer<T>(cfa, w, h, pat, mul, rgb);                                                                                    \
    }                                                                                                                              \
    void or_clip_plane_##S(const T* cfa, size_t w, size_t h, T* out) { clip_plane<T>(cfa, w, h, out); }                              \
    void or_oklab_from_prophoto_##S(const T* rgb, size_t n, T* lab) {                                                              \
        Oklab<T> ok;                                                                                                               \
        for (size_t i = 0; i < n; i++) ok.from_prophoto(rgb + 3 * i, lab + 3 * i);                                                 \
    }                                                                                                                              \
    void or_oklab_to_prophoto_##S(const T* lab, size_t n, T* rgb) {                                                                \
        Oklab<T> ok;                                                                                                               \
        for (size_t i = 0; i < n; i++) ok.to_prophoto(lab + 3 * i, rgb + 3 * i);                                                   \
    }                                                                                                                              \
    void or_oklab_matrices_##S(T* to_lms, T* from_lms, T* m2, T* m2_inv) {                                                         \
        Oklab<T> ok;                                                                                                               \
        std::memcpy(to_lms, ok.to_lms, 9 * sizeof(T));                                                                             \
        std::memcpy(from_lms, ok.from_lms, 9 * sizeof(T));                                                                         \
        std::memcpy(m2, ok.m2, 9 * sizeof(T));                                                                                     \
        std::memcpy(m2_inv, ok.m2_inv, 9 * sizeof(T));                                                                             \
    }                                                                                                                              \
    void or_luminance_##S(const T* rgb, size_t n, T* out) {                                                                        \
        for (size_t i = 0; i < n; i++) out[i] = luma_prophoto(rgb + 3 * i);                                                        \
    }                                                                                                                              \
    T or_estimate_noise_##S(const T* p, size_t w, size_t h) { return estimate_noise<T>(p, w, h); }                                  \
    void or_box_blur_##S(const T* p, size_t w, size_t h, size_t r, T* out) { box_blur<T>(p, w, h, r, out); }                         \
    void or_gaussian_blur_##S(const T* p, size_t w, size_t h, T sigma, T* out) { gaussian_blur<T>(p, w, h, sigma, out); }            \
    void or_guided_filter_##S(const T* I, const T* p, size_t w, size_t h, size_t r, T eps, T* out) {                               \
        guided_filter<T>(I, p, w, h, r, eps, out);                                                                                 \
    }                                                                                                                              \
    void or_nlm_##S(const T* p, size_t w, size_t h, T h_param, long search, size_t patch, int threads, T* out) {                   \
        nlm<T>(p, w, h, h_param, (isz)search, patch, threads, out);                                                                \
    }                                                                                                                              \
    void or_denoise_##S(const T* img, size_t w, size_t h, T amount, long search, size_t patch, int threads, T* out) {              \
        denoise<T>(img, w, h, amount, (isz)search, patch, threads, out);                                                           \
    }                                                                                                                              \
    void or_raw_details_##S(const T* img, size_t w, size_t h, T sigma, size_t iters, T* out) {                                     \
        raw_details<T>(img, w, h, sigma, iters, out);                                                                              \
    }                                                                                                                              \
    void or_cfa_clip_levels_##S(const T* cfa, size_t w, size_t h, const uint8_t* pat, T* out3) {                                    \
        cfa_clip_levels<T>(cfa, w, h, pat, out3);                                                                                  \
    }                                                                                                                              \
    void or_reconstruct_highlights_##S(T* img, const T* clip, size_t w, size_t h, const T* mul, const T* raw_clip) {              \
        reconstruct_highlights<T>(img, clip, w, h, mul, raw_clip);                                                                 \
    }                                                                                                                              \
    void or_demosaic_##S(const T* cfa, size_t w, size_t h, const uint8_t* pat, const T* mul, T* rgb, T* clip) {                    \
        rcd_bayer<T>(cfa, w, h, pat, mul, rgb);                                                                                    \
        if (clip) clip_plane<T>(cfa, w, h, clip);                                                                                  \
    }

OR_BOTH(DEF)
#undef DEF

void or_gaussian_taps_f32(float sigma, float* taps, int* radius) {
    std::vector<float> k;
    isz r;
    gaussian_taps(sigma, k, r);
    *radius = (int)r;
    for (size_t i = 0; i < k.size(); i++) taps[i] = k[i];
}

void or_box_radii_for_gauss(float sigma, size_t* radii) { box_radii_for_gauss(sigma, 3, radii); }

int or_max_threads(void) { return omp_get_max_threads(); }

// Engine::auto_wb_xy statistics (pipeline/mod.rs:421-441): sequential FP64
// accumulation over every third pixel of the image, as the reference loops.
// out: acc[3], n, neutral[3] (neutral only when n > 0; returns 1 then).
int or_auto_wb_stats(const float* rgb, size_t npx, const float* n_as, double* acc, double* n_out, float* neutral) {
    acc[0] = acc[1] = acc[2] = 0.0;
    double n = 0.0;
    for (size_t i = 0; i < npx; i += 3) {
        const float* p = rgb + 3 * i;
        const float raw[3] = {p[0] * n_as[0], p[1] * n_as[1], p[2] * n_as[2]};
        const float m = std::max(std::max(raw[0], raw[1]), raw[2]);
        const float lo = std::min(std::min(raw[0], raw[1]), raw[2]);
        if (m > 0.9f || lo < 0.002f) continue;
        const double w = std::sqrt((double)m);
        for (int c = 0; c < 3; c++) acc[c] += std::log((double)raw[c]) * w;
        n += w;
    }
    *n_out = n;
    if (!(n > 0.0)) return 0;
    const double g = std::exp(acc[1] / n);
    neutral[0] = (float)(std::exp(acc[0] / n) / g);
    neutral[1] = 1.0f;
    neutral[2] = (float)(std::exp(acc[2] / n) / g);
    return 1;
}

}  // extern "C"
