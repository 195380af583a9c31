// Seeded synthetic Bayer scenes: smooth gradients + random convex polygons
// (sharp edges, uniform colours) + sinusoidal texture, linear in
// [0.02, 0.85] of the sensor range, then Poisson-Gaussian noise in DN.
//
// The per-frame scene description is drawn on the host from a sequential
// splitmix64 stream; per-pixel noise uses a counter-based hash of
// (seed, pixel index, draw index), so any pixel can be generated
// independently on the host or on the GPU.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace ds {

constexpr int SYNTH_MAX_POLY = 48;
constexpr int SYNTH_POLY_K = 7;
constexpr int SYNTH_NSIN = 3;

struct SynthPoly {
    float nx[SYNTH_POLY_K], ny[SYNTH_POLY_K], d[SYNTH_POLY_K];
    float x0, y0, x1, y1;  // bounding box
    float col[3];
};

struct SynthSin {
    float kx, ky, phase;  // radians per pixel, phase
    float amp[3];
};

struct SynthScene {
    uint64_t seed;
    int w, h;
    float black, white;  // DN
    float gain;          // DN per electron (0.5 x ISO factor)
    float read_sigma;    // DN
    float grad[3][4];    // per channel: a0 + ax*u + ay*v + axy*u*v, (u, v) in [0,1]
    int npoly;
    SynthPoly poly[SYNTH_MAX_POLY];
    SynthSin sn[SYNTH_NSIN];
    uint8_t pat[4];
};

__host__ __device__ inline uint64_t splitmix64_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct SplitMix64 {
    uint64_t s;
    __host__ __device__ uint64_t next() {
        s += 0x9E3779B97F4A7C15ull;
        return splitmix64_mix(s);
    }
    __host__ __device__ double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    __host__ __device__ double range(double a, double b) { return a + (b - a) * uniform(); }
};

// Uniform in [0,1) for draw `k` of pixel `idx`.
__host__ __device__ inline double synth_u01(uint64_t seed, uint64_t idx, uint32_t k) {
    uint64_t z = seed * 0xD1B54A32D192ED03ull + idx * 0x9E3779B97F4A7C15ull + (uint64_t)k * 0xF1357AEA2E62A9C5ull;
    return (double)(splitmix64_mix(splitmix64_mix(z)) >> 11) * (1.0 / 9007199254740992.0);
}

// Host-side scene description.  bits: 12 or 14; iso: noise multiplier >= 1,
// or 0 for the noise-free frame of the same scene.
inline void synth_scene_init(SynthScene* s, uint64_t seed, int w, int h, int bits, float iso, const uint8_t pat[4]) {
    SplitMix64 r{seed ^ 0xA5A5A5A55A5A5A5Aull};
    s->seed = seed;
    s->w = w;
    s->h = h;
    if (bits == 12) {
        s->black = 128.0f;
        s->white = 4095.0f;
    } else {
        s->black = 512.0f;
        s->white = 16383.0f;
    }
    s->gain = 0.5f * iso;
    s->read_sigma = 4.0f;
    for (int i = 0; i < 4; i++) s->pat[i] = pat[i];
    for (int c = 0; c < 3; c++) {
        s->grad[c][0] = (float)r.range(0.08, 0.45);
        s->grad[c][1] = (float)r.range(-0.2, 0.2);
        s->grad[c][2] = (float)r.range(-0.2, 0.2);
        s->grad[c][3] = (float)r.range(-0.15, 0.15);
    }
    const double md = (double)(w < h ? w : h);
    s->npoly = 16 + (int)(r.uniform() * (SYNTH_MAX_POLY - 16));
    for (int p = 0; p < s->npoly; p++) {
        SynthPoly& q = s->poly[p];
        double cx = r.range(0, w), cy = r.range(0, h);
        double rad = md * r.range(0.03, 0.22);
        // vertices at sorted random angles on a circle: a convex polygon
        double ang[SYNTH_POLY_K];
        for (int k = 0; k < SYNTH_POLY_K; k++) ang[k] = r.range(0, 6.283185307179586);
        for (int i = 1; i < SYNTH_POLY_K; i++)
            for (int j = i; j > 0 && ang[j - 1] > ang[j]; j--) {
                double t = ang[j];
                ang[j] = ang[j - 1];
                ang[j - 1] = t;
            }
        double vx[SYNTH_POLY_K], vy[SYNTH_POLY_K];
        double bx0 = 1e30, by0 = 1e30, bx1 = -1e30, by1 = -1e30;
        for (int k = 0; k < SYNTH_POLY_K; k++) {
            vx[k] = cx + rad * std::cos(ang[k]);
            vy[k] = cy + rad * std::sin(ang[k]);
            bx0 = vx[k] < bx0 ? vx[k] : bx0;
            by0 = vy[k] < by0 ? vy[k] : by0;
            bx1 = vx[k] > bx1 ? vx[k] : bx1;
            by1 = vy[k] > by1 ? vy[k] : by1;
        }
        for (int k = 0; k < SYNTH_POLY_K; k++) {
            int k2 = (k + 1) % SYNTH_POLY_K;
            // counter-clockwise in (x, y): inside where cross(e, p - v) >= 0
            double ex = vx[k2] - vx[k], ey = vy[k2] - vy[k];
            q.nx[k] = (float)(-ey);
            q.ny[k] = (float)ex;
            q.d[k] = (float)(ey * vx[k] - ex * vy[k]);
        }
        q.x0 = (float)bx0;
        q.y0 = (float)by0;
        q.x1 = (float)bx1;
        q.y1 = (float)by1;
        for (int c = 0; c < 3; c++) q.col[c] = (float)r.range(0.03, 0.8);
    }
    for (int i = 0; i < SYNTH_NSIN; i++) {
        double cyc = r.range(0.02, 0.3);  // cycles per pixel
        double f = cyc * 6.283185307179586;
        double th = r.range(0, 3.141592653589793);
        s->sn[i].kx = (float)(f * std::cos(th));
        s->sn[i].ky = (float)(f * std::sin(th));
        s->sn[i].phase = (float)r.range(0, 6.283185307179586);
        // amplitude falls off as 1/sqrt(frequency), like natural texture
        double amax = 0.03 * std::sqrt(0.02 / cyc);
        for (int c = 0; c < 3; c++) s->sn[i].amp[c] = (float)r.range(0.0, amax);
    }
}

// Linear scene value of colour channel c at pixel (x, y), in [0.02, 0.85].
__host__ __device__ inline float synth_value(const SynthScene& s, int x, int y, int c) {
    float u = ((float)x + 0.5f) / (float)s.w, v = ((float)y + 0.5f) / (float)s.h;
    float val = s.grad[c][0] + s.grad[c][1] * u + s.grad[c][2] * v + s.grad[c][3] * u * v;
    float px = (float)x + 0.5f, py = (float)y + 0.5f;
    for (int p = 0; p < s.npoly; p++) {
        const SynthPoly& q = s.poly[p];
        if (px < q.x0 || px > q.x1 || py < q.y0 || py > q.y1) continue;
        bool in = true;
        for (int k = 0; k < SYNTH_POLY_K; k++) in = in && (q.nx[k] * px + q.ny[k] * py + q.d[k] >= 0.0f);
        if (in) val = q.col[c];
    }
    for (int i = 0; i < SYNTH_NSIN; i++) val += s.sn[i].amp[c] * sinf(s.sn[i].kx * px + s.sn[i].ky * py + s.sn[i].phase);
    return fminf(fmaxf(val, 0.02f), 0.85f);
}

// Poisson variate with mean lam from the pixel's counter stream.  Small
// means use Knuth's product-of-uniforms method; larger ones the transformed
// rejection method with squeeze (PTRS) of W. Hoermann, "The transformed
// rejection method for generating Poisson random variables", Insurance:
// Mathematics and Economics 12 (1993).
__host__ __device__ inline double synth_poisson(uint64_t seed, uint64_t idx, uint32_t& k, double lam) {
    if (lam <= 0.0) return 0.0;
    if (lam < 10.0) {
        double L = exp(-lam), p = 1.0;
        double n = -1.0;
        do {
            n += 1.0;
            p *= synth_u01(seed, idx, k++);
        } while (p > L && n < 1000.0);
        return n;
    }
    const double slam = sqrt(lam), loglam = log(lam);
    const double b = 0.931 + 2.53 * slam;
    const double a = -0.059 + 0.02483 * b;
    const double inv_alpha = 1.1239 + 1.1328 / (b - 3.4);
    const double vr = 0.9277 - 3.6224 / (b - 2.0);
    for (int it = 0; it < 64; it++) {
        double U = synth_u01(seed, idx, k++) - 0.5;
        double V = synth_u01(seed, idx, k++);
        double us = 0.5 - fabs(U);
        double kk = floor((2.0 * a / us + b) * U + lam + 0.43);
        if (us >= 0.07 && V <= vr) return kk;
        if (kk < 0.0 || (us < 0.013 && V > us)) continue;
        if (log(V) + log(inv_alpha) - log(a / (us * us) + b) <= -lam + kk * loglam - lgamma(kk + 1.0)) return kk;
    }
    return floor(lam + 0.5);
}

// Noisy raw DN at (x, y): electrons ~ Poisson(S / K), DN = black + K e +
// N(0, read_sigma^2), rounded and clamped to [0, white].
__host__ __device__ inline uint16_t synth_dn(const SynthScene& s, int x, int y) {
    int c = s.pat[((y & 1) << 1) | (x & 1)];
    double sig = (double)synth_value(s, x, y, c) * (double)(s.white - s.black);
    if (s.gain <= 0.0f) {  // noise-free rendering
        double dn = floor((double)s.black + sig + 0.5);
        return (uint16_t)(dn > (double)s.white ? (double)s.white : dn);
    }
    uint64_t idx = (uint64_t)y * (uint64_t)s.w + (uint64_t)x;
    uint32_t k = 0;
    double e = synth_poisson(s.seed, idx, k, sig / (double)s.gain);
    double u1 = synth_u01(s.seed, idx, 1000u), u2 = synth_u01(s.seed, idx, 1001u);
    double g = sqrt(-2.0 * log(u1 + 1e-300)) * cos(6.283185307179586 * u2);
    double dn = (double)s.black + (double)s.gain * e + (double)s.read_sigma * g;
    dn = floor(dn + 0.5);
    if (dn < 0.0) dn = 0.0;
    if (dn > (double)s.white) dn = (double)s.white;
    return (uint16_t)dn;
}

}  // namespace ds
