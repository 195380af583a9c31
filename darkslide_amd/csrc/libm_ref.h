// Single-precision math functions that return the same bits as the x86_64
// glibc the reference links against (Rust's f32::cbrt and f32::exp call the
// platform libm's cbrtf and expf), for host and device.  Each is a
// restatement of a published algorithm with every rounding point kept; the
// tests compare all 2^32 inputs against the host libm.
//
// ds_cbrtf restates the published algorithm of glibc's generic single-
// precision cube root (sysdeps/ieee754/flt-32/s_cbrtf.c, Alboth and
// Drepper, 1997), which x86_64 uses: split |x| = m 2^e with m in [0.5, 1)
// (frexp), a quadratic initial estimate u of cbrt(m) evaluated in double
// and rounded to float, its cube t in float, one Halley step
// u (t + 2m) / (2t + m) in double, multiplied by 2^((e mod 3)/3) in double
// and rounded to float, then scaled by 2^(e/3) (C's truncating / and %).
// Zero, infinity and NaN return x + x.  Every rounding point is kept, so
// with FMA contraction off the result equals glibc's for every input; the
// test compares all 2^32 inputs.
#pragma once

#include <cstdint>
#include <cstring>

#ifndef DS_HD
#define DS_HD __host__ __device__
#endif

namespace ds {

DS_HD inline float ds_cbrtf(float x) {
#pragma clang fp contract(off)
    uint32_t bits;
    memcpy(&bits, &x, 4);
    const uint32_t ab = bits & 0x7fffffffu;
    if (ab == 0 || ab >= 0x7f800000u) return x + x;  // +-0, +-inf, NaN
    // frexp of |x|: m in [0.5, 1), |x| = m 2^e
    int e;
    uint32_t mb;
    if (ab < 0x00800000u) {  // subnormal: normalise
        int lz = 0;
        uint32_t f = ab;
        while (!(f & 0x00800000u)) {
            f <<= 1;
            lz++;
        }
        e = -125 - lz;
        mb = (f & 0x007fffffu) | 0x3f000000u;
    } else {
        e = (int)(ab >> 23) - 126;
        mb = (ab & 0x007fffffu) | 0x3f000000u;
    }
    float xm;
    memcpy(&xm, &mb, 4);
    const double d_xm = (double)xm;
    const float u = (float)(0.492659620528969547 + (0.697570460207922770 - 0.191502161678719066 * d_xm) * d_xm);
    const float t2 = u * u * u;
    const double factor[5] = {1.0 / 1.5874010519681994748, 1.0 / 1.2599210498948731648, 1.0, 1.2599210498948731648,
                              1.5874010519681994748};
    const float ym = (float)((double)u * ((double)t2 + 2.0 * d_xm) / (2.0 * (double)t2 + d_xm) * factor[2 + e % 3]);
    // ldexpf(+-ym, e / 3): ym lies in [0.79, 1.26] and e / 3 in [-49, 42],
    // so the product is a normal float and the scaling is exact
    const int q = e / 3;
    uint32_t yb;
    memcpy(&yb, &ym, 4);
    yb += (uint32_t)q << 23;
    yb |= bits & 0x80000000u;
    float r;
    memcpy(&r, &yb, 4);
    return r;
}

// ds_expf restates the published algorithm of glibc's single-precision
// exponential (sysdeps/ieee754/flt-32/e_expf.c and e_exp2f_data.c, from
// Szabolcs Nagy's optimized-routines, glibc 2.28 on): in double,
// k = x 32/ln2 rounded to nearest by adding and subtracting 1.5 2^52,
// r = x 32/ln2 - k, s = 2^(k/32) from a 32-entry table of
// bits(RN(2^(i/32))) - (i << 47) with k's high bits added to the exponent,
// and exp(x) = s (C0 r^3 + C1 r^2 + C2 r + 1) (minimax cubic for
// 2^(r/32) on |r| <= 1/2) rounded once to float.  |x| >= 88 takes the
// special cases first: -inf -> 0, +inf/NaN -> x + x, overflow above
// ln(2^128) -> +inf, underflow below ln(2^-150) -> 0.
// On x86_64 CPUs with FMA, glibc's ifunc selects its FMA build of this
// function (sysdeps/x86_64/fpu/multiarch), where the compiler fuses the
// multiply-adds; the fused forms are kept here.  The unfused reduction
// r = z - k differs from it on 2 of the 2^32 inputs (by one ulp).
namespace expf_detail {
constexpr uint64_t kTab[32] = {
    0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full, 0x3fef9301d0125b51ull,
    0x3fef72b83c7d517bull, 0x3fef54873168b9aaull, 0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull,
    0x3fef06fe0a31b715ull, 0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,
    0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull, 0x3feea47eb03a5585ull,
    0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull, 0x3feea11473eb0187ull, 0x3feea589994cce13ull,
    0x3feeace5422aa0dbull, 0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,
    0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull, 0x3fef3720dcef9069ull,
    0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full, 0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull,
};
}  // namespace expf_detail

DS_HD inline float ds_expf(float x) {
#pragma clang fp contract(off)
    uint32_t bits;
    memcpy(&bits, &x, 4);
    const uint32_t abstop = (bits >> 20) & 0x7ffu;
    if (abstop >= 0x42bu) {  // |x| >= 88, or NaN
        if (bits == 0xff800000u) return 0.0f;
        if (abstop >= 0x7f8u) return x + x;
        if (x > 0x1.62e42ep6f) return __builtin_inff();
        if (x < -0x1.9fe368p6f) return 0.0f;
    }
    const double xd = (double)x;
    const double inv_ln2_n = 0x1.71547652b82fep+0 * 32.0;
    double kd = __builtin_fma(inv_ln2_n, xd, 0x1.8p+52);
    uint64_t ki;
    memcpy(&ki, &kd, 8);
    kd -= 0x1.8p+52;
    const double r = __builtin_fma(inv_ln2_n, xd, -kd);
    uint64_t t = expf_detail::kTab[ki % 32];
    t += ki << 47;
    double s;
    memcpy(&s, &t, 8);
    const double c0 = 0x1.c6af84b912394p-5 / 32.0 / 32.0 / 32.0, c1 = 0x1.ebfce50fac4f3p-3 / 32.0 / 32.0,
                 c2 = 0x1.62e42ff0c52d6p-1 / 32.0;
    const double zz = __builtin_fma(c0, r, c1);
    const double r2 = r * r;
    double y = __builtin_fma(c2, r, 1.0);
    y = __builtin_fma(zz, r2, y);
    y = y * s;
    return (float)y;
}

// One term of the order-free hash of a function over all 2^32 inputs:
// H = sum over i of mix(i << 32 | bits(f(x_i))) mod 2^64, mix being
// SplitMix64's finaliser and every NaN result taken as 0x7fc00000 (NaN
// payloads are not part of the comparison).  Equal hashes of two
// implementations mean equal outputs up to a 2^-64 chance.
DS_HD inline uint64_t libm_hash_term(uint32_t i, float y) {
    uint32_t b;
    memcpy(&b, &y, 4);
    if ((b & 0x7fffffffu) > 0x7f800000u) b = 0x7fc00000u;
    uint64_t z = ((uint64_t)i << 32 | b) + 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

}  // namespace ds
