// Exhaustive checks of the kernels' libm restatements (darkslide_amd/csrc/
// libm_ref.h) against the host's glibc, which the reference's f32::cbrt and
// f32::exp call.
//   libm_exhaustive check  compares ds_cbrtf / ds_expf (built here for the
//                          CPU) with cbrtf / expf on all 2^32 inputs and
//                          prints the mismatch count of each; exit 1 if any
//   libm_exhaustive hash   prints the order-free hash (libm_hash_term) of
//                          glibc's cbrtf and expf over all 2^32 inputs, to
//                          compare with ds_libm_hash on the GPU
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#define DS_HD
#include "libm_ref.h"

namespace {

// f(i, a, b) over all 2^32 inputs on up to 16 threads, each adding into
// its own pair of counters; returns the sums of both.
template <class F>
void sweep(F f, uint64_t& sa, uint64_t& sb) {
    unsigned nt = std::thread::hardware_concurrency();
    nt = nt < 1 ? 1 : (nt > 16 ? 16 : nt);
    std::vector<uint64_t> ra(nt, 0), rb(nt, 0);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([=, &ra, &rb] {
            const uint64_t lo = ((1ull << 32) * t) / nt, hi = ((1ull << 32) * (t + 1)) / nt;
            uint64_t a = 0, b = 0;
            for (uint64_t i = lo; i < hi; i++) f((uint32_t)i, a, b);
            ra[t] = a;
            rb[t] = b;
        });
    for (auto& x : th) x.join();
    sa = sb = 0;
    for (unsigned t = 0; t < nt; t++) sa += ra[t], sb += rb[t];
}

float as_float(uint32_t i) {
    float x;
    memcpy(&x, &i, 4);
    return x;
}

bool same(float a, float b) {
    uint32_t ab, bb;
    memcpy(&ab, &a, 4);
    memcpy(&bb, &b, 4);
    return ab == bb || (std::isnan(a) && std::isnan(b));
}

}  // namespace

int main(int argc, char** argv) {
    const std::string mode = argc > 1 ? argv[1] : "check";
    if (mode == "check") {
        uint64_t m1, m2;
        sweep(
            [](uint32_t i, uint64_t& a, uint64_t& b) {
                const float x = as_float(i);
                a += !same(cbrtf(x), ds::ds_cbrtf(x));
                b += !same(expf(x), ds::ds_expf(x));
            },
            m1, m2);
        printf("cbrtf mismatches %llu\nexpf mismatches %llu\n", (unsigned long long)m1, (unsigned long long)m2);
        return (m1 || m2) ? 1 : 0;
    }
    if (mode == "hash") {
        uint64_t h1, h2;
        sweep(
            [](uint32_t i, uint64_t& a, uint64_t& b) {
                const float x = as_float(i);
                a += ds::libm_hash_term(i, cbrtf(x));
                b += ds::libm_hash_term(i, expf(x));
            },
            h1, h2);
        printf("cbrtf %016llx\nexpf %016llx\n", (unsigned long long)h1, (unsigned long long)h2);
        return 0;
    }
    fprintf(stderr, "usage: %s check|hash\n", argv[0]);
    return 2;
}
