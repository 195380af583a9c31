#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace ds {

// NLM filter strength: either h_param itself, or (when sigma_dev is set)
// h_param = max(*sigma_dev, 1e-4) * sigma_mul, read on the device, as
// enhance.rs:82-84 derives it from the noise estimate.
struct NlmH {
    float h_param;
    const float* sigma_dev;
    float sigma_mul;
};

// Weight parameters as kernels see them.
struct NlmWeight {
    float h_param;
    const float* sigma_dev;
    float sigma_mul;
    float k;  // norm^2 * log2(e)
    __device__ __forceinline__ float h2() const {
        float hp = sigma_dev ? fmaxf(*sigma_dev, 1e-4f) * sigma_mul : h_param;
        return fmaxf(hp * hp, 1e-12f);  // enhance.rs:40
    }
    __device__ __forceinline__ float scale() const { return -k / h2(); }
};

size_t nlm_tile_lds_bytes(int S, int P);

// nlm(p, h, search, patch); search >= 0.
hipError_t launch_nlm(const float* in, float* out, size_t w, size_t h, NlmH hp, long search, size_t patch, hipStream_t stream);

}  // namespace ds
