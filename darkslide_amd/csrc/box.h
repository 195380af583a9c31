#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace ds {

constexpr int BOX_MAX_PLANES = 6;

struct BoxPlanes {
    float* p[BOX_MAX_PLANES];
    const int* gate = nullptr;  // when set, the pass does nothing while *gate == 0
};
struct CBoxPlanes {
    const float* p[BOX_MAX_PLANES];
};

// Input of a horizontal pass: plane k at pixel i.
struct BoxInPlanes {
    CBoxPlanes src;
    __device__ __forceinline__ float get(size_t i, int k) const { return src.p[k][i]; }
};

// Exact running-sum passes of image.rs:276-311 over `planes` planes.
hipError_t box_h_planes(CBoxPlanes in, BoxPlanes out, int planes, size_t w, size_t h, size_t r, hipStream_t st);
hipError_t box_v_planes(CBoxPlanes in, BoxPlanes out, int planes, size_t w, size_t h, size_t r, hipStream_t st);

// box_blur(in, r) -> out (image.rs:269); tmp: w*h floats.
hipError_t launch_box_blur(const float* in, float* tmp, float* out, size_t w, size_t h, size_t r, hipStream_t st);

}  // namespace ds
