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

// Guided-filter inputs of denoise (enhance.rs:86, image.rs:386-391):
// I = L + (LF - L) m; planes [I, a, b, I a, I b, I I].
struct BoxInGuided {
    const float* L;
    const float* LF;
    const float* A;
    const float* B;
    float m;
    __device__ __forceinline__ float get(size_t i, int k) const {
        const float o = L[i];
        const float I = o + (LF[i] - o) * m;
        switch (k) {
            case 0: return I;
            case 1: return A[i];
            case 2: return B[i];
            case 3: return I * A[i];
            case 4: return I * B[i];
            default: return I * I;
        }
    }
};

// Inputs of guided_filter(I, p) (image.rs:386-391): planes [I, p, I p, I I].
struct BoxInGuide1 {
    const float* I;
    const float* P;
    __device__ __forceinline__ float get(size_t i, int k) const {
        switch (k) {
            case 0: return I[i];
            case 1: return P[i];
            case 2: return I[i] * P[i];
            default: return I[i] * I[i];
        }
    }
};

// Exact running-sum passes of image.rs:276-311 over `planes` planes.
hipError_t box_h_planes(CBoxPlanes in, BoxPlanes out, int planes, size_t w, size_t h, size_t r, hipStream_t st);
hipError_t box_h_guided(BoxInGuided in, BoxPlanes out, size_t w, size_t h, size_t r, hipStream_t st);
hipError_t box_h_guide1(BoxInGuide1 in, BoxPlanes out, size_t w, size_t h, size_t r, hipStream_t st);
hipError_t box_v_planes(CBoxPlanes in, BoxPlanes out, int planes, size_t w, size_t h, size_t r, hipStream_t st);

// box_blur(in, r) -> out (image.rs:269); tmp: w*h floats.
hipError_t launch_box_blur(const float* in, float* tmp, float* out, size_t w, size_t h, size_t r, hipStream_t st);

}  // namespace ds
