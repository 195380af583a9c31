#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

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

// Regions of the learned denoiser's NLM spot check and fallback: square
// blocks of NLM_REGION pixels, row-major, ceil(w / NLM_REGION) per row.
constexpr int NLM_REGION = 256;
constexpr int NLM_SPOT = 16;  // spot-check lattice per region side

size_t nlm_tile_lds_bytes(int S, int P);
size_t nlm_regions(size_t w, size_t h);
size_t nlm_spot_scratch_floats(size_t w, size_t h);

// nlm(p, h, search, patch); search >= 0.
hipError_t launch_nlm(const float* in, float* out, size_t w, size_t h, NlmH hp, long search, size_t patch, hipStream_t stream);

// The denoise composite's guide planes formed from the NLM result lf of L
// (enhance.rs:86-89, image.rs:386-391): I = L + (lf - L) m, I A, I B and I I,
// each rounded as a plain FP32 expression, as k_gf_inputs forms them.
struct NlmGuide {
    const float* L;
    const float* A;
    const float* B;
    float m;
    float* I;
    float* IA;
    float* IB;
    float* II;
};
// NLM of L whose output stage stores the guide planes g instead of lf: the
// tile kernel forms them from each result as it stores it (*fused = true,
// out untouched).  Where the tile kernel does not take (search, patch), lf
// goes to out and *fused is false.
hipError_t launch_nlm_guide(const float* L, float* out, size_t w, size_t h, NlmH hp, long search, size_t patch, const NlmGuide& g, bool* fused,
                            hipStream_t stream);

// NLM over the pixels of the flagged regions only (regions[r] != 0); out
// keeps its values elsewhere.  regions == nullptr: every pixel.
hipError_t launch_nlm_regions(const float* in, float* out, size_t w, size_t h, NlmH hp, long search, size_t patch, const uint8_t* regions,
                              hipStream_t stream);

// Per-region deviation of `cand` from NLM in dB (gap_db) and the fallback
// flags (gap_db > thr_db, or every region when thr_db <= 0); hp.sigma_dev
// must be set; scratch: nlm_spot_scratch_floats(w, h).
hipError_t launch_nlm_spot_check(const float* in, const float* cand, size_t w, size_t h, NlmH hp, long search, size_t patch, float thr_db,
                                 float* gap_db, uint8_t* flags, float* scratch, hipStream_t stream);

}  // namespace ds
