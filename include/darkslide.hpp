// darkslide.hpp - C++ mirror of the reference engine's detail-path API
// (Syha-01/darkslide crates/core/src), header-only over the C ABI in
// darkslide_gpu.h.  Types and functions keep the reference's names,
// layouts and argument meaning:
//   Image3 / Plane        image.rs:7-21
//   CfaPattern / CfaImage decode.rs:16-56
//   demosaic              demosaic.rs:20
//   estimate_noise, nlm, denoise, raw_details   enhance.rs:10, :36, :75, :96
// The reference functions are infallible; here a GPU failure throws
// darkslide::Error (there is no CPU fallback in this library).
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "darkslide_gpu.h"

namespace darkslide {

using Rgb = std::array<float, 3>;

struct Image3 {
    size_t w = 0, h = 0;
    std::vector<Rgb> data;  // interleaved, row-major
    Image3() = default;
    Image3(size_t w_, size_t h_) : w(w_), h(h_), data(w_ * h_, Rgb{0.0f, 0.0f, 0.0f}) {}
    Rgb get(size_t x, size_t y) const { return data[y * w + x]; }
};

struct Plane {
    size_t w = 0, h = 0;
    std::vector<float> data;
    Plane() = default;
    Plane(size_t w_, size_t h_) : w(w_), h(h_), data(w_ * h_, 0.0f) {}
    float get(size_t x, size_t y) const { return data[y * w + x]; }
};

// Colours 0 = R, 1 = G, 2 = B.
struct CfaPattern {
    size_t width = 2, height = 2;
    std::vector<uint8_t> colors;
    size_t color_at(size_t x, size_t y) const { return colors[(y % height) * width + (x % width)]; }
    bool is_bayer() const { return width == 2 && height == 2; }
};

struct CfaImage {
    size_t w = 0, h = 0;
    std::vector<float> data;  // black-subtracted, 1.0 = clipping
    CfaPattern pattern;
};

enum class Quality { Draft, Full };

class Error : public std::runtime_error {
   public:
    Error(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

   private:
    int code_;
};

// One shared context per process on device 0, as the reference shares its
// rayon pool.  Calls on it are serialised by the library.
class Gpu {
   public:
    static ds_ctx* ctx() {
        static std::once_flag once;
        static ds_ctx* c = nullptr;
        static int rc = 0;
        std::call_once(once, [] { rc = ds_ctx_create(0, &c); });
        if (rc != DS_OK) throw Error(rc, std::string("darkslide GPU context: ") + ds_last_error(nullptr));
        return c;
    }
    static void check(int rc) {
        if (rc != DS_OK) throw Error(rc, ds_last_error(ctx()));
    }
};

inline const float* flat(const std::vector<Rgb>& v) { return v.empty() ? nullptr : v.front().data(); }
inline float* flat(std::vector<Rgb>& v) { return v.empty() ? nullptr : v.front().data(); }

// demosaic.rs:20.  On a Bayer CFA, Full quality runs RCD (demosaic.rs:147)
// and the clip plane (demosaic.rs:31), Draft the superpixel
// (demosaic.rs:44-87, one pixel per 2x2 tile), on the GPU.  X-Trans is
// outside this library and throws.
inline std::pair<Image3, Plane> demosaic(const CfaImage& cfa, Rgb mul, Quality quality) {
    if (!cfa.pattern.is_bayer()) throw Error(DS_ERR_INVALID, "only Bayer CFAs run on the GPU");
    uint8_t pat[4];
    for (int i = 0; i < 4; i++) pat[i] = (uint8_t)cfa.pattern.color_at((size_t)(i & 1), (size_t)(i >> 1));
    if (quality == Quality::Draft) {
        Image3 img(cfa.w / 2, cfa.h / 2);
        Plane clip(cfa.w / 2, cfa.h / 2);
        Gpu::check(ds_superpixel_f32(Gpu::ctx(), cfa.data.data(), cfa.w, cfa.h, pat, mul.data(), flat(img.data), clip.data.data()));
        return {std::move(img), std::move(clip)};
    }
    Image3 img(cfa.w, cfa.h);
    Plane clip(cfa.w, cfa.h);
    Gpu::check(ds_demosaic_f32(Gpu::ctx(), cfa.data.data(), cfa.w, cfa.h, pat, mul.data(), DS_ALGO_RCD, flat(img.data), clip.data.data()));
    return {std::move(img), std::move(clip)};
}

// enhance.rs:10
inline float estimate_noise(const Plane& p) {
    float s = 0.0f;
    Gpu::check(ds_estimate_noise(Gpu::ctx(), p.data.data(), p.w, p.h, &s));
    return s;
}

// enhance.rs:36
inline Plane nlm(const Plane& p, float h_param, std::ptrdiff_t search, size_t patch) {
    Plane out(p.w, p.h);
    Gpu::check(ds_nlm(Gpu::ctx(), p.data.data(), p.w, p.h, h_param, (long)search, patch, out.data.data()));
    return out;
}

// enhance.rs:75 (NLM search 3, patch 1 as in the reference)
inline Image3 denoise(const Image3& img, float amount) {
    Image3 out(img.w, img.h);
    Gpu::check(ds_denoise(Gpu::ctx(), flat(img.data), img.w, img.h, amount, flat(out.data)));
    return out;
}

// enhance.rs:75 with the NLM search / patch radii as parameters
inline Image3 denoise_with(const Image3& img, float amount, std::ptrdiff_t search, size_t patch) {
    Image3 out(img.w, img.h);
    Gpu::check(ds_denoise_ex(Gpu::ctx(), flat(img.data), img.w, img.h, amount, (long)search, patch, flat(out.data)));
    return out;
}

// enhance.rs:96
inline Image3 raw_details(const Image3& img, float sigma, size_t iterations) {
    Image3 out(img.w, img.h);
    Gpu::check(ds_raw_details(Gpu::ctx(), flat(img.data), img.w, img.h, sigma, iterations, flat(out.data)));
    return out;
}

// image.rs:269 / :327 / :385
inline Plane box_blur(const Plane& p, size_t r) {
    Plane out(p.w, p.h);
    Gpu::check(ds_box_blur(Gpu::ctx(), p.data.data(), p.w, p.h, r, out.data.data()));
    return out;
}
inline Plane gaussian_blur(const Plane& p, float sigma) {
    Plane out(p.w, p.h);
    Gpu::check(ds_gaussian_blur(Gpu::ctx(), p.data.data(), p.w, p.h, sigma, out.data.data()));
    return out;
}
inline Plane guided_filter(const Plane& i, const Plane& p, size_t r, float eps) {
    Plane out(i.w, i.h);
    Gpu::check(ds_guided_filter(Gpu::ctx(), i.data.data(), p.data.data(), i.w, i.h, r, eps, out.data.data()));
    return out;
}

}  // namespace darkslide
