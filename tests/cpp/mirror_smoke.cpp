// Drives the C++ mirror (include/darkslide.hpp) the way the reference's
// Engine::base_for drives the Rust functions (pipeline/mod.rs:229-252):
// CFA -> demosaic -> denoise -> raw_details, and the Draft demosaic.  Writes
// each stage's output as
// raw float32 files into the directory given as argv[1], for the GPU test to
// compare against the oracle.
#include <cstdio>
#include <string>
#include <vector>

#include "darkslide.hpp"

static bool dump(const std::string& path, const float* p, size_t n) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    size_t k = std::fwrite(p, sizeof(float), n, f);
    std::fclose(f);
    return k == n;
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s OUTDIR\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    const size_t w = 320, h = 240;
    const uint8_t pat[4] = {0, 1, 1, 2};
    std::vector<uint16_t> raw(w * h);
    if (ds_synth_bayer_host(0x5EED0000ull, w, h, 14, 4.0f, pat, raw.data()) != DS_OK) return 1;
    darkslide::CfaImage cfa;
    cfa.w = w;
    cfa.h = h;
    cfa.pattern.colors = {0, 1, 1, 2};
    cfa.data.resize(w * h);
    for (size_t i = 0; i < w * h; i++) cfa.data[i] = ((float)raw[i] - 512.0f) / (16383.0f - 512.0f);  // decode.rs:303
    try {
        auto [img, clip] = darkslide::demosaic(cfa, {1.0f, 1.0f, 1.0f}, darkslide::Quality::Full);
        darkslide::Image3 dn = darkslide::denoise_with(img, 0.0f, 10, 3);
        darkslide::Image3 rl = darkslide::raw_details(dn, 1.0f, 20);
        darkslide::Plane L(w, h);
        for (size_t i = 0; i < w * h; i++) L.data[i] = img.data[i][1];
        float sigma = darkslide::estimate_noise(L);
        darkslide::Plane lf = darkslide::nlm(L, sigma * 0.6f, 10, 3);
        auto [draft, dclip] = darkslide::demosaic(cfa, {2.0f, 1.0f, 1.5f}, darkslide::Quality::Draft);
        bool ok = dump(dir + "/draft.f32", darkslide::flat(draft.data), 3 * (w / 2) * (h / 2)) &&
                  dump(dir + "/draft_clip.f32", dclip.data.data(), (w / 2) * (h / 2)) &&
                  dump(dir + "/cfa.f32", cfa.data.data(), w * h) && dump(dir + "/rgb.f32", darkslide::flat(img.data), 3 * w * h) &&
                  dump(dir + "/clip.f32", clip.data.data(), w * h) && dump(dir + "/denoise.f32", darkslide::flat(dn.data), 3 * w * h) &&
                  dump(dir + "/rl.f32", darkslide::flat(rl.data), 3 * w * h) && dump(dir + "/nlm.f32", lf.data.data(), w * h);
        std::printf("sigma %.9g\n", (double)sigma);
        return ok ? 0 : 1;
    } catch (const darkslide::Error& e) {
        std::fprintf(stderr, "darkslide error %d: %s\n", e.code(), e.what());
        return 3;
    }
}
