// Runs every entry point of the CPU oracle (oracle/oracle.cpp,
// oracle/cnn_oracle.cpp) on small frames, odd and degenerate sizes included,
// in both precisions.  Built with AddressSanitizer and
// UndefinedBehaviorSanitizer (-fno-sanitize-recover), so any out-of-bounds
// access, use after free, signed overflow or invalid shift in the oracle
// ends the run with a report and a non-zero status.  tests/test_oracle_sanitize.py
// runs it.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

extern "C" {
#define DECL(T, S)                                                                                                                  \
    void or_cfa_normalize_u16_##S(const uint16_t*, size_t, size_t, size_t, const uint8_t*, T, const T*, T*);                      \
    void or_cfa_ingest_u16_##S(const uint16_t*, size_t, size_t, size_t, size_t, size_t, const uint8_t*, const T*, size_t, size_t, \
                               size_t, size_t, const T*, size_t, T*, uint8_t*);                                                   \
    void or_superpixel_##S(const T*, size_t, size_t, const uint8_t*, const T*, T*, T*);                                          \
    void or_plane_resize_##S(const T*, size_t, size_t, size_t, size_t, T*);                                                      \
    void or_fast_guided_filter_##S(const T*, const T*, size_t, size_t, size_t, T, size_t, T*);                                   \
    void or_noise_reduction_##S(T*, size_t, size_t, const T*, const T*, T);                                                      \
    void or_sharpen_##S(T*, size_t, size_t, const T*, const T*, T);                                                              \
    void or_super_res_refine_##S(T*, size_t, size_t);                                                                            \
    void or_output_sharpen_##S(T*, size_t, size_t, int, int);                                                                    \
    void or_mhc_bayer_##S(const T*, size_t, size_t, const uint8_t*, const T*, T*);                                               \
    void or_rcd_bayer_##S(const T*, size_t, size_t, const uint8_t*, const T*, T*);                                               \
    void or_clip_plane_##S(const T*, size_t, size_t, T*);                                                                        \
    void or_oklab_from_prophoto_##S(const T*, size_t, T*);                                                                       \
    void or_oklab_to_prophoto_##S(const T*, size_t, T*);                                                                         \
    void or_luminance_##S(const T*, size_t, T*);                                                                                 \
    T or_estimate_noise_##S(const T*, size_t, size_t);                                                                           \
    void or_box_blur_##S(const T*, size_t, size_t, size_t, T*);                                                                  \
    void or_gaussian_blur_##S(const T*, size_t, size_t, T, T*);                                                                  \
    void or_guided_filter_##S(const T*, const T*, size_t, size_t, size_t, T, T*);                                                \
    void or_nlm_##S(const T*, size_t, size_t, T, long, size_t, int, T*);                                                         \
    void or_denoise_##S(const T*, size_t, size_t, T, long, size_t, int, T*);                                                     \
    void or_raw_details_##S(const T*, size_t, size_t, T, size_t, T*);                                                            \
    void or_cfa_clip_levels_##S(const T*, size_t, size_t, const uint8_t*, T*);                                                    \
    void or_reconstruct_highlights_##S(T*, const T*, size_t, size_t, const T*, const T*);
DECL(float, f32)
DECL(double, f64)
#undef DECL
int or_auto_wb_stats(const float*, size_t, const float*, double*, double*, float*);
size_t or_cnn_param_count(void);
void or_cnn_forward_f64(const double*, const float*, int, int, int, int, double*);
double or_cnn_loss_grad_f64(const double*, const float*, const float*, int, int, int, int, int, double*);
void or_cnn_adam_f64(double*, const double*, double*, double*, size_t, int, double);
void or_cnn_input_f64(const float*, int, int, double, double*);
}

static uint64_t rng_state = 0x9E3779B97F4A7C15ull;
static double urand() {  // SplitMix64 -> [0, 1)
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return (double)((z ^ (z >> 31)) >> 11) * (1.0 / 9007199254740992.0);
}
template <class T>
static std::vector<T> rnd(size_t n, double lo, double hi) {
    std::vector<T> v(n);
    for (auto& x : v) x = (T)(lo + (hi - lo) * urand());
    return v;
}

template <class T, class F>
static void run_all(F&& fn) {
    fn((T)0);
}

#define STAGES(T, S)                                                                                                         \
    static void stages_##S(size_t w, size_t h) {                                                                             \
        const size_t n = w * h;                                                                                              \
        const uint8_t pats[4][4] = {{0, 1, 1, 2}, {2, 1, 1, 0}, {1, 0, 2, 1}, {1, 2, 0, 1}};                                 \
        std::vector<uint16_t> raw(n);                                                                                        \
        for (auto& v : raw) v = (uint16_t)(urand() * 16383.0);                                                               \
        const T white[3] = {(T)16383, (T)15000, (T)16000}, mul[3] = {(T)2.0, (T)1.0, (T)1.5};                                 \
        for (const auto& pat : pats) {                                                                                       \
            std::vector<T> cfa(n), rgb(3 * n), clip(n), sp(3 * (w / 2) * (h / 2) + 3), spc((w / 2) * (h / 2) + 1), lv(3);    \
            or_cfa_normalize_u16_##S(raw.data(), w, h, w, pat, (T)512, white, cfa.data());                                    \
            or_rcd_bayer_##S(cfa.data(), w, h, pat, mul, rgb.data());                                                        \
            or_mhc_bayer_##S(cfa.data(), w, h, pat, mul, rgb.data());                                                        \
            or_clip_plane_##S(cfa.data(), w, h, clip.data());                                                                \
            or_superpixel_##S(cfa.data(), w, h, pat, mul, sp.data(), spc.data());                                            \
            or_cfa_clip_levels_##S(cfa.data(), w, h, pat, lv.data());                                                        \
            const T raw_clip[3] = {(T)0.9, (T)0.95, (T)0.92};                                                                \
            or_reconstruct_highlights_##S(rgb.data(), clip.data(), w, h, mul, raw_clip);                                     \
        }                                                                                                                    \
        /* ingestion with a crop and a 2x2 black table of 3 values per entry */                                            \
        if (w > 2 && h > 2) {                                                                                                \
            const T black[12] = {500, 510, 505, 512, 498, 501, 507, 509, 511, 503, 502, 506};                                \
            std::vector<T> out((w - 1) * (h - 2));                                                                           \
            uint8_t pat_out[4];                                                                                              \
            or_cfa_ingest_u16_##S(raw.data(), w, 1, 2, w - 1, h - 2, pats[0], black, 12, 2, 2, 3, white, 3, out.data(), pat_out); \
        }                                                                                                                    \
        auto rgb = rnd<T>(3 * n, 0.0, 1.2);                                                                                  \
        std::vector<T> lab(3 * n), back(3 * n), L(n), o(n), o3(3 * n);                                                      \
        or_oklab_from_prophoto_##S(rgb.data(), n, lab.data());                                                               \
        or_oklab_to_prophoto_##S(lab.data(), n, back.data());                                                                \
        or_luminance_##S(rgb.data(), n, L.data());                                                                           \
        const T sigma = or_estimate_noise_##S(L.data(), w, h);                                                               \
        (void)sigma;                                                                                                         \
        for (size_t r : {0, 1, 3, 9}) or_box_blur_##S(L.data(), w, h, r, o.data());                                          \
        for (T s : {(T)0.2, (T)1.0, (T)2.5, (T)6.0}) or_gaussian_blur_##S(L.data(), w, h, s, o.data());                       \
        auto p2 = rnd<T>(n, 0.0, 1.0);                                                                                       \
        or_guided_filter_##S(L.data(), p2.data(), w, h, 2, (T)1e-3, o.data());                                               \
        for (int threads : {1, 3}) {                                                                                         \
            or_nlm_##S(L.data(), w, h, (T)0.05, 3, 1, threads, o.data());                                                    \
            or_nlm_##S(L.data(), w, h, (T)0.05, 2, 2, threads, o.data());                                                    \
            or_denoise_##S(rgb.data(), w, h, (T)37, 3, 1, threads, o3.data());                                               \
        }                                                                                                                    \
        or_nlm_##S(L.data(), w, h, (T)0.05, -1, 1, 1, o.data());                                                             \
        or_raw_details_##S(rgb.data(), w, h, (T)1.0, 3, o3.data());                                                          \
        or_raw_details_##S(rgb.data(), w, h, (T)0.75, 2, o3.data());                                                         \
        for (size_t w2 : {(size_t)1, w / 2 + 1, w + 3})                                                                      \
            for (size_t h2 : {(size_t)1, h / 3 + 1}) {                                                                       \
                std::vector<T> rs(w2 * h2);                                                                                  \
                or_plane_resize_##S(L.data(), w, h, w2, h2, rs.data());                                                      \
            }                                                                                                                \
        for (size_t sub : {1, 2, 4}) or_fast_guided_filter_##S(L.data(), p2.data(), w, h, 3, (T)1e-3, sub, o.data());         \
        const T nr[6] = {50, 50, 30, 25, 50, 50}, sh[4] = {40, 1.0, 25, 20};                                                 \
        auto field = rnd<T>(n, -1.0, 1.0);                                                                                   \
        std::vector<T> img = rgb;                                                                                            \
        or_noise_reduction_##S(img.data(), w, h, nr, nullptr, (T)1);                                                         \
        or_noise_reduction_##S(img.data(), w, h, nr, field.data(), (T)0.5);                                                  \
        or_sharpen_##S(img.data(), w, h, sh, nullptr, (T)1);                                                                 \
        or_sharpen_##S(img.data(), w, h, sh, field.data(), (T)0.5);                                                          \
        or_super_res_refine_##S(img.data(), w, h);                                                                           \
        for (int mode = 0; mode < 4; mode++)                                                                                 \
            for (int amt = 0; amt < 3; amt++) or_output_sharpen_##S(img.data(), w, h, mode, amt);                            \
    }
STAGES(float, f32)
STAGES(double, f64)
#undef STAGES

static void cnn(int n, int H, int W) {
    const size_t np = or_cnn_param_count();
    auto params = rnd<double>(np, -0.1, 0.1);
    auto x = rnd<float>((size_t)n * H * W, -2.0, 2.0), t = rnd<float>((size_t)n * H * W, -2.0, 2.0);
    std::vector<double> y((size_t)n * H * W), grad(np), m(np, 0.0), v(np, 0.0), xin((size_t)H * W);
    for (int bf16 = 0; bf16 < 2; bf16++) {
        or_cnn_forward_f64(params.data(), x.data(), n, H, W, bf16, y.data());
        or_cnn_loss_grad_f64(params.data(), x.data(), t.data(), n, H, W, H > 4 && W > 4 ? 1 : 0, bf16, grad.data());
    }
    for (int step = 1; step <= 2; step++) or_cnn_adam_f64(params.data(), grad.data(), m.data(), v.data(), np, step, 1e-3);
    or_cnn_input_f64(x.data(), W, H, 0.05, xin.data());
}

int main() {
    const size_t sizes[][2] = {{1, 1}, {2, 2}, {3, 2}, {5, 3}, {7, 5}, {16, 16}, {33, 17}, {64, 48}};
    for (const auto& s : sizes) {
        stages_f32(s[0], s[1]);
        stages_f64(s[0], s[1]);
        std::printf("stages %zux%zu ok\n", s[0], s[1]);
    }
    for (const auto& s : {std::pair<int, int>{1, 1}, {5, 7}, {19, 13}}) {
        cnn(1, s.first, s.second);
        cnn(2, s.first, s.second);
    }
    auto rgb = rnd<float>(3 * 1000, 0.0, 1.0);
    const float n_as[3] = {0.5f, 1.0f, 0.7f};
    double acc[3], cnt;
    float neutral[3];
    or_auto_wb_stats(rgb.data(), 1000, n_as, acc, &cnt, neutral);
    std::printf("cnn, auto-wb ok\n");
    return 0;
}
