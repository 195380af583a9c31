/*
 * darkslide_gpu.h - C ABI of the MI355X (gfx950) raw detail path:
 * Bayer demosaic (RCD / MHC), NLM luminance denoise and Richardson-Lucy
 * "raw details" sharpening.
 *
 * Each entry point replaces one function of the reference engine
 * (Syha-01/darkslide, crates/core/src); the replaced function is cited
 * beside it.  Buffer layouts are the reference's in-memory layouts, so a
 * Rust caller passes `vec.as_ptr()` straight through:
 *   - Plane:  w*h f32, row-major, tightly packed (image.rs:15-21)
 *   - Image3: w*h interleaved [r,g,b] f32 triples (image.rs:7-13)
 *   - CFA:    w*h f32 normalised samples (decode.rs:49-56), or raw u16 DN
 *   - Bayer pattern: 4 bytes, colour (0=R,1=G,2=B) of (x%2, y%2) at
 *     pattern[(y%2)*2 + x%2] (decode.rs:16-28); greens must form a
 *     checkerboard (RGGB, BGGR, GRBG, GBRG).
 *
 * Host-pointer entry points are synchronous: they copy in, run on the
 * context's HIP stream, copy out and return.  The *_dev entry points take
 * device pointers, only enqueue on the context's stream and return; call
 * ds_sync() before reading results.
 *
 * Every call returns DS_OK (0) or a negative DS_ERR_* code;
 * ds_last_error() gives the message of the last failure on that context.
 * Images of 2^32 pixels or more are refused with DS_ERR_INVALID.
 * A context serialises its calls with an internal mutex, so it may be shared
 * between threads; one context per thread gives concurrency.
 */
#ifndef DARKSLIDE_GPU_H
#define DARKSLIDE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DS_OK 0
#define DS_ERR_INVALID (-1) /* bad argument (size, pattern, null pointer) */
#define DS_ERR_DEVICE (-2)  /* no usable GPU, or a HIP runtime failure */
#define DS_ERR_ALLOC (-3)   /* device memory allocation failed */
#define DS_ERR_LAUNCH (-4)  /* kernel launch failed */

#define DS_ALGO_RCD 0 /* demosaic.rs:147 rcd_bayer (MHC below 16 px) */
#define DS_ALGO_MHC 1 /* demosaic.rs:90 mhc_bayer */

typedef struct ds_ctx ds_ctx;

/* Context on HIP device `device` with its own stream. */
int ds_ctx_create(int device, ds_ctx** out);
void ds_ctx_destroy(ds_ctx* ctx);
/* Message of the last failure on ctx (ctx may be NULL: last create failure). */
const char* ds_last_error(const ds_ctx* ctx);
/* Library version string. */
const char* ds_version(void);

/* ---- demosaic ---------------------------------------------------------- */

/* demosaic.rs:20 demosaic(cfa, mul, Quality::Full) for a Bayer CFA:
 * rcd_bayer (demosaic.rs:147) or mhc_bayer (demosaic.rs:90), and
 * clip_plane (demosaic.rs:31).  cfa: w*h normalised samples; mul: as-shot
 * WB multipliers (pipeline/mod.rs:166-169).  rgb_out: w*h*3; clip_out:
 * w*h or NULL. */
int ds_demosaic_f32(ds_ctx* ctx, const float* cfa, size_t w, size_t h, const uint8_t pattern[4], const float mul[3], int algo,
                    float* rgb_out, float* clip_out);

/* Same, fused with the CFA normalisation of decode.rs:291-306:
 * v = (raw - black) / max(white[c] - black, 1).  raw: h rows of `stride`
 * u16 samples. */
int ds_demosaic_u16(ds_ctx* ctx, const uint16_t* raw, size_t w, size_t h, size_t stride, const uint8_t pattern[4], float black,
                    const float white[3], const float mul[3], int algo, float* rgb_out, float* clip_out);

/* ---- raw ingestion and Draft demosaic ----------------------------------- */

/* A Bayer raw buffer as load_raw has it after rawler's decode
 * (decode.rs:188-307): `width` x `height` samples, the crop rectangle kept
 * (crop_area or active_area, decode.rs:251-254), the CFA colour at the
 * buffer origin, the black level table black_at (decode.rs:258-263: one
 * value, or black_w x black_h entries of black_cpp values, channel 0 used)
 * and the white level per colour (white_at, decode.rs:290).  Black tables
 * whose period exceeds 2 (non-Bayer sensors) are refused. */
typedef struct ds_raw_cfa {
    const uint16_t* data; /* host or device per entry point */
    size_t width, height;
    size_t crop_x, crop_y, crop_w, crop_h;
    uint8_t pattern[4];
    const float* black; /* host pointer, nblack values */
    size_t nblack, black_w, black_h, black_cpp;
    float white[3];
} ds_raw_cfa;

#define DS_QUALITY_FULL 0
#define DS_QUALITY_DRAFT 1

/* The CfaImage load_raw builds (decode.rs:292-306): the normalised crop
 * (crop_w x crop_h floats) and its pattern (CfaPattern::shifted, decode.rs:38). */
int ds_cfa_ingest(ds_ctx* ctx, const ds_raw_cfa* raw, float* cfa_out, uint8_t pattern_out[4]);

/* demosaic(cfa, mul, quality) (demosaic.rs:20) of the crop, normalised on
 * the fly: DS_QUALITY_FULL gives crop_w x crop_h (RCD or MHC), DS_QUALITY_DRAFT
 * the superpixel image (demosaic.rs:44-87) of crop_w/2 x crop_h/2.  clip_out
 * (same size as one plane of rgb_out) may be NULL. */
int ds_demosaic_raw(ds_ctx* ctx, const ds_raw_cfa* raw, const float mul[3], int algo, int quality, float* rgb_out, float* clip_out);

/* cfa_clip_levels (demosaic.rs:419) of the crop's normalised CFA. */
int ds_cfa_clip_levels_raw(ds_ctx* ctx, const ds_raw_cfa* raw, float levels_out[3]);

/* demosaic(cfa, mul, Quality::Draft) on a normalised CFA: superpixel
 * (demosaic.rs:44-87), rgb_out (w/2)*(h/2)*3, clip_out (w/2)*(h/2) or NULL. */
int ds_superpixel_f32(ds_ctx* ctx, const float* cfa, size_t w, size_t h, const uint8_t pattern[4], const float mul[3], float* rgb_out,
                      float* clip_out);

/* ---- highlight reconstruction ------------------------------------------ */

/* demosaic.rs:419 cfa_clip_levels(cfa): per-colour max of the normalised
 * CFA over every third row, kept as the clip level when >= 0.95, else
 * FLT_MAX.  levels_out: 3 floats (host). */
int ds_cfa_clip_levels_f32(ds_ctx* ctx, const float* cfa, size_t w, size_t h, const uint8_t pattern[4], float levels_out[3]);
/* Same on raw DN, normalised as in ds_demosaic_u16. */
int ds_cfa_clip_levels_u16(ds_ctx* ctx, const uint16_t* raw, size_t w, size_t h, size_t stride, const uint8_t pattern[4], float black,
                           const float white[3], float levels_out[3]);

/* demosaic.rs:435 reconstruct_highlights(img, clip, mul, raw_clip), in
 * place on rgb (w*h*3, white-balanced camera RGB from ds_demosaic_*); clip:
 * the clip plane ds_demosaic_* returns; raw_clip: ds_cfa_clip_levels'
 * result.  As called by Engine::base (pipeline/mod.rs:186-187). */
int ds_reconstruct_highlights(ds_ctx* ctx, float* rgb, const float* clip, size_t w, size_t h, const float mul[3], const float raw_clip[3]);

/* ---- detail (pipeline/detail.rs) ----------------------------------------- */

/* NoiseReduction (params.rs:319-326) and Sharpening (params.rs:304-309), in
 * the reference's 0..100 units. */
typedef struct ds_nr_params {
    float luminance, luminance_detail, luminance_contrast, color, color_detail, color_smoothness;
} ds_nr_params;
typedef struct ds_sharpen_params {
    float amount, radius, detail, masking;
} ds_sharpen_params;

/* detail::noise_reduction(img, nr, fields, scale, ok) (pipeline/detail.rs:13),
 * in place on interleaved linear ProPhoto (w*h*3).  noise_field: the
 * LocalFields noise plane (w*h, the reference's units) or NULL when no mask
 * sets it; scale: output pixels per full-resolution pixel. */
int ds_noise_reduction(ds_ctx* ctx, float* rgb, size_t w, size_t h, const ds_nr_params* nr, const float* noise_field, float scale);

/* detail::sharpen(img, sh, fields, scale) (pipeline/detail.rs:96), in place;
 * sharpness_field: the LocalFields sharpness plane or NULL. */
int ds_sharpen(ds_ctx* ctx, float* rgb, size_t w, size_t h, const ds_sharpen_params* sh, const float* sharpness_field, float scale);

/* fast_guided_filter(i, p, r, eps, sub) (image.rs:408) and Plane::resize(w2, h2) (image.rs:238). */
int ds_fast_guided_filter(ds_ctx* ctx, const float* I, const float* p, size_t w, size_t h, size_t r, float eps, size_t sub, float* out);
int ds_plane_resize(ds_ctx* ctx, const float* in, size_t w, size_t h, size_t w2, size_t h2, float* out);

/* The export post-steps after render_full (export.rs:244-247), in place on
 * interleaved ProPhoto.  ds_super_res_refine replaces
 * enhance::super_res_refine(img) (enhance.rs:122); ds_output_sharpen replaces
 * output_sharpen(img, mode, amount) (export.rs:199) with mode a SharpenFor
 * (export.rs:63) and amount a SharpenAmount (export.rs:71), numbered in
 * declaration order; other values are DS_ERR_INVALID. */
enum { DS_SHARPEN_FOR_NONE = 0, DS_SHARPEN_FOR_SCREEN = 1, DS_SHARPEN_FOR_MATTE = 2, DS_SHARPEN_FOR_GLOSSY = 3 };
enum { DS_SHARPEN_AMOUNT_LOW = 0, DS_SHARPEN_AMOUNT_STANDARD = 1, DS_SHARPEN_AMOUNT_HIGH = 2 };
int ds_super_res_refine(ds_ctx* ctx, float* rgb, size_t w, size_t h);
int ds_output_sharpen(ds_ctx* ctx, float* rgb, size_t w, size_t h, int mode, int amount);

/* ---- enhance ----------------------------------------------------------- */

/* enhance.rs:10 estimate_noise(p) */
int ds_estimate_noise(ds_ctx* ctx, const float* plane, size_t w, size_t h, float* sigma_out);

/* enhance.rs:36 nlm(p, h_param, search, patch) */
int ds_nlm(ds_ctx* ctx, const float* in, size_t w, size_t h, float h_param, long search, size_t patch, float* out);

/* enhance.rs:36 nlm(p, h_param, search, patch) in the reference's own
 * summation order for a rayon pool of `threads` threads
 * (rayon::current_num_threads(), enhance.rs:43-44): offsets in chunks of
 * ceil((2S+1)^2 / threads), each chunk's sums offset by offset, running
 * box sums along whole rows then whole columns (image.rs:280-309), glibc's
 * expf, chunk partials added in order.  Bit for bit the CPU result; it
 * walks every row and column once per offset, so it is a parity mode, far
 * slower than ds_nlm. */
int ds_nlm_reference_order(ds_ctx* ctx, const float* in, size_t w, size_t h, float h_param, long search, size_t patch, int threads,
                           float* out);

/* enhance.rs:75 denoise(img, amount): NLM with search 3, patch 1. */
int ds_denoise(ds_ctx* ctx, const float* rgb, size_t w, size_t h, float amount, float* rgb_out);

/* enhance.rs:75 denoise with the NLM search / patch radii as parameters
 * (the reference hard-codes 3, 1 at enhance.rs:84). */
int ds_denoise_ex(ds_ctx* ctx, const float* rgb, size_t w, size_t h, float amount, long search, size_t patch, float* rgb_out);

/* enhance.rs:75 denoise with its NLM in the reference's summation order for
 * a pool of `threads` threads (as ds_nlm_reference_order); every other stage
 * is the same as ds_denoise_ex, so the result is the CPU path's, bit for bit. */
int ds_denoise_reference_order(ds_ctx* ctx, const float* rgb, size_t w, size_t h, float amount, long search, size_t patch, int threads,
                               float* rgb_out);

/* enhance.rs:96 raw_details(img, sigma, iterations) */
int ds_raw_details(ds_ctx* ctx, const float* rgb, size_t w, size_t h, float sigma, size_t iterations, float* rgb_out);

/* ---- plane primitives of the path --------------------------------------- */

/* image.rs:269 box_blur(p, r): running-sum window mean, clamped edges. */
int ds_box_blur(ds_ctx* ctx, const float* in, size_t w, size_t h, size_t r, float* out);

/* image.rs:327 gaussian_blur(p, sigma): copy below 0.3, exact separable
 * (image.rs:341 gaussian_blur_exact) below 2, else three box blurs. */
int ds_gaussian_blur(ds_ctx* ctx, const float* in, size_t w, size_t h, float sigma, float* out);

/* image.rs:385 guided_filter(guide, src, r, eps) */
int ds_guided_filter(ds_ctx* ctx, const float* guide, const float* src, size_t w, size_t h, size_t r, float eps, float* out);

/* color.rs:408 Oklab::from_prophoto / color.rs:415 Oklab::to_prophoto on n
 * interleaved triples. */
int ds_oklab_from_prophoto(ds_ctx* ctx, const float* rgb, size_t n, float* lab_out);
int ds_oklab_to_prophoto(ds_ctx* ctx, const float* lab, size_t n, float* rgb_out);

/* image.rs:174 Image3::luminance (color.rs:196 luma_prophoto) of n pixels. */
int ds_luminance(ds_ctx* ctx, const float* rgb, size_t n, float* out);

/* Self-check of the kernels' libm restatements: the cube root behind
 * Oklab::from_prophoto (color.rs:408-412, f32::cbrt -> glibc cbrtf) and the
 * exponential (f32::exp -> glibc expf).  Evaluates the device function on
 * all 2^32 float inputs and returns sum_i mix(i << 32 | bits(f(x_i)))
 * mod 2^64 (SplitMix64 finaliser, NaN results as 0x7fc00000); the host
 * libm's outputs give the same hash when the two agree bit for bit. */
#define DS_LIBM_CBRTF 0
#define DS_LIBM_EXPF 1
int ds_libm_hash(ds_ctx* ctx, int fn, uint64_t* hash_out);

/* ---- device-resident variants (enqueue only) ---------------------------- */

int ds_dev_alloc(ds_ctx* ctx, size_t bytes, void** dptr);
int ds_dev_free(ds_ctx* ctx, void* dptr);
/* Pinned (page-locked) host memory for staging frames: copies from and to it
 * run at the link's rate, where pageable memory is copied through a bounce
 * buffer.  No counterpart in the reference (its buffers are Vec<f32>); freed
 * with the context if not before. */
int ds_host_alloc(ds_ctx* ctx, size_t bytes, void** hptr);
int ds_host_free(ds_ctx* ctx, void* hptr);
int ds_memcpy_h2d(ds_ctx* ctx, void* dst, const void* src, size_t bytes); /* synchronous */
int ds_memcpy_d2h(ds_ctx* ctx, void* dst, const void* src, size_t bytes); /* synchronous */
int ds_sync(ds_ctx* ctx);
/* Device-to-device copy with a plain streaming kernel (16-byte aligned). */
int ds_dev_copy(ds_ctx* ctx, void* dst, const void* src, size_t bytes);

int ds_demosaic_u16_dev(ds_ctx* ctx, const uint16_t* raw, size_t w, size_t h, size_t stride, const uint8_t pattern[4], float black,
                        const float white[3], const float mul[3], int algo, float* rgb_out, float* clip_out);
int ds_demosaic_f32_dev(ds_ctx* ctx, const float* cfa, size_t w, size_t h, const uint8_t pattern[4], const float mul[3], int algo,
                        float* rgb_out, float* clip_out);
/* raw->data, cfa_out, rgb_out, clip_out and levels_out on the device */
int ds_cfa_ingest_dev(ds_ctx* ctx, const ds_raw_cfa* raw, float* cfa_out, uint8_t pattern_out[4]);
int ds_demosaic_raw_dev(ds_ctx* ctx, const ds_raw_cfa* raw, const float mul[3], int algo, int quality, float* rgb_out, float* clip_out);
int ds_cfa_clip_levels_raw_dev(ds_ctx* ctx, const ds_raw_cfa* raw, float* levels_out);
int ds_superpixel_f32_dev(ds_ctx* ctx, const float* cfa, size_t w, size_t h, const uint8_t pattern[4], const float mul[3], float* rgb_out,
                          float* clip_out);
/* detail stages on device buffers (noise_field / sharpness_field too); the
 * noise reduction reads the field's maximum back to the host when a field is given. */
int ds_noise_reduction_dev(ds_ctx* ctx, float* rgb, size_t w, size_t h, const ds_nr_params* nr, const float* noise_field,
                           float scale);
int ds_sharpen_dev(ds_ctx* ctx, float* rgb, size_t w, size_t h, const ds_sharpen_params* sh, const float* sharpness_field, float scale);
int ds_fast_guided_filter_dev(ds_ctx* ctx, const float* I, const float* p, size_t w, size_t h, size_t r, float eps, size_t sub, float* out);
int ds_plane_resize_dev(ds_ctx* ctx, const float* in, size_t w, size_t h, size_t w2, size_t h2, float* out);
int ds_super_res_refine_dev(ds_ctx* ctx, float* rgb, size_t w, size_t h);
int ds_output_sharpen_dev(ds_ctx* ctx, float* rgb, size_t w, size_t h, int mode, int amount);
/* levels_out: 3 device floats */
int ds_cfa_clip_levels_f32_dev(ds_ctx* ctx, const float* cfa, size_t w, size_t h, const uint8_t pattern[4], float* levels_out);
int ds_cfa_clip_levels_u16_dev(ds_ctx* ctx, const uint16_t* raw, size_t w, size_t h, size_t stride, const uint8_t pattern[4],
                               float black, const float white[3], float* levels_out);
/* rgb, clip, raw_clip (3 floats) on the device; mul on the host.  With no
 * positive clip value nothing changes, decided on the device (no wait). */
int ds_reconstruct_highlights_dev(ds_ctx* ctx, float* rgb, const float* clip, size_t w, size_t h, const float mul[3],
                                  const float* raw_clip);
/* sigma_out is a device pointer to one float */
int ds_estimate_noise_dev(ds_ctx* ctx, const float* plane, size_t w, size_t h, float* sigma_out);
/* No counterpart in the reference: ds_estimate_noise_dev for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  It calls no allocating routine:
 * the select's scratch (histograms and samples, 3.2 MB) is a static device
 * array of the library, one per GPU, sized for every plane size; a plane
 * needing more would be refused with DS_ERR_INVALID.  Every call on a GPU
 * shares that array and leaves its contents there (the call clears what must
 * start at zero), so calls that overlap on one GPU (from two contexts, or
 * streams running at once) race; calls on one stream are ordered.  plane and
 * sigma_out are device pointers; enqueue only. */
int ds_estimate_noise_noalloc(ds_ctx* ctx, const float* plane, size_t w, size_t h, float* sigma_out);
int ds_nlm_dev(ds_ctx* ctx, const float* in, size_t w, size_t h, float h_param, long search, size_t patch, float* out);
int ds_denoise_ex_dev(ds_ctx* ctx, const float* rgb, size_t w, size_t h, float amount, long search, size_t patch, float* rgb_out);
int ds_nlm_reference_order_dev(ds_ctx* ctx, const float* in, size_t w, size_t h, float h_param, long search, size_t patch, int threads,
                               float* out);
int ds_denoise_reference_order_dev(ds_ctx* ctx, const float* rgb, size_t w, size_t h, float amount, long search, size_t patch,
                                   int threads, float* rgb_out);
int ds_raw_details_dev(ds_ctx* ctx, const float* rgb, size_t w, size_t h, float sigma, size_t iterations, float* rgb_out);

/* ---- stage timing ------------------------------------------------------ */

/* With profiling on, every kernel launch of the context is bracketed by
 * HIP events on the context's stream; per-kernel totals accumulate until
 * ds_prof_reset.  ds_prof_get fills up to `cap` entries and returns the
 * number of kernels seen (or a negative error). */
typedef struct ds_prof_entry {
    char name[48];
    double total_ms;
    long launches;
} ds_prof_entry;

int ds_prof_enable(ds_ctx* ctx, int on);
int ds_prof_reset(ds_ctx* ctx);
int ds_prof_get(ds_ctx* ctx, ds_prof_entry* out, int cap);

/* ---- synthetic input --------------------------------------------------- */

/* Seeded synthetic Bayer frame (gradients + polygons + sinusoids, linear in
 * [0.02, 0.85] of range, Poisson-Gaussian noise with gain 0.5*iso DN/e and
 * read sigma 4 DN).  bits 14: black 512, white 16383; bits 12: black 128,
 * white 4095.  iso 0 renders the same scene without noise.  raw_out: w*h
 * u16.  The _dev form writes device memory. */
int ds_synth_bayer_host(uint64_t seed, size_t w, size_t h, int bits, float iso, const uint8_t pattern[4], uint16_t* raw_out);
int ds_synth_bayer_dev(ds_ctx* ctx, uint64_t seed, size_t w, size_t h, int bits, float iso, const uint8_t pattern[4], uint16_t* raw_out);

/* ---- learned luminance denoiser ---------------------------------------- */

/* No counterpart in the reference: a residual CNN (8 conv3x3 layers, 32
 * channels, ReLU; BF16 weights and activations, FP32 accumulation) that
 * stands in for the NLM call of denoise (enhance.rs:84), trained on the
 * GPU on pairs (noisy luminance, NLM output) that the pipeline generates
 * from synthetic frames.  Its input is x = clamp((L - mean5x5(L)) / s, -64,
 * 64) with s = max(estimate_noise(L), 1e-4) (enhance.rs:82); its output y
 * gives the luminance L + s (y - x).  Parameters: ds_cnn_param_count()
 * FP32 values, laid out W_in[32][9], b_in[32], six times W[co 32][tap 9]
 * [ci 32], b[32], then W_out[tap 9][ci 32], b_out (tap = 3 (dy+1) + dx+1). */
typedef struct ds_cnn ds_cnn;

size_t ds_cnn_param_count(void);
/* He-normal initial weights from `seed`; Adam state zero. */
int ds_cnn_create(ds_ctx* ctx, uint64_t seed, ds_cnn** out);
void ds_cnn_destroy(ds_cnn* net);
int ds_cnn_get_params(ds_cnn* net, float* params, size_t n);
/* Replaces the weights and resets the Adam state and step count. */
int ds_cnn_set_params(ds_cnn* net, const float* params, size_t n);
long ds_cnn_step(const ds_cnn* net);
/* Training checkpoint: weights, Adam moments and step ("DSCNN001" file). */
int ds_cnn_save(ds_cnn* net, const char* path);
int ds_cnn_load(ds_cnn* net, const char* path);

/* Host batches of n_img images h x w (network input x, target t, output y).
 * ds_cnn_loss_grad: loss = mean over pixels at least `margin` from the
 * image border of (y - t)^2, and its gradient (ds_cnn_param_count floats). */
int ds_cnn_forward(ds_cnn* net, const float* x, size_t n_img, size_t h, size_t w, float* y_out);
/* Inference forward with all layers fused in one kernel (activations stay
 * in LDS); the same result as ds_cnn_forward, bit for bit. */
int ds_cnn_infer(ds_cnn* net, const float* x, size_t n_img, size_t h, size_t w, float* y_out);
int ds_cnn_loss_grad(ds_cnn* net, const float* x, const float* t, size_t n_img, size_t h, size_t w, size_t margin, double* loss,
                     float* grad);
/* One Adam step (lr, beta 0.9 / 0.999, eps 1e-8) on a device batch; with
 * loss_out set, waits for and returns the batch loss before the step. */
int ds_cnn_train_step_dev(ds_cnn* net, const float* x, const float* t, size_t n_img, size_t h, size_t w, size_t margin, float lr,
                          double* loss_out);
/* Training pairs: the synthetic frame `seed` (frame_w x frame_h RGGB, 14-bit,
 * noise level iso) through RCD demosaic, Oklab L, estimate_noise and NLM
 * (h = 0.6 sigma, search / patch radii); n_crops seeded crop x crop crops of
 * x and of the target t = x + (NLM(L) - L) / s, into device buffers. */
int ds_cnn_make_pairs_dev(ds_ctx* ctx, uint64_t seed, size_t frame_w, size_t frame_h, float iso, long search, size_t patch, size_t crop,
                          size_t n_crops, float* x_out, float* t_out);
/* The learned luminance of one plane (host buffers): L + s (y - x). */
int ds_cnn_denoise_plane(ds_cnn* net, const float* L, size_t w, size_t h, float* lf_out);

/* denoise (enhance.rs:74-92) with the learned luminance in place of NLM.
 * Every 256 x 256 region is spot-checked against NLM on a 16 x 16 lattice
 * of pixels: deviation = 10 log10(1 + mean (learned - NLM)^2 / mean var),
 * var = sigma^2 sum w^2 / (sum w)^2 being the white-noise variance NLM
 * leaves at that pixel.  Regions whose deviation exceeds threshold_db (all
 * regions when threshold_db <= 0) are recomputed with NLM.  stats may be
 * NULL; when set, the call waits for the result. */
typedef struct ds_learned_stats {
    size_t regions;
    size_t fallback;
    float mean_gap_db;
    float max_gap_db;
} ds_learned_stats;

int ds_denoise_learned(ds_ctx* ctx, ds_cnn* net, const float* rgb, size_t w, size_t h, float amount, long search, size_t patch,
                       float threshold_db, float* rgb_out, ds_learned_stats* stats);
int ds_denoise_learned_dev(ds_ctx* ctx, ds_cnn* net, const float* rgb, size_t w, size_t h, float amount, long search, size_t patch,
                           float threshold_db, float* rgb_out, ds_learned_stats* stats);

/* ---- batched export ----------------------------------------------------- */

/* Stands in for the sequential per-file loops of the CLI and GUI batch
 * export (cli.rs:142-187, dialogs.rs:310-365), each of which runs
 * Engine::base_for (pipeline/mod.rs:229-252) per image: demosaic ->
 * [denoise] -> [raw_details], with Engine::base's highlight reconstruction
 * (pipeline/mod.rs:186-187) after the demosaic when opts.highlights is set.  Every frame of the batch goes through one
 * hipGraph, captured on the first call and replayed while the frames
 * (pointers, shapes, formats) and options stay the same.  Frames are spread
 * over `streams` parallel graph branches (longest first), each with its own
 * workspace.  With autotune set, tile shapes are chosen per frame shape by
 * timing every variant once (ds_autotune), and with streams == 0 the branch
 * count too.  All pointers are device pointers; the call only enqueues,
 * unless info is set (then it waits and fills it).  With host_frames set the
 * pointers are pinned host buffers instead: the batch then runs without a
 * graph, as a pipeline of uploads, kernels and downloads on three streams
 * (largest frame first), and the call returns when every result is in host
 * memory; streams is ignored and info reports 1 branch. */
#define DS_DENOISE_NONE 0
#define DS_DENOISE_NLM 1
#define DS_DENOISE_LEARNED 2
#define DS_MAX_BRANCHES 8

typedef struct ds_frame {
    const uint16_t* raw; /* h rows of `stride` u16 */
    size_t w, h, stride;
    uint8_t pattern[4];
    float black;
    float white[3];
    float mul[3];
    float* rgb_out; /* w*h interleaved RGB (w/2 * h/2 at DS_QUALITY_DRAFT) */
    float black_site[4]; /* with per_site set: black level of each CFA site, (y & 1) * 2 + (x & 1), overriding black */
    int per_site;
} ds_frame;

typedef struct ds_pipeline_opts {
    int algo;             /* DS_ALGO_RCD / DS_ALGO_MHC */
    int denoise;          /* DS_DENOISE_* */
    float amount;         /* denoise amount (enhance.rs:75) */
    long search;          /* NLM search radius */
    size_t patch;         /* NLM patch radius */
    int raw_details;      /* run RL */
    float rl_sigma;
    size_t rl_iters;
    int reference_gating; /* 1: base_for's rules: amount.max(1), RL (0.75, 8) only for a long edge > 3200 */
    ds_cnn* net;          /* DS_DENOISE_LEARNED */
    float threshold_db;   /* learned: spot-check threshold */
    int streams;          /* graph branches, 1..DS_MAX_BRANCHES; 0 = 2, or tuned when autotune is set */
    int autotune;
    int highlights;       /* 1: reconstruct_highlights after demosaic, as Engine::base does (pipeline/mod.rs:186-187) */
    int quality;          /* DS_QUALITY_FULL (RCD / MHC) or DS_QUALITY_DRAFT (superpixel, half size; no RL under reference_gating) */
    int host_frames;      /* 1: every frame's raw and rgb_out are pinned host buffers (ds_host_alloc); each frame is uploaded
                             and its result downloaded on copy streams while other frames' kernels run */
} ds_pipeline_opts;

typedef struct ds_batch_info {
    size_t frames;
    size_t branches;
    int graph_reused;        /* 1: replayed a graph captured by an earlier call */
    double capture_ms;       /* host time spent capturing (0 when reused) */
    size_t fallback_regions; /* learned: regions recomputed with NLM */
    size_t regions;
} ds_batch_info;

int ds_pipeline_batch(ds_ctx* ctx, const ds_frame* frames, size_t n, const ds_pipeline_opts* opts, ds_batch_info* info);

/* Tile autotuner.  Variants of the RCD and MHC demosaic tiles and of the RL
 * iteration tile compute the same results; ds_autotune times each on a
 * synthetic frame of the shape and keeps the fastest for that shape (used
 * by every later call on the context).  *_ms: median time per variant. */
typedef struct ds_tuning {
    int rcd_variant, mhc_variant, rl_variant;
    float rcd_ms[4], mhc_ms[4], rl_ms[4];
} ds_tuning;

int ds_autotune(ds_ctx* ctx, size_t w, size_t h, ds_tuning* out);
int ds_set_tuning(ds_ctx* ctx, size_t w, size_t h, const ds_tuning* tuning);
int ds_get_tuning(ds_ctx* ctx, size_t w, size_t h, ds_tuning* out);

/* ---- per-image statistics (FP64 on the GPU) ------------------------------ */

/* The statistics of Engine::auto_wb_xy (pipeline/mod.rs:421-441) over an
 * RGB image: every third pixel, raw = p * as_shot_neutral, skipped when
 * max > 0.9 or min < 0.002; w = sqrt(max); acc[c] = sum ln(raw[c]) w,
 * n = sum w, accumulated in FP64; neutral = [exp(acc0/n)/g, 1, exp(acc2/n)/g]
 * with g = exp(acc1/n) (valid = 0 and neutral 1 when n == 0). */
typedef struct ds_wb_stats {
    double acc[3];
    double n;
    float neutral[3];
    int valid;
} ds_wb_stats;

int ds_auto_wb_stats(ds_ctx* ctx, const float* rgb, size_t w, size_t h, const float as_shot_neutral[3], ds_wb_stats* out);
int ds_auto_wb_stats_dev(ds_ctx* ctx, const float* rgb, size_t w, size_t h, const float as_shot_neutral[3], ds_wb_stats* out);

#ifdef __cplusplus
}
#endif

#endif /* DARKSLIDE_GPU_H */
