"""CPU oracle of the raw detail path (TEST INFRASTRUCTURE ONLY).

ctypes binding of oracle/liboracle.so, the C++ restatement of the reference
Rust (oracle.cpp).  Only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline use it, as the checker and as the timed CPU path; the product
library never loads it.

Every function takes float32 or float64 numpy arrays and dispatches to the
FP32 (bit-faithful) or FP64 instantiation accordingly.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
    return _lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _sfx(dtype):
    if dtype == np.float32:
        return "f32", ctypes.c_float
    if dtype == np.float64:
        return "f64", ctypes.c_double
    raise TypeError(dtype)


def _fn(name, dtype, restype=None):
    sfx, ct = _sfx(dtype)
    f = getattr(lib(), f"{name}_{sfx}")
    f.restype = restype
    return f, ct


def _pat(pattern):
    return (ctypes.c_uint8 * 4)(*[int(v) for v in pattern])


def threads():
    f = lib().or_max_threads
    f.restype = ctypes.c_int
    return f()


def cfa_normalize_u16(raw, pattern, black, white, dtype=np.float32):
    raw = np.ascontiguousarray(raw, np.uint16)
    h, w = raw.shape
    out = np.empty((h, w), dtype)
    f, ct = _fn("or_cfa_normalize_u16", dtype)
    wh = np.asarray(white, dtype)
    f(_p(raw), ctypes.c_size_t(w), ctypes.c_size_t(h), ctypes.c_size_t(w), _pat(pattern), ct(black), _p(wh), _p(out))
    return out


def _demo(name, cfa, pattern, mul):
    cfa = np.ascontiguousarray(cfa)
    h, w = cfa.shape
    out = np.empty((h, w, 3), cfa.dtype)
    f, _ = _fn(name, cfa.dtype)
    m = np.asarray(mul, cfa.dtype)
    f(_p(cfa), ctypes.c_size_t(w), ctypes.c_size_t(h), _pat(pattern), _p(m), _p(out))
    return out


def mhc_bayer(cfa, pattern=(0, 1, 1, 2), mul=(1, 1, 1)):
    return _demo("or_mhc_bayer", cfa, pattern, mul)


def rcd_bayer(cfa, pattern=(0, 1, 1, 2), mul=(1, 1, 1)):
    return _demo("or_rcd_bayer", cfa, pattern, mul)


def clip_plane(cfa):
    cfa = np.ascontiguousarray(cfa)
    h, w = cfa.shape
    out = np.empty((h, w), cfa.dtype)
    f, _ = _fn("or_clip_plane", cfa.dtype)
    f(_p(cfa), ctypes.c_size_t(w), ctypes.c_size_t(h), _p(out))
    return out


def demosaic(cfa, pattern=(0, 1, 1, 2), mul=(1, 1, 1)):
    return rcd_bayer(cfa, pattern, mul), clip_plane(cfa)


def _pix3(name, arr):
    arr = np.ascontiguousarray(arr)
    out = np.empty_like(arr)
    f, _ = _fn(name, arr.dtype)
    f(_p(arr), ctypes.c_size_t(arr.size // 3), _p(out))
    return out


def oklab_from_prophoto(rgb):
    return _pix3("or_oklab_from_prophoto", rgb)


def oklab_to_prophoto(lab):
    return _pix3("or_oklab_to_prophoto", lab)


def oklab_matrices(dtype=np.float32):
    mats = [np.empty((3, 3), dtype) for _ in range(4)]
    f, _ = _fn("or_oklab_matrices", dtype)
    f(*[_p(m) for m in mats])
    return dict(zip(["to_lms", "from_lms", "m2", "m2_inv"], mats))


def luminance(rgb):
    rgb = np.ascontiguousarray(rgb)
    out = np.empty(rgb.shape[:-1], rgb.dtype)
    f, _ = _fn("or_luminance", rgb.dtype)
    f(_p(rgb), ctypes.c_size_t(rgb.size // 3), _p(out))
    return out


def estimate_noise(p):
    p = np.ascontiguousarray(p)
    h, w = p.shape
    _, ct = _sfx(p.dtype)
    f, _ = _fn("or_estimate_noise", p.dtype, ct)
    return float(f(_p(p), ctypes.c_size_t(w), ctypes.c_size_t(h)))


def box_blur(p, r):
    p = np.ascontiguousarray(p)
    h, w = p.shape
    out = np.empty_like(p)
    f, _ = _fn("or_box_blur", p.dtype)
    f(_p(p), ctypes.c_size_t(w), ctypes.c_size_t(h), ctypes.c_size_t(r), _p(out))
    return out


def gaussian_blur(p, sigma):
    p = np.ascontiguousarray(p)
    h, w = p.shape
    out = np.empty_like(p)
    f, ct = _fn("or_gaussian_blur", p.dtype)
    f(_p(p), ctypes.c_size_t(w), ctypes.c_size_t(h), ct(sigma), _p(out))
    return out


def guided_filter(I, p, r, eps):
    I = np.ascontiguousarray(I)
    p = np.ascontiguousarray(p, I.dtype)
    h, w = I.shape
    out = np.empty_like(I)
    f, ct = _fn("or_guided_filter", I.dtype)
    f(_p(I), _p(p), ctypes.c_size_t(w), ctypes.c_size_t(h), ctypes.c_size_t(r), ct(eps), _p(out))
    return out


def nlm(p, h_param, search, patch, nthreads=None):
    p = np.ascontiguousarray(p)
    h, w = p.shape
    out = np.empty_like(p)
    f, ct = _fn("or_nlm", p.dtype)
    f(_p(p), ctypes.c_size_t(w), ctypes.c_size_t(h), ct(h_param), ctypes.c_long(search), ctypes.c_size_t(patch),
      ctypes.c_int(nthreads or threads()), _p(out))
    return out


def denoise(rgb, amount, search=3, patch=1, nthreads=None):
    rgb = np.ascontiguousarray(rgb)
    h, w, _ = rgb.shape
    out = np.empty_like(rgb)
    f, ct = _fn("or_denoise", rgb.dtype)
    f(_p(rgb), ctypes.c_size_t(w), ctypes.c_size_t(h), ct(amount), ctypes.c_long(search), ctypes.c_size_t(patch),
      ctypes.c_int(nthreads or threads()), _p(out))
    return out


def raw_details(rgb, sigma, iters):
    rgb = np.ascontiguousarray(rgb)
    h, w, _ = rgb.shape
    out = np.empty_like(rgb)
    f, ct = _fn("or_raw_details", rgb.dtype)
    f(_p(rgb), ctypes.c_size_t(w), ctypes.c_size_t(h), ct(sigma), ctypes.c_size_t(iters), _p(out))
    return out


def gaussian_taps(sigma):
    taps = np.zeros(32, np.float32)
    r = ctypes.c_int()
    f = lib().or_gaussian_taps_f32
    f(ctypes.c_float(sigma), _p(taps), ctypes.byref(r))
    return taps[: 2 * r.value + 1], r.value


def box_radii_for_gauss(sigma):
    radii = (ctypes.c_size_t * 3)()
    lib().or_box_radii_for_gauss(ctypes.c_float(sigma), radii)
    return list(radii)


# ---- learned denoiser (cnn_oracle.cpp): FP64 forward / loss / gradient


def cnn_param_count():
    f = lib().or_cnn_param_count
    f.restype = ctypes.c_size_t
    return int(f())


def bf16_round(a):
    """Round float64 values to the nearest bfloat16 (ties to even)."""
    f = lib().or_cnn_bf16_round
    f.restype = ctypes.c_double
    f.argtypes = [ctypes.c_double]
    return np.vectorize(lambda v: f(float(v)))(np.asarray(a, dtype=np.float64))


def cnn_forward(params, x, bf16=False):
    """x: (n, H, W) float32 -> out (n, H, W) float64."""
    params = np.ascontiguousarray(params, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float32)
    n, H, W = x.shape
    out = np.empty((n, H, W), np.float64)
    f = lib().or_cnn_forward_f64
    f.restype = None
    f(_p(params), _p(x), n, H, W, int(bool(bf16)), _p(out))
    return out


def cnn_loss_grad(params, x, tgt, margin, bf16=False, grad=True):
    """Mean squared error inside the margin and its gradient (float64)."""
    params = np.ascontiguousarray(params, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.float32)
    tgt = np.ascontiguousarray(tgt, dtype=np.float32)
    n, H, W = x.shape
    g = np.empty(cnn_param_count(), np.float64) if grad else None
    f = lib().or_cnn_loss_grad_f64
    f.restype = ctypes.c_double
    loss = f(_p(params), _p(x), _p(tgt), n, H, W, int(margin), int(bool(bf16)), _p(g))
    return (loss, g) if grad else loss


def cnn_adam(w, g, m, v, step, lr):
    """One Adam step in place on float64 arrays."""
    f = lib().or_cnn_adam_f64
    f.restype = None
    f(_p(w), _p(g), _p(m), _p(v), ctypes.c_size_t(w.size), int(step), ctypes.c_double(lr))


def cnn_input(L, sigma):
    L = np.ascontiguousarray(L, dtype=np.float32)
    h, w = L.shape
    x = np.empty((h, w), np.float64)
    f = lib().or_cnn_input_f64
    f.restype = None
    f(_p(L), w, h, ctypes.c_double(sigma), _p(x))
    return x


def auto_wb_stats(rgb, as_shot_neutral):
    """pipeline/mod.rs:421-441 statistics (sequential FP64) -> (acc[3], n, neutral or None)."""
    rgb = np.ascontiguousarray(rgb, dtype=np.float32)
    npx = rgb.size // 3
    acc = np.zeros(3, np.float64)
    n = ctypes.c_double()
    neutral = np.zeros(3, np.float32)
    nas = np.ascontiguousarray(as_shot_neutral, dtype=np.float32)
    f = lib().or_auto_wb_stats
    f.restype = ctypes.c_int
    ok = f(_p(rgb), ctypes.c_size_t(npx), _p(nas), _p(acc), ctypes.byref(n), _p(neutral))
    return acc, n.value, (neutral if ok else None)


def cfa_clip_levels(cfa, pattern=(0, 1, 1, 2)):
    """demosaic.rs:419-430 -> 3 clip levels (f32::MAX where untrusted)."""
    cfa = np.ascontiguousarray(cfa)
    h, w = cfa.shape
    out = np.empty(3, cfa.dtype)
    f, _ = _fn("or_cfa_clip_levels", cfa.dtype)
    f(_p(cfa), ctypes.c_size_t(w), ctypes.c_size_t(h), _pat(pattern), _p(out))
    return out


def reconstruct_highlights(img, clip, mul, raw_clip):
    """demosaic.rs:435-502 on a copy of img (interleaved RGB)."""
    img = np.array(img, copy=True, order="C")
    clip = np.ascontiguousarray(clip, img.dtype)
    h, w, _ = img.shape
    f, _ = _fn("or_reconstruct_highlights", img.dtype)
    m = np.asarray(mul, img.dtype)
    rc = np.asarray(raw_clip, img.dtype)
    f(_p(img), _p(clip), ctypes.c_size_t(w), ctypes.c_size_t(h), _p(m), _p(rc))
    return img


def cfa_ingest_u16(raw, crop, pattern, black, white, black_dims=None, dtype=np.float32):
    """decode.rs:276-307 CFA branch: raw (H, W) uint16 buffer, crop = (x, y, w, h), pattern at the
    buffer origin, black = values (1, or bw*bh*bcpp with black_dims = (bw, bh, bcpp)), white = 1 or
    >= 3 values -> (cfa (h, w), pattern of the crop)."""
    raw = np.ascontiguousarray(raw, np.uint16)
    H, W = raw.shape
    cx, cy, cw, ch = crop
    assert cx + cw <= W and cy + ch <= H
    b = np.ascontiguousarray(black, dtype).ravel()
    wv = np.ascontiguousarray(white, dtype).ravel()
    bw, bh, bcpp = black_dims if black_dims else (1, 1, 1)
    out = np.empty((ch, cw), dtype)
    pat_out = (ctypes.c_uint8 * 4)()
    f, _ = _fn("or_cfa_ingest_u16", dtype)
    S = ctypes.c_size_t
    f(_p(raw), S(W), S(cx), S(cy), S(cw), S(ch), _pat(pattern), _p(b), S(b.size), S(bw), S(bh), S(bcpp), _p(wv), S(wv.size),
      _p(out), pat_out)
    return out, tuple(pat_out)


def superpixel(cfa, pattern=(0, 1, 1, 2), mul=(1, 1, 1)):
    """demosaic.rs:44-87 (Bayer) -> (rgb (h/2, w/2, 3), clip (h/2, w/2))."""
    cfa = np.ascontiguousarray(cfa)
    h, w = cfa.shape
    rgb = np.empty((h // 2, w // 2, 3), cfa.dtype)
    clip = np.empty((h // 2, w // 2), cfa.dtype)
    f, _ = _fn("or_superpixel", cfa.dtype)
    m = np.asarray(mul, cfa.dtype)
    f(_p(cfa), ctypes.c_size_t(w), ctypes.c_size_t(h), _pat(pattern), _p(m), _p(rgb), _p(clip))
    return rgb, clip


def plane_resize(p, w2, h2):
    """image.rs:238-262 Plane::resize."""
    p = np.ascontiguousarray(p)
    h, w = p.shape
    out = np.empty((h2, w2), p.dtype)
    f, _ = _fn("or_plane_resize", p.dtype)
    f(_p(p), ctypes.c_size_t(w), ctypes.c_size_t(h), ctypes.c_size_t(w2), ctypes.c_size_t(h2), _p(out))
    return out


def fast_guided_filter(I, p, r, eps, sub):
    """image.rs:408-431."""
    I = np.ascontiguousarray(I)
    p = np.ascontiguousarray(p, I.dtype)
    h, w = I.shape
    out = np.empty_like(I)
    f, ct = _fn("or_fast_guided_filter", I.dtype)
    f(_p(I), _p(p), ctypes.c_size_t(w), ctypes.c_size_t(h), ctypes.c_size_t(r), ct(eps), ctypes.c_size_t(sub), _p(out))
    return out


def noise_reduction(img, nr, field=None, scale=1.0):
    """pipeline/detail.rs:13-92 on a copy of img; nr = (luminance, luminance_detail, luminance_contrast,
    color, color_detail, color_smoothness) in the reference's 0..100 units."""
    img = np.array(img, copy=True, order="C")
    h, w, _ = img.shape
    f, ct = _fn("or_noise_reduction", img.dtype)
    prm = np.asarray(nr, img.dtype)
    fld = None if field is None else np.ascontiguousarray(field, img.dtype)
    f(_p(img), ctypes.c_size_t(w), ctypes.c_size_t(h), _p(prm), _p(fld) if fld is not None else None, ct(scale))
    return img


def sharpen(img, sh, field=None, scale=1.0):
    """pipeline/detail.rs:96-139 on a copy of img; sh = (amount, radius, detail, masking)."""
    img = np.array(img, copy=True, order="C")
    h, w, _ = img.shape
    f, ct = _fn("or_sharpen", img.dtype)
    prm = np.asarray(sh, img.dtype)
    fld = None if field is None else np.ascontiguousarray(field, img.dtype)
    f(_p(img), ctypes.c_size_t(w), ctypes.c_size_t(h), _p(prm), _p(fld) if fld is not None else None, ct(scale))
    return img


def super_res_refine(img):
    """enhance.rs:122-131 on a copy of img."""
    img = np.array(img, copy=True, order="C")
    h, w, _ = img.shape
    f, _ct = _fn("or_super_res_refine", img.dtype)
    f(_p(img), ctypes.c_size_t(w), ctypes.c_size_t(h))
    return img


def output_sharpen(img, mode, amount):
    """export.rs:199-223 on a copy of img; mode 0..3 = None, Screen, Matte,
    Glossy; amount 0..2 = Low, Standard, High."""
    img = np.array(img, copy=True, order="C")
    h, w, _ = img.shape
    f, _ct = _fn("or_output_sharpen", img.dtype)
    f(_p(img), ctypes.c_size_t(w), ctypes.c_size_t(h), ctypes.c_int(mode), ctypes.c_int(amount))
    return img


def tone_enc(v, dtype=np.float32):
    f, ct = _fn("or_tone_enc", dtype, None)
    f.restype = ct
    return float(f(ct(v)))


def tone_dec(v, dtype=np.float32):
    f, ct = _fn("or_tone_dec", dtype, None)
    f.restype = ct
    return float(f(ct(v)))
