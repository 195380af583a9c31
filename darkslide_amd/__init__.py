"""Python binding of the MI355X raw detail path (libdarkslide_gpu.so).

The functions mirror the reference engine's stage functions
(Syha-01/darkslide crates/core/src): ``demosaic`` (demosaic.rs:20),
``estimate_noise`` / ``nlm`` / ``denoise`` / ``raw_details``
(enhance.rs:10, :36, :75, :96).  Arrays follow the reference layouts:
a Plane is an ``(h, w)`` float32 array, an Image3 an ``(h, w, 3)`` float32
array, a CFA an ``(h, w)`` float32 (normalised) or uint16 (raw DN) array.

Everything runs on the GPU through the C ABI in include/darkslide_gpu.h.
There is no CPU fallback: without the built library or a GPU the calls raise.
"""
import atexit
import ctypes
import os
import threading
import weakref

import numpy as np

__all__ = [
    "DarkslideError", "Context", "lib", "default_context", "RGGB", "ALGO_RCD", "ALGO_MHC",
    "demosaic", "demosaic_u16", "estimate_noise", "nlm", "denoise", "raw_details", "synth_bayer", "CNN", "LearnedStats", "Frame", "PipelineOpts", "DENOISE_NONE", "DENOISE_NLM", "DENOISE_LEARNED",
]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdarkslide_gpu.so")

ALGO_RCD = 0
ALGO_MHC = 1
RGGB = (0, 1, 1, 2)

_lib = None
_lib_lock = threading.Lock()


class DarkslideError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"darkslide error {code}: {msg}")
        self.code = code


class ProfEntry(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 48), ("total_ms", ctypes.c_double), ("launches", ctypes.c_long)]


DENOISE_NONE, DENOISE_NLM, DENOISE_LEARNED = 0, 1, 2
QUALITY_FULL, QUALITY_DRAFT = 0, 1


class NrParams(ctypes.Structure):
    """ds_nr_params: NoiseReduction (params.rs:319-326), 0..100 units."""
    _fields_ = [("luminance", ctypes.c_float), ("luminance_detail", ctypes.c_float), ("luminance_contrast", ctypes.c_float),
                ("color", ctypes.c_float), ("color_detail", ctypes.c_float), ("color_smoothness", ctypes.c_float)]


class SharpenParams(ctypes.Structure):
    """ds_sharpen_params: Sharpening (params.rs:304-309)."""
    _fields_ = [("amount", ctypes.c_float), ("radius", ctypes.c_float), ("detail", ctypes.c_float), ("masking", ctypes.c_float)]


# SharpenFor (export.rs:63-68) and SharpenAmount (export.rs:71-75)
SHARPEN_FOR_NONE, SHARPEN_FOR_SCREEN, SHARPEN_FOR_MATTE, SHARPEN_FOR_GLOSSY = 0, 1, 2, 3
SHARPEN_AMOUNT_LOW, SHARPEN_AMOUNT_STANDARD, SHARPEN_AMOUNT_HIGH = 0, 1, 2


class RawCfa(ctypes.Structure):
    """ds_raw_cfa: a Bayer raw buffer with crop, black level table and white levels."""
    _fields_ = [("data", ctypes.c_void_p), ("width", ctypes.c_size_t), ("height", ctypes.c_size_t), ("crop_x", ctypes.c_size_t),
                ("crop_y", ctypes.c_size_t), ("crop_w", ctypes.c_size_t), ("crop_h", ctypes.c_size_t), ("pattern", ctypes.c_uint8 * 4),
                ("black", ctypes.c_void_p), ("nblack", ctypes.c_size_t), ("black_w", ctypes.c_size_t), ("black_h", ctypes.c_size_t),
                ("black_cpp", ctypes.c_size_t), ("white", ctypes.c_float * 3)]


class Frame(ctypes.Structure):
    """ds_frame: one raw frame of a batch (device pointers)."""
    _fields_ = [("raw", ctypes.c_void_p), ("w", ctypes.c_size_t), ("h", ctypes.c_size_t), ("stride", ctypes.c_size_t),
                ("pattern", ctypes.c_uint8 * 4), ("black", ctypes.c_float), ("white", ctypes.c_float * 3),
                ("mul", ctypes.c_float * 3), ("rgb_out", ctypes.c_void_p),
                ("black_site", ctypes.c_float * 4), ("per_site", ctypes.c_int)]


class PipelineOpts(ctypes.Structure):
    _fields_ = [("algo", ctypes.c_int), ("denoise", ctypes.c_int), ("amount", ctypes.c_float), ("search", ctypes.c_long),
                ("patch", ctypes.c_size_t), ("raw_details", ctypes.c_int), ("rl_sigma", ctypes.c_float), ("rl_iters", ctypes.c_size_t),
                ("reference_gating", ctypes.c_int), ("net", ctypes.c_void_p), ("threshold_db", ctypes.c_float),
                ("streams", ctypes.c_int), ("autotune", ctypes.c_int), ("highlights", ctypes.c_int),
                ("quality", ctypes.c_int), ("host_frames", ctypes.c_int)]


class BatchInfo(ctypes.Structure):
    _fields_ = [("frames", ctypes.c_size_t), ("branches", ctypes.c_size_t), ("graph_reused", ctypes.c_int),
                ("capture_ms", ctypes.c_double), ("fallback_regions", ctypes.c_size_t), ("regions", ctypes.c_size_t)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Tuning(ctypes.Structure):
    _fields_ = [("rcd_variant", ctypes.c_int), ("mhc_variant", ctypes.c_int), ("rl_variant", ctypes.c_int),
                ("rcd_ms", ctypes.c_float * 4), ("mhc_ms", ctypes.c_float * 4), ("rl_ms", ctypes.c_float * 4)]

    def as_dict(self):
        return {"rcd_variant": self.rcd_variant, "mhc_variant": self.mhc_variant, "rl_variant": self.rl_variant,
                "rcd_ms": list(self.rcd_ms)[:3], "mhc_ms": list(self.mhc_ms)[:3], "rl_ms": list(self.rl_ms)[:3]}


class WbStats(ctypes.Structure):
    _fields_ = [("acc", ctypes.c_double * 3), ("n", ctypes.c_double), ("neutral", ctypes.c_float * 3), ("valid", ctypes.c_int)]


class LearnedStats(ctypes.Structure):
    _fields_ = [("regions", ctypes.c_size_t), ("fallback", ctypes.c_size_t), ("mean_gap_db", ctypes.c_float),
                ("max_gap_db", ctypes.c_float)]

    def as_dict(self):
        return {"regions": self.regions, "fallback": self.fallback, "mean_gap_db": self.mean_gap_db, "max_gap_db": self.max_gap_db}


_P = ctypes.c_void_p
_SZ = ctypes.c_size_t
_F = ctypes.c_float
_I = ctypes.c_int
_L = ctypes.c_long

# name -> argtypes (all return int status unless listed in _RESTYPES)
_SIGS = {
    "ds_ctx_create": [_I, ctypes.POINTER(_P)],
    "ds_ctx_destroy": [_P],
    "ds_last_error": [_P],
    "ds_version": [],
    "ds_demosaic_f32": [_P, _P, _SZ, _SZ, _P, _P, _I, _P, _P],
    "ds_cfa_clip_levels_f32": [_P, _P, _SZ, _SZ, _P, _P],
    "ds_cfa_ingest": [_P, _P, _P, _P],
    "ds_noise_reduction": [_P, _P, _SZ, _SZ, _P, _P, _F],
    "ds_noise_reduction_dev": [_P, _P, _SZ, _SZ, _P, _P, _F],
    "ds_sharpen": [_P, _P, _SZ, _SZ, _P, _P, _F],
    "ds_sharpen_dev": [_P, _P, _SZ, _SZ, _P, _P, _F],
    "ds_fast_guided_filter": [_P, _P, _P, _SZ, _SZ, _SZ, _F, _SZ, _P],
    "ds_fast_guided_filter_dev": [_P, _P, _P, _SZ, _SZ, _SZ, _F, _SZ, _P],
    "ds_plane_resize": [_P, _P, _SZ, _SZ, _SZ, _SZ, _P],
    "ds_plane_resize_dev": [_P, _P, _SZ, _SZ, _SZ, _SZ, _P],
    "ds_super_res_refine": [_P, _P, _SZ, _SZ],
    "ds_super_res_refine_dev": [_P, _P, _SZ, _SZ],
    "ds_output_sharpen": [_P, _P, _SZ, _SZ, _I, _I],
    "ds_output_sharpen_dev": [_P, _P, _SZ, _SZ, _I, _I],
    "ds_cfa_ingest_dev": [_P, _P, _P, _P],
    "ds_demosaic_raw": [_P, _P, _P, _I, _I, _P, _P],
    "ds_demosaic_raw_dev": [_P, _P, _P, _I, _I, _P, _P],
    "ds_cfa_clip_levels_raw": [_P, _P, _P],
    "ds_cfa_clip_levels_raw_dev": [_P, _P, _P],
    "ds_superpixel_f32": [_P, _P, _SZ, _SZ, _P, _P, _P, _P],
    "ds_superpixel_f32_dev": [_P, _P, _SZ, _SZ, _P, _P, _P, _P],
    "ds_cfa_clip_levels_u16": [_P, _P, _SZ, _SZ, _SZ, _P, _F, _P, _P],
    "ds_reconstruct_highlights": [_P, _P, _P, _SZ, _SZ, _P, _P],
    "ds_cfa_clip_levels_f32_dev": [_P, _P, _SZ, _SZ, _P, _P],
    "ds_cfa_clip_levels_u16_dev": [_P, _P, _SZ, _SZ, _SZ, _P, _F, _P, _P],
    "ds_reconstruct_highlights_dev": [_P, _P, _P, _SZ, _SZ, _P, _P],
    "ds_demosaic_u16": [_P, _P, _SZ, _SZ, _SZ, _P, _F, _P, _P, _I, _P, _P],
    "ds_estimate_noise": [_P, _P, _SZ, _SZ, _P],
    "ds_nlm": [_P, _P, _SZ, _SZ, _F, _L, _SZ, _P],
    "ds_denoise": [_P, _P, _SZ, _SZ, _F, _P],
    "ds_denoise_ex": [_P, _P, _SZ, _SZ, _F, _L, _SZ, _P],
    "ds_nlm_reference_order": [_P, _P, _SZ, _SZ, _F, _L, _SZ, _I, _P],
    "ds_denoise_reference_order": [_P, _P, _SZ, _SZ, _F, _L, _SZ, _I, _P],
    "ds_nlm_reference_order_dev": [_P, _P, _SZ, _SZ, _F, _L, _SZ, _I, _P],
    "ds_denoise_reference_order_dev": [_P, _P, _SZ, _SZ, _F, _L, _SZ, _I, _P],
    "ds_raw_details": [_P, _P, _SZ, _SZ, _F, _SZ, _P],
    "ds_box_blur": [_P, _P, _SZ, _SZ, _SZ, _P],
    "ds_gaussian_blur": [_P, _P, _SZ, _SZ, _F, _P],
    "ds_guided_filter": [_P, _P, _P, _SZ, _SZ, _SZ, _F, _P],
    "ds_oklab_from_prophoto": [_P, _P, _SZ, _P],
    "ds_oklab_to_prophoto": [_P, _P, _SZ, _P],
    "ds_luminance": [_P, _P, _SZ, _P],
    "ds_libm_hash": [_P, _I, ctypes.POINTER(ctypes.c_uint64)],
    "ds_dev_alloc": [_P, _SZ, ctypes.POINTER(_P)],
    "ds_dev_free": [_P, _P],
    "ds_host_alloc": [_P, _SZ, ctypes.POINTER(_P)],
    "ds_host_free": [_P, _P],
    "ds_memcpy_h2d": [_P, _P, _P, _SZ],
    "ds_memcpy_d2h": [_P, _P, _P, _SZ],
    "ds_sync": [_P],
    "ds_dev_copy": [_P, _P, _P, _SZ],
    "ds_demosaic_u16_dev": [_P, _P, _SZ, _SZ, _SZ, _P, _F, _P, _P, _I, _P, _P],
    "ds_demosaic_f32_dev": [_P, _P, _SZ, _SZ, _P, _P, _I, _P, _P],
    "ds_estimate_noise_dev": [_P, _P, _SZ, _SZ, _P],
    "ds_estimate_noise_noalloc": [_P, _P, _SZ, _SZ, _P],
    "ds_nlm_dev": [_P, _P, _SZ, _SZ, _F, _L, _SZ, _P],
    "ds_denoise_ex_dev": [_P, _P, _SZ, _SZ, _F, _L, _SZ, _P],
    "ds_raw_details_dev": [_P, _P, _SZ, _SZ, _F, _SZ, _P],
    "ds_prof_enable": [_P, _I],
    "ds_prof_reset": [_P],
    "ds_prof_get": [_P, _P, _I],
    "ds_synth_bayer_host": [ctypes.c_uint64, _SZ, _SZ, _I, _F, _P, _P],
    "ds_synth_bayer_dev": [_P, ctypes.c_uint64, _SZ, _SZ, _I, _F, _P, _P],
    "ds_cnn_param_count": [],
    "ds_cnn_create": [_P, ctypes.c_uint64, ctypes.POINTER(_P)],
    "ds_cnn_destroy": [_P],
    "ds_cnn_get_params": [_P, _P, _SZ],
    "ds_cnn_set_params": [_P, _P, _SZ],
    "ds_cnn_step": [_P],
    "ds_cnn_save": [_P, ctypes.c_char_p],
    "ds_cnn_load": [_P, ctypes.c_char_p],
    "ds_cnn_forward": [_P, _P, _SZ, _SZ, _SZ, _P],
    "ds_cnn_infer": [_P, _P, _SZ, _SZ, _SZ, _P],
    "ds_cnn_loss_grad": [_P, _P, _P, _SZ, _SZ, _SZ, _SZ, ctypes.POINTER(ctypes.c_double), _P],
    "ds_cnn_train_step_dev": [_P, _P, _P, _SZ, _SZ, _SZ, _SZ, _F, ctypes.POINTER(ctypes.c_double)],
    "ds_cnn_make_pairs_dev": [_P, ctypes.c_uint64, _SZ, _SZ, _F, _L, _SZ, _SZ, _SZ, _P, _P],
    "ds_cnn_denoise_plane": [_P, _P, _SZ, _SZ, _P],
    "ds_denoise_learned": [_P, _P, _P, _SZ, _SZ, _F, _L, _SZ, _F, _P, _P],
    "ds_denoise_learned_dev": [_P, _P, _P, _SZ, _SZ, _F, _L, _SZ, _F, _P, _P],
    "ds_pipeline_batch": [_P, _P, _SZ, _P, _P],
    "ds_autotune": [_P, _SZ, _SZ, _P],
    "ds_set_tuning": [_P, _SZ, _SZ, _P],
    "ds_get_tuning": [_P, _SZ, _SZ, _P],
    "ds_auto_wb_stats": [_P, _P, _SZ, _SZ, _P, _P],
    "ds_auto_wb_stats_dev": [_P, _P, _SZ, _SZ, _P, _P],
}
_RESTYPES = {"ds_ctx_destroy": None, "ds_last_error": ctypes.c_char_p, "ds_version": ctypes.c_char_p,
             "ds_cnn_param_count": ctypes.c_size_t, "ds_cnn_destroy": None, "ds_cnn_step": ctypes.c_long}


def lib():
    """Loads libdarkslide_gpu.so (built by darkslide_amd/build.py)."""
    global _lib
    with _lib_lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise DarkslideError(-2, f"{LIB_PATH} is not built; run darkslide_amd/build.py")
            L = ctypes.CDLL(LIB_PATH)
            for name, args in _SIGS.items():
                f = getattr(L, name)
                f.argtypes = args
                f.restype = _RESTYPES.get(name, _I)
            _lib = L
    return _lib


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


def _plane(a, name, hw=None, dtype=np.float32):
    """a as a C-contiguous (h, w) array; with hw, its shape must be hw."""
    a = np.ascontiguousarray(a, dtype)
    if a.ndim != 2:
        raise ValueError(f"{name} must be an (h, w) array, got shape {a.shape}")
    if hw is not None and a.shape != tuple(hw):
        raise ValueError(f"{name} must have shape {tuple(hw)}, got {a.shape}")
    return a


def _rgb(a, name, hw=None, copy=False):
    """a as a C-contiguous (h, w, 3) float32 array (a fresh copy with copy=True)."""
    a = np.array(a, np.float32, order="C", copy=True) if copy else np.ascontiguousarray(a, np.float32)
    if a.ndim != 3 or a.shape[2] != 3:
        raise ValueError(f"{name} must be an (h, w, 3) array, got shape {a.shape}")
    if hw is not None and a.shape[:2] != tuple(hw):
        raise ValueError(f"{name} must have shape {tuple(hw) + (3,)}, got {a.shape}")
    return a


def _triples(a, name):
    """a as a C-contiguous float32 array of interleaved triples (last axis 3)."""
    a = np.ascontiguousarray(a, np.float32)
    if a.ndim < 1 or a.shape[-1] != 3:
        raise ValueError(f"{name} must have a last axis of 3, got shape {a.shape}")
    return a


def _batch(a, name, shape=None):
    """a as a C-contiguous (n, h, w) float32 batch; with shape, its shape must be shape."""
    a = np.ascontiguousarray(a, np.float32)
    if a.ndim != 3:
        raise ValueError(f"{name} must be an (n, h, w) array, got shape {a.shape}")
    if shape is not None and a.shape != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}, got {a.shape}")
    return a


def _pat(pattern):
    return (ctypes.c_uint8 * 4)(*[int(v) for v in pattern])


def _f3(v):
    return (ctypes.c_float * 3)(*[float(x) for x in v])


class HostBuffer:
    """Pinned host allocation owned by a Context (ds_host_alloc; freed with it or by .free())."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        ctx._check(lib().ds_host_alloc(ctx.handle, self.nbytes, ctypes.byref(p)))
        self.ptr = p

    def array(self, shape, dtype):
        """A numpy view of the buffer (valid until .free())."""
        dt = np.dtype(dtype)
        count = int(np.prod(shape))
        if count * dt.itemsize > self.nbytes:
            raise ValueError("view larger than buffer")
        raw = (ctypes.c_uint8 * (count * dt.itemsize)).from_address(self.ptr.value)
        return np.frombuffer(raw, dtype=dt, count=count).reshape(shape)

    def free(self):
        if self.ptr is not None:
            self.ctx._check(lib().ds_host_free(self.ctx.handle, self.ptr))
            self.ptr = None


class DeviceBuffer:
    """Device allocation owned by a Context (freed with it or by .free())."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        ctx._check(lib().ds_dev_alloc(ctx.handle, self.nbytes, ctypes.byref(p)))
        self.ptr = p

    def at(self, offset):
        """Device address `offset` bytes into the buffer."""
        if not 0 <= offset <= self.nbytes:
            raise ValueError("offset outside the buffer")
        return ctypes.c_void_p(self.ptr.value + int(offset))

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes:
            raise ValueError("array larger than buffer")
        self.ctx._check(lib().ds_memcpy_h2d(self.ctx.handle, self.ptr, _ptr(arr), arr.nbytes))

    def download(self, shape, dtype):
        out = np.empty(shape, dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("requested more than buffer holds")
        self.ctx._check(lib().ds_memcpy_d2h(self.ctx.handle, _ptr(out), self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr is not None:
            self.ctx._check(lib().ds_dev_free(self.ctx.handle, self.ptr))
            self.ptr = None


# Contexts still open when the interpreter exits are closed by an atexit
# handler, which runs before the HIP runtime's own teardown; left to
# finalisation, their destructors would call into a runtime already being
# torn down.
_live_contexts = weakref.WeakSet()


def _close_live_contexts():
    for c in list(_live_contexts):
        try:
            c.close()
        except Exception:
            pass


atexit.register(_close_live_contexts)


class Context:
    """One GPU context (device + HIP stream) of the C ABI."""

    def __init__(self, device=0):
        L = lib()
        h = ctypes.c_void_p()
        rc = L.ds_ctx_create(int(device), ctypes.byref(h))
        if rc != 0:
            raise DarkslideError(rc, L.ds_last_error(None).decode())
        self.handle = h
        self._nets = weakref.WeakSet()  # networks are destroyed before their context
        _live_contexts.add(self)

    def close(self):
        if self.handle:
            for net in list(self._nets):
                net.close()
            lib().ds_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise DarkslideError(rc, lib().ds_last_error(self.handle).decode())
        return rc

    # ---- device memory helpers
    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def host_alloc(self, nbytes):
        return HostBuffer(self, nbytes)

    def sync(self):
        self._check(lib().ds_sync(self.handle))

    # ---- profiling
    def prof_enable(self, on=True):
        self._check(lib().ds_prof_enable(self.handle, 1 if on else 0))

    def prof_reset(self):
        self._check(lib().ds_prof_reset(self.handle))

    def prof(self):
        n = lib().ds_prof_get(self.handle, None, 0)
        if n < 0:
            self._check(n)
        arr = (ProfEntry * max(n, 1))()
        lib().ds_prof_get(self.handle, ctypes.cast(arr, ctypes.c_void_p), n)
        return {arr[i].name.decode(): (arr[i].total_ms, arr[i].launches) for i in range(n)}

    # ---- stage functions (host arrays in, host arrays out)
    def demosaic(self, cfa, pattern=RGGB, mul=(1.0, 1.0, 1.0), algo=ALGO_RCD, clip=True):
        cfa = _plane(cfa, "cfa")
        h, w = cfa.shape
        rgb = np.empty((h, w, 3), np.float32)
        cl = np.empty((h, w), np.float32) if clip else None
        self._check(lib().ds_demosaic_f32(self.handle, _ptr(cfa), w, h, _pat(pattern), _f3(mul), algo, _ptr(rgb), _ptr(cl)))
        return (rgb, cl) if clip else rgb

    def demosaic_u16(self, raw, pattern=RGGB, black=512.0, white=(16383.0,) * 3, mul=(1.0, 1.0, 1.0), algo=ALGO_RCD, clip=True):
        raw = _plane(raw, "raw", dtype=np.uint16)
        h, w = raw.shape
        rgb = np.empty((h, w, 3), np.float32)
        cl = np.empty((h, w), np.float32) if clip else None
        self._check(lib().ds_demosaic_u16(self.handle, _ptr(raw), w, h, w, _pat(pattern), float(black), _f3(white), _f3(mul), algo,
                                          _ptr(rgb), _ptr(cl)))
        return (rgb, cl) if clip else rgb

    @staticmethod
    def _raw_cfa(raw, crop, pattern, black, white, black_dims):
        raw = _plane(raw, "raw", dtype=np.uint16)
        H, W = raw.shape
        cx, cy, cw, ch = crop if crop is not None else (0, 0, W, H)
        b = np.ascontiguousarray(np.asarray(black, np.float32).ravel())
        bw, bh, bcpp = black_dims if black_dims else (1, 1, 1)
        wv = list(np.asarray(white, np.float32).ravel())
        wv = wv * 3 if len(wv) == 1 else wv[:3]
        r = RawCfa(raw.ctypes.data, W, H, cx, cy, cw, ch, _pat(pattern), b.ctypes.data, b.size, bw, bh, bcpp, _f3(wv))
        return r, (raw, b), (cw, ch)

    def cfa_ingest(self, raw, crop=None, pattern=RGGB, black=(512.0,), white=(16383.0,), black_dims=None):
        """decode.rs:276-307 CFA branch -> (normalised crop (h, w) float32, pattern of the crop)."""
        r, keep, (cw, ch) = self._raw_cfa(raw, crop, pattern, black, white, black_dims)
        out = np.empty((ch, cw), np.float32)
        pat = (ctypes.c_uint8 * 4)()
        self._check(lib().ds_cfa_ingest(self.handle, ctypes.byref(r), _ptr(out), pat))
        return out, tuple(pat)

    def demosaic_raw(self, raw, crop=None, pattern=RGGB, black=(512.0,), white=(16383.0,), mul=(1.0, 1.0, 1.0), algo=ALGO_RCD,
                     quality=QUALITY_FULL, black_dims=None):
        """demosaic(cfa, mul, quality) of the normalised crop -> (rgb, clip)."""
        r, keep, (cw, ch) = self._raw_cfa(raw, crop, pattern, black, white, black_dims)
        ow, oh = (cw // 2, ch // 2) if quality == QUALITY_DRAFT else (cw, ch)
        rgb = np.empty((oh, ow, 3), np.float32)
        cl = np.empty((oh, ow), np.float32)
        self._check(lib().ds_demosaic_raw(self.handle, ctypes.byref(r), _f3(mul), algo, quality, _ptr(rgb), _ptr(cl)))
        return rgb, cl

    def cfa_clip_levels_raw(self, raw, crop=None, pattern=RGGB, black=(512.0,), white=(16383.0,), black_dims=None):
        r, keep, _ = self._raw_cfa(raw, crop, pattern, black, white, black_dims)
        out = np.zeros(3, np.float32)
        self._check(lib().ds_cfa_clip_levels_raw(self.handle, ctypes.byref(r), _ptr(out)))
        return out

    def superpixel(self, cfa, pattern=RGGB, mul=(1.0, 1.0, 1.0)):
        """demosaic.rs:44-87 on a normalised CFA -> (rgb (h/2, w/2, 3), clip (h/2, w/2))."""
        cfa = _plane(cfa, "cfa")
        h, w = cfa.shape
        rgb = np.empty((h // 2, w // 2, 3), np.float32)
        cl = np.empty((h // 2, w // 2), np.float32)
        self._check(lib().ds_superpixel_f32(self.handle, _ptr(cfa), w, h, _pat(pattern), _f3(mul), _ptr(rgb), _ptr(cl)))
        return rgb, cl

    def noise_reduction(self, rgb, nr, field=None, scale=1.0):
        """pipeline/detail.rs:13 noise_reduction -> new (h, w, 3) array; nr = (luminance, luminance_detail,
        luminance_contrast, color, color_detail, color_smoothness)."""
        out = _rgb(rgb, "rgb", copy=True)
        h, w, _ = out.shape
        f = None if field is None else _plane(field, "field", (h, w))
        self._check(lib().ds_noise_reduction(self.handle, _ptr(out), w, h, ctypes.byref(NrParams(*nr)), _ptr(f), float(scale)))
        return out

    def sharpen(self, rgb, sh, field=None, scale=1.0):
        """pipeline/detail.rs:96 sharpen -> new array; sh = (amount, radius, detail, masking)."""
        out = _rgb(rgb, "rgb", copy=True)
        h, w, _ = out.shape
        f = None if field is None else _plane(field, "field", (h, w))
        self._check(lib().ds_sharpen(self.handle, _ptr(out), w, h, ctypes.byref(SharpenParams(*sh)), _ptr(f), float(scale)))
        return out

    def super_res_refine(self, rgb):
        """enhance.rs:122 super_res_refine -> new array."""
        out = _rgb(rgb, "rgb", copy=True)
        h, w, _ = out.shape
        self._check(lib().ds_super_res_refine(self.handle, _ptr(out), w, h))
        return out

    def output_sharpen(self, rgb, mode, amount):
        """export.rs:199 output_sharpen -> new array; mode SHARPEN_FOR_*, amount SHARPEN_AMOUNT_*."""
        out = _rgb(rgb, "rgb", copy=True)
        h, w, _ = out.shape
        self._check(lib().ds_output_sharpen(self.handle, _ptr(out), w, h, int(mode), int(amount)))
        return out

    def fast_guided_filter(self, I, p, r, eps, sub):
        I = _plane(I, "I")
        h, w = I.shape
        p = _plane(p, "p", (h, w))
        out = np.empty_like(I)
        self._check(lib().ds_fast_guided_filter(self.handle, _ptr(I), _ptr(p), w, h, int(r), float(eps), int(sub), _ptr(out)))
        return out

    def plane_resize(self, p, w2, h2):
        p = _plane(p, "p")
        h, w = p.shape
        out = np.empty((h2, w2), np.float32)
        self._check(lib().ds_plane_resize(self.handle, _ptr(p), w, h, int(w2), int(h2), _ptr(out)))
        return out

    def cfa_clip_levels(self, cfa, pattern=RGGB):
        """demosaic.rs:419 cfa_clip_levels on a normalised CFA -> [3] float32."""
        cfa = _plane(cfa, "cfa")
        h, w = cfa.shape
        out = np.zeros(3, np.float32)
        self._check(lib().ds_cfa_clip_levels_f32(self.handle, _ptr(cfa), w, h, _pat(pattern), _ptr(out)))
        return out

    def cfa_clip_levels_u16(self, raw, pattern=RGGB, black=512.0, white=(16383.0,) * 3):
        """cfa_clip_levels of raw DN normalised as decode.rs:291-306 -> [3] float32."""
        raw = _plane(raw, "raw", dtype=np.uint16)
        h, w = raw.shape
        out = np.zeros(3, np.float32)
        self._check(lib().ds_cfa_clip_levels_u16(self.handle, _ptr(raw), w, h, w, _pat(pattern), float(black), _f3(white), _ptr(out)))
        return out

    def reconstruct_highlights(self, rgb, clip, mul, raw_clip):
        """demosaic.rs:435 reconstruct_highlights -> new (h, w, 3) array (the input is not changed)."""
        out = _rgb(rgb, "rgb", copy=True)
        h, w, _ = out.shape
        clip = _plane(clip, "clip", (h, w))
        self._check(lib().ds_reconstruct_highlights(self.handle, _ptr(out), _ptr(clip), w, h, _f3(mul), _f3(raw_clip)))
        return out

    def estimate_noise(self, plane):
        plane = _plane(plane, "plane")
        h, w = plane.shape
        s = np.zeros(1, np.float32)
        self._check(lib().ds_estimate_noise(self.handle, _ptr(plane), w, h, _ptr(s)))
        return float(s[0])

    def nlm(self, plane, h_param, search, patch):
        plane = _plane(plane, "plane")
        h, w = plane.shape
        out = np.empty_like(plane)
        self._check(lib().ds_nlm(self.handle, _ptr(plane), w, h, float(h_param), int(search), int(patch), _ptr(out)))
        return out

    def denoise(self, rgb, amount, search=3, patch=1):
        rgb = _rgb(rgb, "rgb")
        h, w, _ = rgb.shape
        out = np.empty_like(rgb)
        self._check(lib().ds_denoise_ex(self.handle, _ptr(rgb), w, h, float(amount), int(search), int(patch), _ptr(out)))
        return out

    def nlm_reference_order(self, plane, h_param, search, patch, threads):
        """nlm in the reference's summation order for a rayon pool of `threads` threads (bit for bit)."""
        plane = _plane(plane, "plane")
        h, w = plane.shape
        out = np.empty_like(plane)
        self._check(lib().ds_nlm_reference_order(self.handle, _ptr(plane), w, h, float(h_param), int(search), int(patch), int(threads),
                                                 _ptr(out)))
        return out

    def denoise_reference_order(self, rgb, amount, search, patch, threads):
        """denoise with its NLM in the reference's summation order for `threads` threads (bit for bit)."""
        rgb = _rgb(rgb, "rgb")
        h, w, _ = rgb.shape
        out = np.empty_like(rgb)
        self._check(lib().ds_denoise_reference_order(self.handle, _ptr(rgb), w, h, float(amount), int(search), int(patch), int(threads),
                                                     _ptr(out)))
        return out

    def box_blur(self, plane, r):
        plane = _plane(plane, "plane")
        h, w = plane.shape
        out = np.empty_like(plane)
        self._check(lib().ds_box_blur(self.handle, _ptr(plane), w, h, int(r), _ptr(out)))
        return out

    def gaussian_blur(self, plane, sigma):
        plane = _plane(plane, "plane")
        h, w = plane.shape
        out = np.empty_like(plane)
        self._check(lib().ds_gaussian_blur(self.handle, _ptr(plane), w, h, float(sigma), _ptr(out)))
        return out

    def guided_filter(self, guide, src, r, eps):
        guide = _plane(guide, "guide")
        h, w = guide.shape
        src = _plane(src, "src", (h, w))
        out = np.empty_like(guide)
        self._check(lib().ds_guided_filter(self.handle, _ptr(guide), _ptr(src), w, h, int(r), float(eps), _ptr(out)))
        return out

    def oklab_from_prophoto(self, rgb):
        rgb = _triples(rgb, "rgb")
        out = np.empty_like(rgb)
        self._check(lib().ds_oklab_from_prophoto(self.handle, _ptr(rgb), rgb.size // 3, _ptr(out)))
        return out

    def oklab_to_prophoto(self, lab):
        lab = _triples(lab, "lab")
        out = np.empty_like(lab)
        self._check(lib().ds_oklab_to_prophoto(self.handle, _ptr(lab), lab.size // 3, _ptr(out)))
        return out

    def libm_hash(self, fn):
        """ds_libm_hash: order-free hash of ds_cbrtf (fn 0) or ds_expf (fn 1) over all 2^32 inputs, on the GPU."""
        v = ctypes.c_uint64()
        self._check(lib().ds_libm_hash(self.handle, int(fn), ctypes.byref(v)))
        return int(v.value)

    def luminance(self, rgb):
        rgb = _triples(rgb, "rgb")
        out = np.empty(rgb.shape[:-1], np.float32)
        self._check(lib().ds_luminance(self.handle, _ptr(rgb), rgb.size // 3, _ptr(out)))
        return out

    def raw_details(self, rgb, sigma, iterations):
        rgb = _rgb(rgb, "rgb")
        h, w, _ = rgb.shape
        out = np.empty_like(rgb)
        self._check(lib().ds_raw_details(self.handle, _ptr(rgb), w, h, float(sigma), int(iterations), _ptr(out)))
        return out

    def pipeline_batch(self, frames, algo=ALGO_RCD, denoise=DENOISE_NLM, amount=0.0, search=10, patch=3, raw_details=True,
                       rl_sigma=1.0, rl_iters=20, reference_gating=False, net=None, threshold_db=0.3, streams=0, autotune=False,
                       info=True, highlights=False, quality=QUALITY_FULL, host_frames=False):
        """Runs the batch graph over `frames` (a list of Frame); returns the batch info (or None with info=False).
        host_frames: the frames' raw and rgb_out are pinned host buffers (HostBuf)."""
        arr = (Frame * len(frames))(*frames)
        o = PipelineOpts(algo, denoise, amount, search, patch, int(raw_details), rl_sigma, rl_iters, int(reference_gating),
                         net.handle.value if net is not None else None, threshold_db, streams, int(autotune), int(highlights),
                         int(quality), int(host_frames))
        bi = BatchInfo()
        self._check(lib().ds_pipeline_batch(self.handle, arr, len(frames), ctypes.byref(o), ctypes.byref(bi) if info else None))
        return bi.as_dict() if info else None

    def autotune(self, w, h):
        t = Tuning()
        self._check(lib().ds_autotune(self.handle, w, h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, w, h, rcd=0, mhc=0, rl=0):
        t = Tuning(rcd, mhc, rl)
        self._check(lib().ds_set_tuning(self.handle, w, h, ctypes.byref(t)))

    def tuning(self, w, h):
        t = Tuning()
        self._check(lib().ds_get_tuning(self.handle, w, h, ctypes.byref(t)))
        return (t.rcd_variant, t.mhc_variant, t.rl_variant)

    def auto_wb_stats(self, rgb, as_shot_neutral):
        """pipeline/mod.rs:421-441 statistics in FP64 -> dict(acc, n, neutral, valid)."""
        rgb = _rgb(rgb, "rgb")
        h, w, _ = rgb.shape
        st = WbStats()
        self._check(lib().ds_auto_wb_stats(self.handle, _ptr(rgb), w, h, _f3(as_shot_neutral), ctypes.byref(st)))
        return {"acc": list(st.acc), "n": st.n, "neutral": list(st.neutral), "valid": st.valid}

    def denoise_learned(self, net, rgb, amount, search=10, patch=3, threshold_db=0.3):
        """denoise with the learned luminance and the per-region NLM fallback -> (rgb, stats)."""
        rgb = _rgb(rgb, "rgb")
        h, w, _ = rgb.shape
        out = np.empty_like(rgb)
        st = LearnedStats()
        self._check(lib().ds_denoise_learned(self.handle, net.handle, _ptr(rgb), w, h, float(amount), int(search), int(patch),
                                             float(threshold_db), _ptr(out), ctypes.byref(st)))
        return out, st.as_dict()

    def make_pairs(self, seed, frame_w, frame_h, iso, crop, n_crops, search=10, patch=3):
        """Training pairs (x, t), each (n_crops, crop, crop), generated on the GPU."""
        nb = n_crops * crop * crop * 4
        bx, bt = self.alloc(nb), self.alloc(nb)
        try:
            self._check(lib().ds_cnn_make_pairs_dev(self.handle, int(seed), frame_w, frame_h, float(iso), int(search), int(patch), crop,
                                                    n_crops, bx.ptr, bt.ptr))
            return bx.download((n_crops, crop, crop), np.float32), bt.download((n_crops, crop, crop), np.float32)
        finally:
            bx.free()
            bt.free()


class CNN:
    """The learned luminance denoiser (include/darkslide_gpu.h, ds_cnn_*).

    Parameters, forward, loss gradient and Adam steps all run on the GPU;
    the host sees FP32 parameter vectors and float32 batches."""

    def __init__(self, ctx, seed=1):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx._check(lib().ds_cnn_create(ctx.handle, int(seed), ctypes.byref(h)))
        self.handle = h
        ctx._nets.add(self)

    def close(self):
        if self.handle:
            lib().ds_cnn_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def param_count():
        return int(lib().ds_cnn_param_count())

    @property
    def step(self):
        return int(lib().ds_cnn_step(self.handle))

    def params(self):
        p = np.empty(self.param_count(), np.float32)
        self.ctx._check(lib().ds_cnn_get_params(self.handle, _ptr(p), p.size))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        self.ctx._check(lib().ds_cnn_set_params(self.handle, _ptr(p), p.size))

    def save(self, path):
        self.ctx._check(lib().ds_cnn_save(self.handle, os.fsencode(path)))

    def load(self, path):
        self.ctx._check(lib().ds_cnn_load(self.handle, os.fsencode(path)))

    def forward(self, x):
        x = _batch(x, "x")
        n, h, w = x.shape
        y = np.empty_like(x)
        self.ctx._check(lib().ds_cnn_forward(self.handle, _ptr(x), n, h, w, _ptr(y)))
        return y

    def infer(self, x):
        """Fused inference forward (same result as forward)."""
        x = _batch(x, "x")
        n, h, w = x.shape
        y = np.empty_like(x)
        self.ctx._check(lib().ds_cnn_infer(self.handle, _ptr(x), n, h, w, _ptr(y)))
        return y

    def loss_grad(self, x, t, margin):
        x = _batch(x, "x")
        t = _batch(t, "t", x.shape)
        n, h, w = x.shape
        g = np.empty(self.param_count(), np.float32)
        loss = ctypes.c_double()
        self.ctx._check(lib().ds_cnn_loss_grad(self.handle, _ptr(x), _ptr(t), n, h, w, int(margin), ctypes.byref(loss), _ptr(g)))
        return loss.value, g

    def train_step_dev(self, x_dev, t_dev, n, h, w, margin, lr, want_loss=False):
        loss = ctypes.c_double()
        self.ctx._check(lib().ds_cnn_train_step_dev(self.handle, x_dev, t_dev, n, h, w, int(margin), float(lr),
                                                    ctypes.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def denoise_plane(self, L):
        L = _plane(L, "L")
        h, w = L.shape
        out = np.empty_like(L)
        self.ctx._check(lib().ds_cnn_denoise_plane(self.handle, _ptr(L), w, h, _ptr(out)))
        return out


_default = None


def default_context():
    global _default
    if _default is None:
        _default = Context(0)
    return _default


def demosaic(cfa, pattern=RGGB, mul=(1.0, 1.0, 1.0), algo=ALGO_RCD):
    """demosaic.rs:20 demosaic(cfa, mul, Quality::Full) -> (rgb, clip)."""
    return default_context().demosaic(cfa, pattern, mul, algo, clip=True)


def demosaic_u16(raw, pattern=RGGB, black=512.0, white=(16383.0,) * 3, mul=(1.0, 1.0, 1.0), algo=ALGO_RCD):
    """CFA normalisation (decode.rs:291-306) fused with demosaic -> (rgb, clip)."""
    return default_context().demosaic_u16(raw, pattern, black, white, mul, algo, clip=True)


def estimate_noise(plane):
    """enhance.rs:10 estimate_noise."""
    return default_context().estimate_noise(plane)


def nlm(plane, h_param, search, patch):
    """enhance.rs:36 nlm."""
    return default_context().nlm(plane, h_param, search, patch)


def denoise(rgb, amount, search=3, patch=1):
    """enhance.rs:75 denoise (search / patch default to the reference's 3, 1)."""
    return default_context().denoise(rgb, amount, search, patch)


def raw_details(rgb, sigma, iterations):
    """enhance.rs:96 raw_details."""
    return default_context().raw_details(rgb, sigma, iterations)


def synth_bayer(seed, w, h, bits=14, iso=1.0, pattern=RGGB):
    """Seeded synthetic Bayer frame, generated on the host (uint16 (h, w))."""
    out = np.empty((h, w), np.uint16)
    rc = lib().ds_synth_bayer_host(int(seed), int(w), int(h), int(bits), float(iso), _pat(pattern), _ptr(out))
    if rc != 0:
        raise DarkslideError(rc, "invalid synthetic frame arguments")
    return out
