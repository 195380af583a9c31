"""ds_estimate_noise_noalloc, the allocation-free form of
ds_estimate_noise_dev, called through the C ABI: the same sigma bit for bit
(and the oracle's), on consecutive calls that leave the shared static scratch
holding a larger plane's samples before a smaller one, and at the size with
the most samples (399,999)."""
import ctypes
import sys

import numpy as np
import pytest

from conftest import RGGB, synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def luminance_plane(O, seed, w, h, iso):
    raw = synth(seed, w, h, 14, iso)
    rgb = O.rcd_bayer(O.cfa_normalize_u16(raw, RGGB, 512.0, [16383.0] * 3))
    return np.ascontiguousarray(O.oklab_from_prophoto(rgb)[..., 0])


def test_noalloc_matches_dev_entry_point(ctx, oracle_mod):
    import darkslide_amd as D
    L = D.lib()
    f = L.ds_estimate_noise_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_void_p]
    f.restype = ctypes.c_int
    # largest first, then smaller: each call must clear what the last one left
    for (w, h, iso) in [(2649, 151 * 3, 8.0), (1200, 800, 2.0), (453, 883, 16.0), (97, 61, 4.0), (5, 4, 1.0), (2, 2, 1.0)]:
        plane = luminance_plane(oracle_mod, 900 + w, w, h, iso) if w >= 16 and h >= 16 else \
            np.random.default_rng(w).random((h, w)).astype(np.float32)
        buf = ctx.alloc(plane.nbytes)
        buf.upload(plane)
        s_dev, s_na = ctx.alloc(4), ctx.alloc(4)
        say(f"estimate_noise {w}x{h}: dev entry point, then noalloc")
        ctx._check(L.ds_estimate_noise_dev(ctx.handle, buf.ptr, w, h, s_dev.ptr))
        ctx._check(f(ctx.handle, buf.ptr, w, h, s_na.ptr))
        ctx.sync()
        a = s_dev.download((1,), np.float32)[0]
        b = s_na.download((1,), np.float32)[0]
        want = np.float32(oracle_mod.estimate_noise(plane))
        say(f"  sigma dev {a!r} noalloc {b!r} oracle {want!r}")
        assert a.tobytes() == b.tobytes() == want.tobytes()
        for x in (buf, s_dev, s_na):
            x.free()


def test_noalloc_rejects_bad_arguments(ctx):
    import darkslide_amd as D
    f = D.lib().ds_estimate_noise_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_void_p]
    f.restype = ctypes.c_int
    s = ctx.alloc(4)
    assert f(ctx.handle, None, 64, 64, s.ptr) == -1
    assert f(ctx.handle, s.ptr, (1 << 31), (1 << 31), s.ptr) == -1
    assert f(None, s.ptr, 4, 4, s.ptr) == -1
    s.free()
