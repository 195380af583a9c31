"""C ABI checks that need no GPU: the library loads, exports every symbol
include/darkslide_gpu.h declares, reports a missing device as an error, and
its host-side synthetic generator is deterministic."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT, RGGB

HEADER = os.path.join(ROOT, "include", "darkslide_gpu.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b(ds_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


@pytest.fixture(scope="module")
def L():
    import darkslide_amd as D
    return D.lib()


def test_exports_every_declared_symbol(L):
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_binding_covers_header():
    import darkslide_amd as D
    assert set(declared_symbols()) <= set(D._SIGS)


def test_version(L):
    assert b"gfx950" in L.ds_version()


def test_no_device_is_an_error(L):
    import darkslide_amd as D
    h = ctypes.c_void_p()
    rc = L.ds_ctx_create(0, ctypes.byref(h))
    if rc == 0:
        L.ds_ctx_destroy(h)
        pytest.skip("a GPU is present")
    assert rc == -2
    assert L.ds_last_error(None)
    with pytest.raises(D.DarkslideError):
        D.Context(0)


def test_null_context_rejected(L):
    assert L.ds_sync(None) == -1
    assert L.ds_prof_get(None, None, 0) == -1


def test_synth_host_deterministic_and_in_range():
    import darkslide_amd as D
    a = D.synth_bayer(0x5EED0000, 96, 64)
    b = D.synth_bayer(0x5EED0000, 96, 64)
    c = D.synth_bayer(0x5EED0001, 96, 64)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a.min() >= 0 and a.max() <= 16383
    lin = (a.astype(np.float64) - 512) / 15871
    assert 0.0 < lin.mean() < 0.85
    a12 = D.synth_bayer(1, 64, 64, bits=12)
    assert a12.max() <= 4095
    with pytest.raises(D.DarkslideError):
        D.synth_bayer(1, 8, 8, pattern=(0, 0, 1, 2))
    with pytest.raises(D.DarkslideError):
        D.synth_bayer(1, 8, 8, bits=10)


def test_synth_noise_scales_with_iso(oracle_mod):
    """Poisson-Gaussian noise: the robust noise estimate of the demosaiced
    luminance grows about as sqrt(ISO) (shot noise dominates read noise)."""
    import darkslide_amd as D
    O = oracle_mod
    def sigma(iso):
        raw = D.synth_bayer(7, 256, 256, iso=iso)
        cfa = O.cfa_normalize_u16(raw, RGGB, 512.0, [16383.0] * 3)
        return O.estimate_noise(O.luminance(O.rcd_bayer(cfa)))
    s1, s4, s16 = sigma(1.0), sigma(4.0), sigma(16.0)
    assert s1 < s4 < s16
    assert 1.8 < s16 / s1 < 5.0
