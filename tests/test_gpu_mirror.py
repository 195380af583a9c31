"""The C++ mirror of the reference API (include/darkslide.hpp) end to end:
tests/cpp/mirror_smoke runs demosaic -> denoise -> raw_details (and the Draft
demosaic) through it and the stages are compared with the oracle."""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import ROOT, maxabs

pytestmark = pytest.mark.gpu
TOL = 2e-5


def test_cpp_mirror_pipeline(oracle_mod, tmp_path):
    O = oracle_mod
    exe = os.path.join(ROOT, "tests", "cpp", "mirror_smoke")
    assert os.path.exists(exe), "build() links tests/cpp/mirror_smoke"
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, file=sys.stderr)
    assert r.returncode == 0
    w, h = 320, 240
    ld = lambda name, *shape: np.fromfile(tmp_path / name, np.float32).reshape(*shape)
    cfa = ld("cfa.f32", h, w)
    rgb, clip = ld("rgb.f32", h, w, 3), ld("clip.f32", h, w)
    ref, rclip = O.demosaic(cfa)
    assert maxabs(rgb, ref) <= TOL and np.array_equal(clip, rclip)
    # Quality::Draft: the superpixel, bit for bit
    drgb, dclip = O.superpixel(cfa, (0, 1, 1, 2), (2.0, 1.0, 1.5))
    assert np.array_equal(ld("draft.f32", h // 2, w // 2, 3), drgb) and np.array_equal(ld("draft_clip.f32", h // 2, w // 2), dclip)
    dn = ld("denoise.f32", h, w, 3)
    dref = O.denoise(ref, 0.0, 10, 3)
    env = max(maxabs(O.denoise(ref, 0.0, 10, 3, nthreads=t), dref) for t in (1, 4, 16))
    assert maxabs(dn, dref) <= max(TOL, 1.5 * env)
    rl = ld("rl.f32", h, w, 3)
    assert maxabs(rl, O.raw_details(dn, 1.0, 20)) <= TOL
    L = np.ascontiguousarray(ref[..., 1])
    sigma = np.float32(r.stdout.split()[1])
    assert sigma == np.float32(O.estimate_noise(L))
    assert maxabs(ld("nlm.f32", h, w), O.nlm(L, np.float32(sigma) * np.float32(0.6), 10, 3)) <= TOL
