"""Pins of the oracle's detail stages (pipeline/detail.rs) and their helpers:
the sRGB tone curves (color.rs:185-192, pipeline/tone.rs:328-335),
Plane::resize (image.rs:238-262), fast_guided_filter (image.rs:408-431),
noise_reduction (detail.rs:13-92) and sharpen (detail.rs:96-139).

No golden vectors exist for these; each pin is a closed form, an identity
the reference's early returns imply, or a reduction to an already pinned
primitive, chosen so a plausible slip (a swapped coefficient, a wrong
resize footprint, a dropped clamp, a sign) fails.  FP32 and FP64 agree.
"""
import numpy as np
import pytest

import oracle as O


def scene(h=96, w=128, noise=0.02, seed=0):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    base = (0.3 + 0.2 * np.sin(xx / 9.0) * np.cos(yy / 7.0))[..., None] * np.array([1.0, 0.9, 0.7])
    return (base + rng.normal(0, noise, base.shape)).clip(0, None).astype(np.float32)


def test_tone_curves():
    assert O.tone_enc(0.001) == pytest.approx(0.01292, abs=1e-9)
    assert O.tone_dec(0.01) == pytest.approx(0.01 / 12.92, rel=1e-6)
    assert O.tone_dec(2.0) == pytest.approx(1.0 + 2.4, rel=1e-7)
    assert O.tone_enc(3.4) == pytest.approx(1.0 + 2.4 / 2.4, rel=1e-7)
    for v in [0.0005, 0.003, 0.01, 0.18, 0.5, 0.99, 1.0, 1.7]:
        assert O.tone_dec(O.tone_enc(v)) == pytest.approx(v, rel=3e-6)
    # continuous at the knee and at 1
    assert O.tone_enc(0.0031308) == pytest.approx(O.tone_enc(0.00313081), abs=1e-6)
    assert O.tone_enc(1.0) == pytest.approx(O.tone_enc(1.0000001), abs=1e-6)


def test_resize_shrink_is_block_mean():
    rng = np.random.default_rng(1)
    p = rng.random((40, 60)).astype(np.float32)
    got = O.plane_resize(p, 15, 10)  # 4x exactly
    want = p.reshape(10, 4, 15, 4).mean(axis=(1, 3))
    assert np.abs(got - want).max() < 1e-6
    # non-integer shrink: every output is a mean of a footprint of the source
    got = O.plane_resize(p, 23, 17)
    assert got.shape == (17, 23) and got.min() >= p.min() and got.max() <= p.max()
    assert np.array_equal(O.plane_resize(p, 60, 40), p)


def test_resize_grow_is_bilinear():
    ramp = np.tile(np.arange(16, dtype=np.float32), (8, 1))
    up = O.plane_resize(ramp, 32, 16)
    # interior samples of a linear ramp stay on it: centre x' maps to (x' + 0.5) / 2 - 0.5
    xs = (np.arange(32) + 0.5) / 2 - 0.5
    assert np.allclose(up[:, 1:-1], np.broadcast_to(xs[1:-1], (16, 30)), atol=1e-6)
    assert np.all(up[:, 0] == 0.0) and np.all(up[:, -1] == 15.0)  # clamped edges


def test_fast_guided_filter_closed_forms():
    rng = np.random.default_rng(2)
    I = rng.random((64, 80)).astype(np.float32)
    p = rng.random((64, 80)).astype(np.float32)
    # sub 1 (and too small a grid) is the plain guided filter
    assert np.array_equal(O.fast_guided_filter(I, p, 3, 1e-3, 1), O.guided_filter(I, p, 3, 1e-3))
    assert np.array_equal(O.fast_guided_filter(I[:20, :20], p[:20, :20], 3, 1e-3, 4), O.guided_filter(I[:20, :20], p[:20, :20], 3, 1e-3))
    # a constant input passes unchanged; a linear function of the guide is reproduced
    c = np.full_like(I, 0.37)
    assert np.abs(O.fast_guided_filter(I, c, 6, 1e-3, 2) - 0.37).max() < 1e-5  # running-sum rounding only
    lin = (2.0 * I + 0.1).astype(np.float32)
    assert np.abs(O.fast_guided_filter(I, lin, 6, 1e-9, 2) - lin).max() < 2e-4
    # and the output is not the guide itself (the coefficients really are applied)
    assert np.abs(O.fast_guided_filter(I, p, 6, 1e-3, 2) - I).max() > 0.1


def test_noise_reduction_identities():
    img = scene()
    assert np.array_equal(O.noise_reduction(img, (0, 50, 0, 0, 50, 50)), img)
    assert np.array_equal(O.noise_reduction(img, (60, 50, 0, 40, 50, 50), scale=0.3), img)
    # a zero noise field is the same as no field
    nr = (35, 40, 20, 30, 50, 50)
    assert np.array_equal(O.noise_reduction(img, nr, field=np.zeros(img.shape[:2])), O.noise_reduction(img, nr))
    # colour NR on a neutral image leaves it neutral (Oklab a = b = 0 stays 0)
    grey = np.repeat(img[..., 1:2], 3, axis=2)
    out = O.noise_reduction(grey, (0, 50, 0, 60, 50, 50))
    assert np.abs(out - grey).max() < 2e-6
    # luminance NR of a flat image changes nothing
    flat = np.full_like(img, 0.25)
    assert np.abs(O.noise_reduction(flat, (80, 20, 50, 0, 50, 50)) - flat).max() < 2e-6


def test_noise_reduction_reduces_noise_and_field_gates_it():
    flat = np.full((64, 64, 3), 0.2, np.float32)
    noisy = (flat + np.random.default_rng(3).normal(0, 0.01, flat.shape)).astype(np.float32)
    out = O.noise_reduction(noisy, (80, 20, 0, 60, 30, 50))
    assert np.std(out - flat) < 0.6 * np.std(noisy - flat)
    # lum 0 and a field only on the left half: the right half's luminance is untouched
    field = np.zeros((64, 64), np.float32)
    field[:, :32] = 70.0
    out = O.noise_reduction(noisy, (0, 20, 0, 0, 50, 50), field=field)
    assert np.abs(out[:, 40:] - noisy[:, 40:]).max() < 2e-6
    assert np.abs(out[:, :24] - noisy[:, :24]).max() > 1e-3


@pytest.mark.parametrize("nr", [(50, 50, 0, 0, 50, 50), (0, 50, 0, 25, 50, 50), (60, 30, 40, 40, 20, 80)])
def test_noise_reduction_fp32_matches_fp64(nr):
    img = scene()
    a = O.noise_reduction(img, nr)
    b = O.noise_reduction(img.astype(np.float64), nr)
    assert np.abs(a - b).max() < 5e-5 and np.abs(a - img).max() > 0.01


def test_sharpen_identities_and_signs():
    img = scene()
    assert np.array_equal(O.sharpen(img, (0, 1, 25, 0)), img)
    assert np.array_equal(O.sharpen(img, (40, 1, 25, 0), scale=0.2), img)  # sigma 0.2 < 0.25
    flat = np.full((32, 32, 3), 0.3, np.float32)
    assert np.array_equal(O.sharpen(flat, (90, 2, 25, 0)), flat)
    assert np.array_equal(O.sharpen(img, (40, 1, 25, 0), field=np.zeros(img.shape[:2])), O.sharpen(img, (40, 1, 25, 0)))
    # a vertical step: the bright side overshoots up, the dark side down
    step = np.full((16, 32, 3), 0.1, np.float32)
    step[:, 16:] = 0.5
    out = O.sharpen(step, (100, 1.5, 100, 0))
    assert out[8, 16, 1] > 0.5 and out[8, 15, 1] < 0.1
    # far from the edge only the blur's rounding is left
    assert np.abs(out[:, :8] - step[:, :8]).max() < 1e-6 and np.abs(out[:, 26:] - step[:, 26:]).max() < 1e-6


def test_sharpen_masking_spares_flat_texture():
    rng = np.random.default_rng(4)
    tex = (0.3 + rng.normal(0, 1e-4, (32, 32, 1)) * np.ones(3)).astype(np.float32)
    # masking 100: t = 0.06, gradients of a 1e-4 texture sit far below t / 2 -> mask 0 -> unchanged
    assert np.array_equal(O.sharpen(tex, (80, 1, 25, 100)), tex)
    assert not np.array_equal(O.sharpen(tex, (80, 1, 25, 0)), tex)


@pytest.mark.parametrize("sh", [(40, 1, 25, 0), (80, 2.5, 10, 40), (40, 0.5, 100, 0)])
def test_sharpen_fp32_matches_fp64(sh):
    img = scene()
    a = O.sharpen(img, sh)
    b = O.sharpen(img.astype(np.float64), sh)
    assert np.abs(a - b).max() < 5e-6 and np.abs(a - img).max() > 0.01
