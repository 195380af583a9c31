"""Pins of the oracle's export post-steps: super_res_refine (enhance.rs:122-131)
and output_sharpen (export.rs:199-223).

The reference holds no tests for either.  Each pin is a closed form that the
clamps fix exactly (a step edge whose blur difference exceeds the clamp, an
isolated pixel whose ratio exceeds 2), the (mode, amount) table read off the
above-1 branch of the tone curve, or a float64 restatement with an
independent gaussian at a column where no clamp acts.
"""
import math

import numpy as np
import pytest

import oracle as O

LUMA_SUM = np.float32(0.2880402) + np.float32(0.7118741) + np.float32(0.0000857)


def grey(h, w, v):
    return np.full((h, w, 3), v, np.float32)


def enc64(v):
    if v > 1.0:
        return 1.0 + (v - 1.0) / 2.4
    return 12.92 * max(v, 0.0) if v <= 0.0031308 else 1.055 * v ** (1 / 2.4) - 0.055


def dec64(e):
    if e > 1.0:
        return 1.0 + (e - 1.0) * 2.4
    return e / 12.92 if e <= 0.04045 else ((e + 0.055) / 1.055) ** 2.4


def gauss_row(row, sigma):
    """Independent 1-D gaussian in float64 (normalised exp(-d^2/2s^2) taps,
    radius ceil(3s), clamp-to-edge)."""
    r = math.ceil(3 * sigma)
    k = np.exp(-np.arange(-r, r + 1) ** 2 / (2 * sigma * sigma))
    k /= k.sum()
    n = len(row)
    return np.array([sum(k[j + r] * row[min(max(x + j, 0), n - 1)] for j in range(-r, r + 1)) for x in range(n)])


def test_super_res_identities():
    z = grey(16, 16, 0.0)
    assert np.array_equal(O.super_res_refine(z), z)
    flat = grey(24, 24, 0.3)
    assert np.abs(O.super_res_refine(flat) - flat).max() < 1e-6


def test_super_res_step_is_clamped_in_closed_form():
    # luminance step y = 0.2 | 0.8 (grey 0.04 | 0.64): the two columns either
    # side of the edge differ from their blur by more than 0.05, so
    # d = +-0.05 and r = ((y +- 0.045) / y)^2 exactly.
    img = grey(16, 32, 0.04)
    img[:, 16:] = 0.64
    out = O.super_res_refine(img)
    yl, yr = math.sqrt(0.04 * LUMA_SUM), math.sqrt(0.64 * LUMA_SUM)
    for x, y, s in [(14, yl, -1), (15, yl, -1), (16, yr, 1), (17, yr, 1)]:
        r = ((y + s * 0.045) / y) ** 2
        assert out[8, x, 1] == pytest.approx(img[8, x, 1] * r, rel=2e-6), x
    # far from the edge only the blur's rounding is left
    assert np.abs(out[:, :10] - img[:, :10]).max() < 1e-6 and np.abs(out[:, 22:] - img[:, 22:]).max() < 1e-6


def test_super_res_ratio_clamps():
    # an isolated pixel on black: d = 0.889 y before the clamp
    img = grey(15, 15, 0.0)
    img[7, 7] = 0.0025  # y = 0.05: unclamped ratio (1 + 0.9 * 0.889)^2 = 3.2 -> 2
    assert O.super_res_refine(img)[7, 7, 0] == pytest.approx(0.005, rel=1e-6)
    img[7, 7] = 0.09  # y = 0.3: d clamps to 0.05 -> ((0.3 + 0.045) / 0.3)^2
    y = math.sqrt(0.09 * LUMA_SUM)
    assert O.super_res_refine(img)[7, 7, 0] == pytest.approx(0.09 * ((y + 0.045) / y) ** 2, rel=2e-6)
    # a dark pixel beside a bright area: (y - 0.045)^2 / y^2 < 0.5 -> 0.5
    img = grey(16, 32, 0.0025)
    img[:, 16:] = 0.8
    assert O.super_res_refine(img)[8, 15, 2] == pytest.approx(0.00125, rel=1e-6)


def test_super_res_keeps_chroma():
    rng = np.random.default_rng(1)
    img = (0.05 + rng.random((32, 40, 3)) * 0.6).astype(np.float32)
    out = O.super_res_refine(img)
    r = out / img
    assert np.abs(r[..., 0] - r[..., 1]).max() < 1e-6 and np.abs(r[..., 2] - r[..., 1]).max() < 1e-6
    assert r.min() >= 0.5 - 1e-6 and r.max() <= 2.0 + 1e-6 and np.abs(r - 1).max() > 0.05


def test_output_sharpen_none_and_flat():
    rng = np.random.default_rng(2)
    img = rng.random((20, 30, 3)).astype(np.float32)
    for amt in range(3):
        assert np.array_equal(O.output_sharpen(img, 0, amt), img)
    flat = grey(24, 24, 0.4)
    for mode in (1, 2, 3):
        assert np.abs(O.output_sharpen(flat, mode, 1) - flat).max() < 1e-6


MODES = {1: (0.6, 0.45), 2: (1.0, 0.8), 3: (0.8, 0.6)}
AMOUNTS = {0: 0.6, 1: 1.0, 2: 1.5}


@pytest.mark.parametrize("mode", [1, 2, 3])
@pytest.mark.parametrize("amt", [0, 1, 2])
def test_output_sharpen_table_at_a_step(mode, amt):
    # grey 0.05 | 1.0: both edge columns clamp (|d| > 0.1 for every radius)
    radius, base = MODES[mode]
    k = base * AMOUNTS[amt]
    img = grey(12, 32, 0.05)
    img[:, 16:] = 1.0
    out = O.output_sharpen(img, mode, amt)
    # bright side: e2 = 1 + 0.1 k lies on the tone curve's linear extension
    assert out[6, 16, 1] == pytest.approx(1.0 + 0.24 * k, rel=1e-6)
    # dark side: e2 = enc(0.05) - 0.1 k
    l = enc64(0.05)
    assert out[6, 15, 1] == pytest.approx(0.05 * dec64(l - 0.1 * k) / dec64(l), rel=1e-5)
    # the second bright column is not clamped: its blur from an independent gaussian
    row = np.where(np.arange(32) < 16, l, 1.0)
    d = 1.0 - gauss_row(row, radius)[17]
    assert d < 0.1
    assert out[6, 17, 1] == pytest.approx(1.0 + 2.4 * k * d, rel=1e-5)
    # beyond the taps (radius ceil(3 sigma)) nothing moves
    r = math.ceil(3 * radius)
    assert np.abs(out[:, 16 + r:] - 1.0).max() < 1e-6 and np.abs(out[:, : 16 - r - 1] - 0.05).max() < 1e-6


def test_export_post_steps_fp32_matches_fp64():
    rng = np.random.default_rng(3)
    yy, xx = np.mgrid[0:64, 0:80]
    img = ((0.3 + 0.2 * np.sin(xx / 5.0) * np.cos(yy / 4.0))[..., None] * np.array([1.0, 0.9, 0.7])
           + rng.normal(0, 0.02, (64, 80, 3))).clip(0, None).astype(np.float32)
    a, b = O.super_res_refine(img), O.super_res_refine(img.astype(np.float64))
    assert np.abs(a - b).max() < 5e-6 and np.abs(a - img).max() > 1e-3
    for mode in (1, 2, 3):
        a, b = O.output_sharpen(img, mode, 2), O.output_sharpen(img.astype(np.float64), mode, 2)
        assert np.abs(a - b).max() < 5e-6 and np.abs(a - img).max() > 1e-3
