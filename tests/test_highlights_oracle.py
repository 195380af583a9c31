"""Pins of the oracle's highlight reconstruction (demosaic.rs:418-502).

The reference ships no golden vectors for these two functions, so they are
pinned by known answers a plausible slip would break:
  * cfa_clip_levels: planted maxima on sampled rows (y % 3 == 0) are found,
    maxima on other rows are not, the 0.95 trust threshold, NaN skipped,
    and the colour lookup follows the CFA pattern;
  * reconstruct_highlights: the early return (no positive clip value leaves
    even over-level pixels alone), all-three-clipped pixels become neutral
    at their max, and in a scene of one hue a clipped channel is rebuilt to
    that hue's ratio and then faded by ((n/3)^2)/2 toward the pixel's top
    value -- a closed form a swapped ratio, a dropped trust factor or a
    wrong fade coefficient all miss;
  * the FP32 and FP64 restatements agree on random scenes whose near-clip
    test is exact.  The reference decides "near a clipped sample" by
    box_blur(clip, 3) > 0, and its vertical running sum leaves residues of
    ~1e-9 (of either sign) below every clipped sample, so in FP32 that test
    reads rounding noise; FP64 leaves different residues.  With the clipped
    samples in the last rows no running sum ever removes a nonzero value
    before the bottom edge and both precisions classify alike.
"""
import numpy as np
import pytest

import oracle as O

F32MAX = float(np.finfo(np.float32).max)
RGGB = (0, 1, 1, 2)
GBRG = (1, 2, 0, 1)


def _cfa(w, h, v=0.3, seed=0):
    rng = np.random.default_rng(seed)
    return (rng.random((h, w)) * v).astype(np.float32)


def test_clip_levels_planted_maxima():
    cfa = _cfa(40, 31)
    cfa[6, 4] = 0.97  # R site (even x, even y) on a sampled row
    cfa[3, 4] = 0.5  # G site (even x, odd y) on a sampled row
    cfa[4, 4] = 0.99  # row 4 is not sampled: ignored
    cfa[1, 6] = 0.98  # G at (even x, odd y), row 1 not sampled
    cfa[6, 9] = 0.96  # G at (odd x, even y), sampled
    cfa[9, 11] = 1.02  # B at (odd x, odd y), sampled
    lv = O.cfa_clip_levels(cfa, RGGB)
    assert lv.tolist() == [np.float32(0.97), np.float32(0.96), np.float32(1.02)]


def test_clip_levels_threshold_and_pattern():
    cfa = _cfa(16, 12)
    cfa[0, 0] = np.float32(0.95)  # exactly the threshold: trusted
    cfa[0, 1] = np.float32(np.nextafter(np.float32(0.95), np.float32(0)))  # just below: untrusted
    lv = O.cfa_clip_levels(cfa, RGGB)
    assert lv[0] == np.float32(0.95)
    assert lv[1] == np.float32(F32MAX) and lv[2] == np.float32(F32MAX)
    # the same samples under GBRG: (0,0) is green, (1,0) is blue
    lv = O.cfa_clip_levels(cfa, GBRG)
    assert lv[1] == np.float32(0.95) and lv[2] == np.float32(F32MAX) and lv[0] == np.float32(F32MAX)


def test_clip_levels_nan_and_empty():
    cfa = _cfa(8, 7)
    cfa[0, 0] = np.nan
    cfa[6, 0] = 0.99  # R on a sampled row
    lv = O.cfa_clip_levels(cfa, RGGB)
    assert lv[0] == np.float32(0.99)
    assert O.cfa_clip_levels(np.zeros((0, 5), np.float32)).tolist() == [F32MAX] * 3


def test_no_clip_is_identity():
    rng = np.random.default_rng(1)
    img = (rng.random((20, 30, 3)) * 3).astype(np.float32)  # many values far above the level
    clip = np.zeros((20, 30), np.float32)
    out = O.reconstruct_highlights(img, clip, (1, 1, 1), (1, 1, 1))
    assert np.array_equal(out, img)


def _scene(w=1200, h=64):
    """One hue everywhere (R/G = 4/3, B/G = 11/9, bright enough that every
    unclipped pixel votes with weight 1), a disc whose true colour is the
    same hue at G = 1.0 with G clipped to 0.98, and a disc clipped in every
    channel.  mul = (2, 1, 2), raw_clip = 1: levels (1.94, 0.97, 1.94)."""
    img = np.empty((h, w, 3), np.float32)
    img[...] = (1.2, 0.9, 1.1)
    clip = np.zeros((h, w), np.float32)
    yy, xx = np.mgrid[0:h, 0:w]
    d1 = (xx - 300) ** 2 + (yy - 32) ** 2 <= 9
    d3 = (xx - 900) ** 2 + (yy - 32) ** 2 <= 9
    img[d1] = (np.float32(4 / 3), np.float32(0.98), np.float32(11 / 9))
    img[d3] = (2.5, 0.99, 2.0)
    clip[d1] = 1.0
    clip[d3] = 1.0
    return img, clip, d1, d3


DTYPES = [np.float32, np.float64]


@pytest.mark.parametrize("dt", DTYPES)
def test_one_clipped_channel_rebuilt_to_scene_hue(dt):
    img, clip, d1, d3 = _scene()
    img, clip = img.astype(dt), clip.astype(dt)
    out = O.reconstruct_highlights(img, clip, (2, 1, 2), (1, 1, 1))
    # disc 1: G rebuilt from R and B at the scene ratios -> 1.0, then the fade
    # p += (top - p) * (1/3)^2 * 0.5 with top = R = 4/3
    r, g, b = 4 / 3, 1.0, 11 / 9
    t = (1 / 3) ** 2 * 0.5
    want = np.array([r, g + (r - g) * t, b + (r - b) * t])
    got = out[d1].astype(np.float64)
    assert np.abs(got - want).max() < 2e-5, (got[0], want)
    # disc 3: every channel at or over its level -> neutral at the max
    assert np.array_equal(out[d3], np.full((int(d3.sum()), 3), 2.5, dt))
    # everything else: unclipped, untouched
    rest = ~(d1 | d3)
    assert np.array_equal(out[rest], img[rest])


@pytest.mark.parametrize("dt", DTYPES)
def test_two_clipped_channels_fade(dt):
    """R and G clipped, B below: g from B at the scene ratio, both rebuilt,
    then a fade of ((2/3)^2)/2."""
    img, clip, d1, _ = _scene()
    img[d1] = (np.float32(1.95), np.float32(0.98), np.float32(11 / 9))
    img, clip = img.astype(dt), clip.astype(dt)
    out = O.reconstruct_highlights(img, clip, (2, 1, 2), (1, 1, 1))
    g = (11 / 9) / (11 / 9)  # B / ratio_B
    r = max(1.95, g * 4 / 3)
    b = 11 / 9
    top = max(r, g, b)
    t = (2 / 3) ** 2 * 0.5
    want = np.array([r + (top - r) * t, g + (top - g) * t, b + (top - b) * t])
    got = out[d1].astype(np.float64)
    assert np.abs(got - want).max() < 2e-5, (got[0], want)


@pytest.mark.parametrize("dt", DTYPES)
def test_without_evidence_ratios_fall_back_to_neutral(dt):
    """No pixel votes (every unclipped pixel is too dark: G <= 1e-4), so
    wv = 0, the ratios are 1 and a lone clipped G becomes max of the other
    channels before the fade."""
    img, clip, d1, _ = _scene()
    img[...] = (0.0, 0.0, 0.0)
    img[d1] = (0.5, 0.98, 0.6)
    img, clip = img.astype(dt), clip.astype(dt)
    out = O.reconstruct_highlights(img, clip, (2, 1, 2), (1, 1, 1))
    g = max(0.98, 0.5, 0.6)
    top = g
    t = (1 / 3) ** 2 * 0.5
    want = np.array([0.5 + (top - 0.5) * t, g, 0.6 + (top - 0.6) * t])
    assert np.abs(out[d1].astype(np.float64) - want).max() < 1e-6


@pytest.mark.parametrize("seed", [3, 4])
def test_fp32_matches_fp64(seed):
    rng = np.random.default_rng(seed)
    h, w = 96, 250
    img = (rng.random((h, w, 3)) * 1.2).astype(np.float32)
    clip = np.zeros((h, w), np.float32)
    clip[h - 4:] = (rng.random((4, w)) > 0.7).astype(np.float32)
    mul, rc = (1.8, 1.0, 1.4), (0.99, 1.0, F32MAX)
    a = O.reconstruct_highlights(img, clip, mul, rc)
    b = O.reconstruct_highlights(img.astype(np.float64), clip.astype(np.float64), mul, rc)
    assert (a != img).any(axis=-1).mean() > 0.1
    assert np.abs(a - b).max() < 1e-5
