"""Pins of the oracle's raw ingestion (decode.rs:276-307) and Draft superpixel
demosaic (demosaic.rs:44-87).

The reference's own test flat_field_is_preserved (demosaic.rs:514-537) is
carried over for the Bayer Draft case.  The rest are closed forms a slip in
the crop offset, the black-table indexing, the white lookup, the pattern
shift or the superpixel's clip rule would break; FP32 and FP64 agree.
"""
import numpy as np
import pytest

import oracle as O

RGGB = (0, 1, 1, 2)


def flat_cfa(v, pat, w, h):
    yy, xx = np.mgrid[0:h, 0:w]
    col = np.array(pat)[((yy & 1) << 1) | (xx & 1)]
    return np.array(v, np.float32)[col]


def test_flat_field_is_preserved_draft():
    """demosaic.rs:514-522, Quality::Draft on the RGGB flat field."""
    cfa = flat_cfa([0.2, 0.4, 0.3], RGGB, 32, 32)
    img, _ = O.superpixel(cfa, RGGB, (1.0, 1.0, 1.0))
    p = img[img.shape[0] // 2, img.shape[1] // 2]
    assert abs(p[0] - 0.2) < 1e-4 and abs(p[1] - 0.4) < 1e-4 and abs(p[2] - 0.3) < 1e-4
    assert img.shape == (16, 16, 3)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_superpixel_tile_rules(dt):
    cfa = flat_cfa([0.2, 0.4, 0.3], RGGB, 6, 4).astype(dt)
    cfa[0, 1] = 0.5       # tile (0,0): G samples 0.5 and 0.4 -> mean 0.45
    cfa[0, 3] = 0.999     # tile (1,0): one G clipped -> G keeps the max, clip = 1/4 * 2 = 0.5
    cfa[0, 4] = 1.2       # tile (2,0): R clipped ...
    cfa[1, 4] = 0.996     # ... and one G clipped -> clip = min(2/4 * 2, 1) = 1
    img, clip = O.superpixel(cfa, RGGB, (2.0, 1.0, 0.5))
    assert img.shape == (2, 3, 3) and clip.shape == (2, 3)
    assert np.allclose(img[0, 0], [0.4, 0.45, 0.15], atol=1e-7)
    assert np.allclose(img[0, 1], [0.4, 0.999, 0.15], atol=1e-7) and clip[0, 1] == 0.5
    assert np.allclose(img[0, 2], [2.4, 0.996, 0.15], atol=1e-6) and clip[0, 2] == 1.0
    assert clip[0, 0] == 0.0 and np.all(clip[1] == 0.0)


def test_superpixel_odd_size_and_pattern():
    # odd sizes drop the last row / column; GBRG puts B at (1, 0)
    cfa = flat_cfa([0.1, 0.5, 0.9], (1, 2, 0, 1), 7, 5)
    img, _ = O.superpixel(cfa, (1, 2, 0, 1))
    assert img.shape == (2, 3, 3)
    assert np.allclose(img, [0.1, 0.5, 0.9], atol=1e-7)


def test_ingest_scalar_black_is_normalize():
    rng = np.random.default_rng(0)
    raw = rng.integers(0, 16384, (21, 30)).astype(np.uint16)
    out, pat = O.cfa_ingest_u16(raw, (0, 0, 30, 21), RGGB, [512.0], [16383.0, 16000.0, 15000.0])
    ref = O.cfa_normalize_u16(raw, RGGB, 512.0, [16383.0, 16000.0, 15000.0])
    assert np.array_equal(out, ref) and pat == RGGB


def test_ingest_crop_shifts_pattern_and_origin():
    rng = np.random.default_rng(1)
    raw = rng.integers(0, 16384, (40, 50)).astype(np.uint16)
    white = [16383.0, 16000.0, 15000.0]
    out, pat = O.cfa_ingest_u16(raw, (3, 1, 40, 30), RGGB, [512.0], white)
    assert pat == (2, 1, 1, 0)  # RGGB shifted by (3, 1): BGGR
    ref = O.cfa_normalize_u16(np.ascontiguousarray(raw[1:31, 3:43]), pat, 512.0, white)
    assert np.array_equal(out, ref)


def test_ingest_black_table_and_white():
    raw = np.full((4, 6), 3000, np.uint16)
    # 2x2 table, 3 values per entry (channel 0 is used), pattern RGGB
    black = np.array([[[500, 9, 9], [510, 9, 9]], [[520, 9, 9], [530, 9, 9]]], np.float32)
    white = [4000.0, 5000.0, 6000.0]
    out, _ = O.cfa_ingest_u16(raw, (1, 1, 4, 2), RGGB, black.ravel(), white, black_dims=(2, 2, 3))
    # crop pixel (0, 0) is buffer (1, 1): black 530, colour B -> (3000 - 530) / (6000 - 530)
    assert out[0, 0] == np.float32(3000 - 530) / np.float32(6000 - 530)
    # crop (1, 0) = buffer (2, 1): black 520, colour G
    assert out[0, 1] == np.float32(3000 - 520) / np.float32(5000 - 520)
    # crop (0, 1) = buffer (1, 2): black 510, colour G
    assert out[1, 0] == np.float32(3000 - 510) / np.float32(5000 - 510)
    # crop (1, 1) = buffer (2, 2): black 500, colour R
    assert out[1, 1] == np.float32(3000 - 500) / np.float32(4000 - 500)
    # one white value serves every colour (white_at for len < 3)
    out1, _ = O.cfa_ingest_u16(raw, (0, 0, 2, 2), RGGB, [512.0], [4095.0])
    assert np.all(out1 == np.float32(3000 - 512) / np.float32(4095 - 512))
    # white - black below 1 is clamped to 1 (decode.rs:303)
    out2, _ = O.cfa_ingest_u16(raw, (0, 0, 2, 2), RGGB, [3000.5], [3000.0])
    assert np.all(out2 == np.float32(-0.5))


def test_ingest_fp32_matches_fp64():
    rng = np.random.default_rng(2)
    raw = rng.integers(0, 16384, (33, 47)).astype(np.uint16)
    black = [600.0, 610.0, 620.0, 630.0]
    a, pa = O.cfa_ingest_u16(raw, (5, 2, 40, 30), RGGB, black, [16383.0], black_dims=(2, 2, 1))
    b, pb = O.cfa_ingest_u16(raw, (5, 2, 40, 30), RGGB, black, [16383.0], black_dims=(2, 2, 1), dtype=np.float64)
    assert pa == pb and np.abs(a - b).max() < 1e-7
