"""Raw ingestion and the Draft superpixel demosaic on the GPU against the FP32
oracle, bit for bit: decode.rs:276-307 normalisation with crops, per-site
black tables and per-colour white levels; demosaic(cfa, mul, Full | Draft)
straight from the raw buffer; cfa_clip_levels of a crop; the batch pipeline
at Draft quality with per-site black levels and highlight reconstruction.
"""
import ctypes
import sys

import numpy as np
import pytest

from conftest import PATTERNS

pytestmark = pytest.mark.gpu

MUL = (2.1, 1.0, 1.6)
WHITE3 = [16383.0, 16000.0, 15500.0]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def raw_buffer(seed, W, H, pattern, saturate=True):
    import darkslide_amd as D
    raw = D.synth_bayer(seed, W, H, 14, 2.0, pattern)
    if saturate:
        yy, xx = np.mgrid[0:H, 0:W]
        raw[(xx - W // 3) ** 2 + (yy - H // 2) ** 2 <= (min(W, H) // 6) ** 2] = 16383
    return raw


CASES = [  # buffer W, H, crop (x, y, w, h), black table, dims
    (64, 48, (0, 0, 64, 48), [512.0], None),
    (70, 51, (3, 1, 60, 44), [512.0], None),
    (70, 51, (2, 2, 66, 40), [600.0, 610.0, 620.0, 630.0], (2, 2, 1)),
    (90, 60, (5, 3, 80, 50), [500, 9, 9, 510, 9, 9, 520, 9, 9, 530, 9, 9], (2, 2, 3)),
    (90, 60, (1, 0, 77, 59), [505.0, 515.0], (2, 1, 1)),
    (33, 20, (1, 1, 31, 17), [512.0, 500.0], (1, 2, 1)),
]


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("pat", ["RGGB", "GBRG"])
def test_ingest_matches_oracle(ctx, oracle_mod, case, pat):
    O = oracle_mod
    W, H, crop, black, dims = CASES[case]
    p = PATTERNS[pat]
    raw = raw_buffer(10 + case, W, H, p)
    want, wpat = O.cfa_ingest_u16(raw, crop, p, black, WHITE3, black_dims=dims)
    got, gpat = ctx.cfa_ingest(raw, crop, p, black, WHITE3, black_dims=dims)
    assert gpat == wpat
    assert np.array_equal(got, want)
    # clip levels of the crop
    assert np.array_equal(ctx.cfa_clip_levels_raw(raw, crop, p, black, WHITE3, black_dims=dims), O.cfa_clip_levels(want, wpat))


@pytest.mark.parametrize("case", range(len(CASES)))
def test_demosaic_raw_full_and_draft(ctx, oracle_mod, case):
    import darkslide_amd as D
    O = oracle_mod
    W, H, crop, black, dims = CASES[case]
    p = PATTERNS["RGGB"]
    raw = raw_buffer(20 + case, W, H, p)
    cfa, cpat = O.cfa_ingest_u16(raw, crop, p, black, WHITE3, black_dims=dims)
    say("demosaic_raw", W, H, crop, dims)
    rgb, clip = ctx.demosaic_raw(raw, crop, p, black, WHITE3, MUL, D.ALGO_RCD, D.QUALITY_FULL, black_dims=dims)
    assert np.array_equal(rgb, O.rcd_bayer(cfa, cpat, MUL))
    assert np.array_equal(clip, O.clip_plane(cfa))
    mhc, _ = ctx.demosaic_raw(raw, crop, p, black, WHITE3, MUL, D.ALGO_MHC, D.QUALITY_FULL, black_dims=dims)
    assert np.array_equal(mhc, O.mhc_bayer(cfa, cpat, MUL))
    srgb, sclip = ctx.demosaic_raw(raw, crop, p, black, WHITE3, MUL, D.ALGO_RCD, D.QUALITY_DRAFT, black_dims=dims)
    wrgb, wclip = O.superpixel(cfa, cpat, MUL)
    assert np.array_equal(srgb, wrgb) and np.array_equal(sclip, wclip)
    assert sclip.max() > 0


def test_superpixel_f32_matches_oracle(ctx, oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(5)
    for (h, w) in [(32, 32), (17, 23), (2, 2), (101, 64)]:
        cfa = (rng.random((h, w)) * 1.1 - 0.05).astype(np.float32)
        cfa[0, 0] = np.nan
        for pat in PATTERNS.values():
            got = ctx.superpixel(cfa, pat, MUL)
            want = O.superpixel(cfa, pat, MUL)
            assert np.array_equal(got[0], want[0], equal_nan=True) and np.array_equal(got[1], want[1])


def test_raw_rejects_bad_tables(ctx):
    import darkslide_amd as D
    raw = np.zeros((8, 8), np.uint16)
    with pytest.raises(D.DarkslideError):  # period 3 black table (non-Bayer)
        ctx.cfa_ingest(raw, (0, 0, 8, 8), (0, 1, 1, 2), [1.0] * 9, [100.0], black_dims=(3, 3, 1))
    with pytest.raises(D.DarkslideError):  # table shorter than its dimensions
        ctx.cfa_ingest(raw, (0, 0, 8, 8), (0, 1, 1, 2), [1.0, 2.0], [100.0], black_dims=(2, 2, 1))
    with pytest.raises(D.DarkslideError):  # crop outside the buffer
        ctx.cfa_ingest(raw, (4, 0, 8, 8), (0, 1, 1, 2), [1.0], [100.0])


FRAMES = [  # W, H, pattern, seed, per-site black
    (400, 300, (0, 1, 1, 2), 51, (510.0, 512.0, 514.0, 516.0)),
    (257, 190, (2, 1, 1, 0), 52, None),
    (130, 98, (1, 0, 2, 1), 53, (500.0, 505.0, 500.0, 505.0)),
]


def test_batch_draft_matches_single_frames(ctx, oracle_mod):
    say("batch draft")
    import darkslide_amd as D
    O = oracle_mod
    L = D.lib()
    bufs, frames = [], []
    try:
        for (W, H, pat, seed, bs) in FRAMES:
            raw = raw_buffer(seed, W, H, pat)
            rb = ctx.alloc(W * H * 2)
            rb.upload(raw)
            out = ctx.alloc((W // 2) * (H // 2) * 12)
            f = D.Frame(rb.ptr.value, W, H, W, (ctypes.c_uint8 * 4)(*pat), 512.0, (16383.0,) * 3, MUL, out.ptr.value)
            if bs is not None:
                f.black_site = (ctypes.c_float * 4)(*bs)
                f.per_site = 1
            frames.append(f)
            bufs.append((rb, out, raw, W, H, pat, bs))
        # superpixel + highlights only: the FP32 oracle
        ctx.pipeline_batch(frames, denoise=D.DENOISE_NONE, raw_details=False, highlights=True, quality=D.QUALITY_DRAFT, streams=2)
        for k, (rb, out, raw, W, H, pat, bs) in enumerate(bufs):
            black, dims = ([512.0], None) if bs is None else (list(bs), (2, 2, 1))
            cfa, cpat = O.cfa_ingest_u16(raw, (0, 0, W, H), pat, black, [16383.0], black_dims=dims)
            rgb, clip = O.superpixel(cfa, cpat, MUL)
            want = O.reconstruct_highlights(rgb, clip, MUL, O.cfa_clip_levels(cfa, cpat))
            got = out.download((H // 2, W // 2, 3), np.float32)
            assert (want != rgb).any(), k
            assert np.array_equal(got, want), k
        # with denoise under reference gating (no RL at Draft): the single-frame chain
        ctx.pipeline_batch(frames, amount=20.0, highlights=True, quality=D.QUALITY_DRAFT, reference_gating=True, streams=2)
        for k, (rb, out, raw, W, H, pat, bs) in enumerate(bufs):
            black, dims = ([512.0], None) if bs is None else (list(bs), (2, 2, 1))
            got = out.download((H // 2, W // 2, 3), np.float32)
            rgb, clip = ctx.demosaic_raw(raw, None, pat, black, [16383.0], MUL, D.ALGO_RCD, D.QUALITY_DRAFT, black_dims=dims)
            lv = ctx.cfa_clip_levels_raw(raw, None, pat, black, [16383.0], black_dims=dims)
            hl = ctx.reconstruct_highlights(rgb, clip, MUL, lv)
            ref = ctx.denoise(hl, 20.0, 10, 3)
            assert np.array_equal(got, ref), k
    finally:
        for (rb, out, *_) in bufs:
            rb.free()
            out.free()
