"""Highlight reconstruction on the GPU (demosaic.rs:418-502) against the FP32
oracle, bit for bit: cfa_clip_levels on raw and normalised CFAs,
reconstruct_highlights on demosaiced frames with saturated areas (whole
discs, green-only discs, a red-only disc), on a dense random clip plane
whose near-clip test reads running-sum residue, the no-clip early return,
and the batch pipeline with highlights switched on.
"""
import ctypes
import sys

import numpy as np
import pytest

from conftest import PATTERNS

pytestmark = pytest.mark.gpu

F32MAX = float(np.finfo(np.float32).max)
MUL = (2.1, 1.0, 1.6)


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def saturated_raw(seed, w, h, bits=14, pattern=(0, 1, 1, 2)):
    """Synthetic frame with saturated discs: every site (neutral clip), green
    sites only, and red sites only."""
    import darkslide_amd as D
    raw = D.synth_bayer(seed, w, h, bits, 2.0, pattern)
    white = 16383 if bits == 14 else 4095
    yy, xx = np.mgrid[0:h, 0:w]
    col = np.array(pattern, np.uint8)[((yy & 1) << 1) | (xx & 1)]
    rad = max(3, min(w, h) // 8)
    rng = np.random.default_rng(seed)
    for kind in ("all", "green", "green", "red"):
        cx, cy = int(rng.integers(0, w)), int(rng.integers(0, h))
        disc = (xx - cx) ** 2 + (yy - cy) ** 2 <= rad * rad
        if kind == "green":
            disc &= col == 1
        elif kind == "red":
            disc &= col == 0
        raw[disc] = white
    black = 512.0 if bits == 14 else 128.0
    return raw, black, float(white)


def oracle_base(O, raw, pattern, black, white, mul):
    cfa = O.cfa_normalize_u16(raw, pattern, black, [white] * 3)
    rgb = O.rcd_bayer(cfa, pattern, mul)
    clip = O.clip_plane(cfa)
    return cfa, rgb, clip, O.cfa_clip_levels(cfa, pattern)


@pytest.mark.parametrize("w,h,bits,pat", [(300, 200, 14, "RGGB"), (257, 131, 12, "BGGR"), (96, 250, 14, "GRBG"),
                                          (1201, 64, 14, "GBRG"), (5, 4, 12, "RGGB")])
def test_clip_levels_match_oracle(ctx, oracle_mod, w, h, bits, pat):
    say("clip levels", w, h, bits, pat)
    O = oracle_mod
    p = PATTERNS[pat]
    raw, black, white = saturated_raw(7 + w, w, h, bits, p)
    cfa = O.cfa_normalize_u16(raw, p, black, [white] * 3)
    want = O.cfa_clip_levels(cfa, p)
    got = ctx.cfa_clip_levels_u16(raw, p, black, (white,) * 3)
    assert np.array_equal(got, want), (got, want)
    assert np.array_equal(ctx.cfa_clip_levels(cfa, p), want)
    # an unsaturated frame: nothing trusted
    plain = O.cfa_normalize_u16(raw.clip(0, int(black + 0.9 * (white - black))), p, black, [white] * 3)
    assert ctx.cfa_clip_levels(plain, p).tolist() == [F32MAX] * 3


def test_clip_levels_nan_negative_and_threshold(ctx, oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(2)
    cfa = (rng.random((61, 77)) * 0.9 - 0.2).astype(np.float32)
    cfa[0, 0] = np.nan  # skipped, as f32::max skips it
    cfa[3, 5] = np.float32(0.95)  # exactly the trust threshold
    cfa[6, 6] = np.float32(np.nextafter(np.float32(0.95), np.float32(0)))  # just below it
    cfa[9, 2] = np.inf
    for p in PATTERNS.values():
        want = O.cfa_clip_levels(cfa, p)
        got = ctx.cfa_clip_levels(cfa, p)
        assert np.array_equal(got, want), (p, got, want)


@pytest.mark.parametrize("w,h,bits,pat", [(1200, 80, 14, "RGGB"), (257, 131, 12, "BGGR"), (640, 427, 14, "GRBG"),
                                          (97, 1500, 14, "GBRG"), (33, 17, 12, "RGGB")])
def test_reconstruct_matches_oracle(ctx, oracle_mod, w, h, bits, pat):
    say("reconstruct", w, h, bits, pat)
    O = oracle_mod
    p = PATTERNS[pat]
    raw, black, white = saturated_raw(100 + h, w, h, bits, p)
    _, rgb, clip, lv = oracle_base(O, raw, p, black, white, MUL)
    want = O.reconstruct_highlights(rgb, clip, MUL, lv)
    got = ctx.reconstruct_highlights(rgb, clip, MUL, lv)
    changed = (want != rgb).any(axis=-1).mean()
    say(f"  levels {lv}, changed {changed:.4f}, max |gpu - oracle| {np.abs(got - want).max():.3g}")
    assert changed > 0.002
    assert np.array_equal(got, want)


def test_reconstruct_dense_random_clip(ctx, oracle_mod):
    """Clipped samples everywhere: box_blur(clip, 3) > 0 then reads the
    running sums' rounding residue, which the GPU passes reproduce."""
    O = oracle_mod
    rng = np.random.default_rng(3)
    h, w = 96, 250
    img = (rng.random((h, w, 3)) * 1.2).astype(np.float32)
    clip = (rng.random((h, w)) > 0.97).astype(np.float32)
    for mul, rc in [((1.8, 1.0, 1.4), (0.99, 1.0, F32MAX)), ((1.0, 1.0, 1.0), (1.0, 1.0, 1.0))]:
        want = O.reconstruct_highlights(img, clip, mul, rc)
        got = ctx.reconstruct_highlights(img, clip, mul, rc)
        assert (want != img).any(axis=-1).mean() > 0.1
        assert np.array_equal(got, want)


def test_no_clip_is_identity(ctx):
    rng = np.random.default_rng(4)
    img = (rng.random((70, 90, 3)) * 3).astype(np.float32)
    out = ctx.reconstruct_highlights(img, np.zeros((70, 90), np.float32), (1, 1, 1), (1, 1, 1))
    assert np.array_equal(out, img)
    # empty frame
    e = ctx.reconstruct_highlights(np.zeros((0, 4, 3), np.float32), np.zeros((0, 4), np.float32), (1, 1, 1), (1, 1, 1))
    assert e.shape == (0, 4, 3)


def test_reconstruct_rejects_bad_input(ctx):
    import darkslide_amd as D
    with pytest.raises(D.DarkslideError):
        ctx._check(D.lib().ds_reconstruct_highlights(ctx.handle, None, None, 4, 4, D._f3((1, 1, 1)), D._f3((1, 1, 1))))
    with pytest.raises(D.DarkslideError):
        ctx.cfa_clip_levels(np.zeros((4, 4), np.float32), (0, 0, 0, 0))


# ---- batch pipeline with highlights

FRAMES = [  # w, h, bits, pattern, seed
    (400, 300, 14, (0, 1, 1, 2), 41),
    (257, 190, 12, (2, 1, 1, 0), 42),
    (130, 99, 14, (1, 0, 2, 1), 43),
]


class SatBatch:
    def __init__(self, ctx):
        import darkslide_amd as D
        self.ctx = ctx
        self.frames, self.bufs = [], []
        for (w, h, bits, pat, seed) in FRAMES:
            raw, black, white = saturated_raw(seed, w, h, bits, pat)
            rb = ctx.alloc(w * h * 2)
            rb.upload(raw)
            out = ctx.alloc(w * h * 12)
            self.frames.append(D.Frame(rb.ptr.value, w, h, w, (ctypes.c_uint8 * 4)(*pat), black, (white,) * 3, MUL, out.ptr.value))
            self.bufs.append((rb, out, raw, w, h, black, white, pat))

    def outputs(self):
        return [out.download((h, w, 3), np.float32) for (_, out, _, w, h, *_) in self.bufs]

    def free(self):
        for (rb, out, *_) in self.bufs:
            rb.free()
            out.free()


def test_batch_with_highlights(ctx, oracle_mod):
    say("batch highlights")
    import darkslide_amd as D
    O = oracle_mod
    L = D.lib()
    b = SatBatch(ctx)
    try:
        # demosaic + highlights only: the FP32 oracle, bit for bit
        info = ctx.pipeline_batch(b.frames, denoise=D.DENOISE_NONE, raw_details=False, highlights=True, streams=2)
        say(f"  {info}")
        for k, (got, (_, _, raw, w, h, black, white, pat)) in enumerate(zip(b.outputs(), b.bufs)):
            _, rgb, clip, lv = oracle_base(O, raw, pat, black, white, MUL)
            want = O.reconstruct_highlights(rgb, clip, MUL, lv)
            assert (want != rgb).any(), k
            assert np.array_equal(got, want), k
        # the full chain: equal to the single-frame device entry points
        ctx.pipeline_batch(b.frames, amount=20.0, highlights=True, streams=2)
        for k, (got, (rb, _, _, w, h, black, white, pat)) in enumerate(zip(b.outputs(), b.bufs)):
            rgb, clip, lv, dn, out = (ctx.alloc(w * h * 12), ctx.alloc(w * h * 4), ctx.alloc(64), ctx.alloc(w * h * 12),
                                      ctx.alloc(w * h * 12))
            try:
                ctx._check(L.ds_demosaic_u16_dev(ctx.handle, rb.ptr, w, h, w, D._pat(pat), black, D._f3((white,) * 3), D._f3(MUL),
                                                 D.ALGO_RCD, rgb.ptr, clip.ptr))
                ctx._check(L.ds_cfa_clip_levels_u16_dev(ctx.handle, rb.ptr, w, h, w, D._pat(pat), black, D._f3((white,) * 3), lv.ptr))
                ctx._check(L.ds_reconstruct_highlights_dev(ctx.handle, rgb.ptr, clip.ptr, w, h, D._f3(MUL), lv.ptr))
                ctx._check(L.ds_denoise_ex_dev(ctx.handle, rgb.ptr, w, h, 20.0, 10, 3, dn.ptr))
                ctx._check(L.ds_raw_details_dev(ctx.handle, dn.ptr, w, h, 1.0, 20, out.ptr))
                ref = out.download((h, w, 3), np.float32)
            finally:
                for x in (rgb, clip, lv, dn, out):
                    x.free()
            assert np.array_equal(got, ref), k
    finally:
        b.free()
