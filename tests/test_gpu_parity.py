"""GPU parity of the C ABI against the CPU oracle.

Gate (enhance / demosaic outputs on [0,1] linear data): max-abs <= 2e-5 against
the FP32 restatement of the reference.  Small and config-1 sizes compare whole
frames; the 6000x4000 frame compares whole frames where the oracle is fast
(demosaic, noise estimate) and crops elsewhere, using the locality of each
stage: a pixel of NLM depends only on pixels within search + patch, a pixel of
RL on pixels within 2 r iterations, so the interior of an oracle run on a
crop with that margin equals the full-frame result.
"""
import sys

import numpy as np
import pytest

from conftest import PATTERNS, RGGB, maxabs, synth

pytestmark = pytest.mark.gpu

TOL = 2e-5
WHITE = (16383.0, 16383.0, 16383.0)


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def oracle_demosaic(O, raw, pat, mul, algo, black=512.0, white=WHITE):
    cfa = O.cfa_normalize_u16(raw, pat, black, list(white))
    rgb = O.rcd_bayer(cfa, pat, mul) if algo == 0 else O.mhc_bayer(cfa, pat, mul)
    return rgb, O.clip_plane(cfa), cfa


# ------------------------------------------------------------------ demosaic

def test_demosaic_rcd_config1(ctx, oracle_mod):
    """1024x1024 RGGB uint16 14-bit, black 512, white 16383 (the CPU config)."""
    O = oracle_mod
    raw = synth(0x5EED0000, 1024, 1024)
    rgb, clip = ctx.demosaic_u16(raw, RGGB, 512.0, WHITE, (1, 1, 1), 0)
    want, wclip, _ = oracle_demosaic(O, raw, RGGB, (1, 1, 1), 0)
    err = maxabs(rgb, want)
    say(f"rcd 1024^2 max-abs {err:.3g}, bit-equal {np.mean(rgb == want):.6f}")
    assert err <= TOL
    assert np.array_equal(clip, wclip)


@pytest.mark.parametrize("pname", list(PATTERNS))
@pytest.mark.parametrize("algo", [0, 1])
@pytest.mark.parametrize("w,h", [(64, 64), (97, 61), (16, 16), (15, 40), (130, 67)])
def test_demosaic_patterns_sizes(ctx, oracle_mod, pname, algo, w, h):
    O = oracle_mod
    pat = PATTERNS[pname]
    raw = synth(1000 + w + h, w, h, 14, 2.0, pat)
    mul = (2.1, 1.0, 1.6)
    rgb, clip = ctx.demosaic_u16(raw, pat, 512.0, WHITE, mul, algo)
    want, wclip, cfa = oracle_demosaic(O, raw, pat, mul, algo)
    assert maxabs(rgb, want) <= TOL
    assert np.array_equal(clip, wclip)
    # the f32 entry point on the normalised CFA gives the same image
    rgb2, clip2 = ctx.demosaic(cfa, pat, mul, algo)
    assert maxabs(rgb2, want) <= TOL
    assert np.array_equal(clip2, wclip)


def test_demosaic_clipped_and_negative(ctx, oracle_mod):
    """Clipped highlights set the clip plane; below-black DNs go negative in
    the CFA and are floored by RCD/MHC as in the reference."""
    O = oracle_mod
    raw = synth(77, 200, 120)
    raw[30:50, 40:90] = 16383
    raw[80:90, 10:20] = 300
    white = (16383.0, 16000.0, 15000.0)
    rgb, clip = ctx.demosaic_u16(raw, RGGB, 512.0, white, (1.9, 1.0, 1.4), 0)
    want, wclip, _ = oracle_demosaic(O, raw, RGGB, (1.9, 1.0, 1.4), 0, 512.0, white)
    assert clip.sum() > 0
    assert np.array_equal(clip, wclip)
    assert maxabs(rgb, want) <= TOL


def test_demosaic_strided_12bit(ctx, oracle_mod):
    """12-bit frame (black 128, white 4095) read from a padded row stride."""
    import darkslide_amd as D
    O = oracle_mod
    w, h = 300, 200
    raw = synth(91, w, h, 12, 3.0)
    padded = np.zeros((h, w + 13), np.uint16)
    padded[:, :w] = raw
    rgb = np.empty((h, w, 3), np.float32)
    L = D.lib()
    rc = L.ds_demosaic_u16(ctx.handle, D._ptr(padded), w, h, w + 13, D._pat(RGGB), 128.0, D._f3((4095,) * 3), D._f3((1, 1, 1)), 0,
                           D._ptr(rgb), None)
    assert rc == 0
    want, _, _ = oracle_demosaic(O, raw, RGGB, (1, 1, 1), 0, 128.0, (4095.0,) * 3)
    assert maxabs(rgb, want) <= TOL


def test_demosaic_rejects_bad_input(ctx):
    import darkslide_amd as D
    raw = synth(5, 32, 32)
    with pytest.raises(D.DarkslideError) as e:
        ctx.demosaic_u16(raw, (0, 0, 1, 2))
    assert e.value.code == -1
    with pytest.raises(D.DarkslideError):
        ctx.demosaic_u16(raw, RGGB, algo=7)


# ------------------------------------------------------------------ noise / NLM

@pytest.mark.parametrize("w,h", [(64, 64), (3, 3), (2, 9), (640, 480), (1024, 1024)])
def test_estimate_noise_exact(ctx, oracle_mod, w, h):
    O = oracle_mod
    rng = np.random.default_rng(w * 7 + h)
    p = rng.random((h, w)).astype(np.float32)
    assert ctx.estimate_noise(p) == O.estimate_noise(p)


@pytest.mark.parametrize("kind", ["constant", "quantized", "octaves", "smooth"])
def test_estimate_noise_concentrated(ctx, oracle_mod, kind):
    """Sample sets whose rank falls in a crowded bin: all-equal laplacians
    (constant), a few repeated values (quantized: the chosen bins hold far
    more samples than fit in LDS), values over 30 octaves, and a smooth
    gradient with tiny noise at a large size."""
    O = oracle_mod
    rng = np.random.default_rng(11)
    w, h = (1024, 768) if kind != "smooth" else (3000, 2000)
    if kind == "constant":
        p = np.full((h, w), 0.37, np.float32)
    elif kind == "quantized":
        p = (rng.integers(0, 4, (h, w)) / np.float32(4)).astype(np.float32)
    elif kind == "octaves":
        p = (10.0 ** rng.uniform(-9, 0, (h, w))).astype(np.float32)
    else:
        yy, xx = np.mgrid[0:h, 0:w].astype(np.float32)
        p = (0.2 + 1e-4 * xx + 5e-5 * yy + 1e-6 * rng.standard_normal((h, w))).astype(np.float32)
    assert ctx.estimate_noise(p) == O.estimate_noise(p)


def luminance_plane(O, seed, w, h, iso=1.0):
    raw = synth(seed, w, h, 14, iso)
    cfa = O.cfa_normalize_u16(raw, RGGB, 512.0, list(WHITE))
    return np.ascontiguousarray(O.oklab_from_prophoto(O.rcd_bayer(cfa))[..., 0])


# search + patch <= 13 takes the narrow tile instance, 14..24 the wide one,
# beyond that the direct kernel
@pytest.mark.parametrize("S,P", [(3, 1), (10, 3), (1, 0), (2, 2), (5, 5), (0, 2), (20, 5), (11, 3), (10, 4), (15, 4), (19, 5)])
@pytest.mark.parametrize("w,h", [(96, 64), (130, 70), (7, 5)])
def test_nlm(ctx, oracle_mod, S, P, w, h):
    O = oracle_mod
    L = luminance_plane(O, 300 + S * 10 + P, w, h, iso=4.0)
    s = O.estimate_noise(L) if w * h > 100 else 0.01
    hp = max(s, 1e-4) * 0.6
    got = ctx.nlm(L, hp, S, P)
    want = O.nlm(L, hp, S, P)
    err = maxabs(got, want)
    say(f"nlm S={S} P={P} {w}x{h} max-abs {err:.3g}")
    assert err <= TOL


def test_nlm_negative_search_is_zero(ctx):
    p = np.random.default_rng(1).random((20, 30)).astype(np.float32)
    assert np.array_equal(ctx.nlm(p, 0.1, -1, 1), np.zeros_like(p))


def test_nlm_reference_test(ctx, oracle_mod):
    """enhance.rs:137-151 on the GPU path."""
    from test_oracle_pins import lcg_noise_plane
    p, noisy = lcg_noise_plane()
    s = ctx.estimate_noise(noisy)
    assert 0.01 < s < 0.06
    d = ctx.nlm(noisy, s * 1.5, 3, 1)
    err = lambda q: float(((q.astype(np.float64) - p) ** 2).sum())
    assert err(d) < err(noisy) * 0.5
    assert maxabs(d, oracle_mod.nlm(noisy, s * 1.5, 3, 1)) <= TOL


# ------------------------------------------------------------------ denoise / RL

def rgb_frame(O, seed, w, h, iso=4.0):
    raw = synth(seed, w, h, 14, iso)
    cfa = O.cfa_normalize_u16(raw, RGGB, 512.0, list(WHITE))
    return O.rcd_bayer(cfa)


def thread_runs(O, rgb, amount, S, P):
    """The reference's own denoise at rayon pool sizes 1..16: NLM's chunked
    sums (enhance.rs:43-70) change L by ~1e-6, and the guided filter's
    variance term amplifies that (test_oracle_pins pins it)."""
    return [O.denoise(rgb, amount, S, P, nthreads=t) for t in (1, 2, 4, 8, 16)]


def thread_envelope(O, rgb, amount, S, P):
    runs = thread_runs(O, rgb, amount, S, P)
    return max(maxabs(a, b) for a in runs for b in runs)


@pytest.mark.parametrize("amount,S,P", [(0.0, 10, 3), (37.0, 3, 1), (100.0, 3, 1)])
def test_denoise_stages(ctx, oracle_mod, amount, S, P):
    """Every stage of denoise on identical inputs within 2e-5: Oklab split,
    noise estimate, NLM, both guided filters, Oklab back."""
    O = oracle_mod
    rgb = rgb_frame(O, 400, 256, 192, iso=8.0)
    lab = O.oklab_from_prophoto(rgb)
    glab = ctx.oklab_from_prophoto(rgb)
    assert np.array_equal(glab, lab)
    L, A, B = (np.ascontiguousarray(lab[..., i]) for i in range(3))
    sigma = max(O.estimate_noise(L), np.float32(1e-4))
    assert ctx.estimate_noise(L) == O.estimate_noise(L)
    k = np.float32(amount) / np.float32(100)
    hp = np.float32(sigma) * (np.float32(0.6) + np.float32(1.6) * k)
    lf = O.nlm(L, hp, S, P)
    assert maxabs(ctx.nlm(L, hp, S, P), lf) <= TOL
    l2 = (L + (lf - L) * (np.float32(0.5) + np.float32(0.5) * k)).astype(np.float32)
    r = int(max((np.float32(256) / np.float32(600)) * (np.float32(1) + np.float32(2) * k), np.float32(2)))
    eps = np.float32(np.float32(sigma) * np.float32(2)) ** 2
    a2, b2 = O.guided_filter(l2, A, r, eps), O.guided_filter(l2, B, r, eps)
    ga2, gb2 = ctx.guided_filter(l2, A, r, eps), ctx.guided_filter(l2, B, r, eps)
    say(f"guided filter on identical inputs: max-abs {maxabs(ga2, a2):.3g} / {maxabs(gb2, b2):.3g}")
    assert maxabs(ga2, a2) <= TOL and maxabs(gb2, b2) <= TOL
    back = np.stack([l2, a2, b2], -1).astype(np.float32)
    assert maxabs(ctx.oklab_to_prophoto(back), O.oklab_to_prophoto(back)) <= TOL


@pytest.mark.parametrize("amount,S,P", [(0.0, 10, 3), (37.0, 3, 1), (100.0, 19, 5), (50.0, 22, 3)])
def test_denoise_is_its_stages(ctx, oracle_mod, amount, S, P):
    """denoise equals, bit for bit, the GPU's own stages run one by one:
    Oklab split, noise estimate, NLM (tile kernel and, for search + patch
    > 24, the direct one), both guided filters on l2, Oklab back, max 0."""
    rgb = rgb_frame(oracle_mod, 401, 250, 190, iso=8.0)
    lab = ctx.oklab_from_prophoto(rgb)
    L, A, B = (np.ascontiguousarray(lab[..., i]) for i in range(3))
    sigma = np.float32(max(np.float32(ctx.estimate_noise(L)), np.float32(1e-4)))
    k = np.float32(amount) / np.float32(100)
    hp = np.float32(sigma * (np.float32(0.6) + np.float32(1.6) * k))
    lf = ctx.nlm(L, hp, S, P)
    l2 = (L + (lf - L) * (np.float32(0.5) + np.float32(0.5) * k)).astype(np.float32)
    r = int(max((np.float32(250) / np.float32(600)) * (np.float32(1) + np.float32(2) * k), np.float32(2)))
    eps = np.float32(sigma * np.float32(2)) ** 2
    a2, b2 = ctx.guided_filter(l2, A, r, eps), ctx.guided_filter(l2, B, r, eps)
    want = np.maximum(ctx.oklab_to_prophoto(np.stack([l2, a2, b2], -1).astype(np.float32)), np.float32(0))
    got = ctx.denoise(rgb, amount, S, P)
    assert np.array_equal(got, want), f"bit-equal {np.mean(got == want):.6f}"


@pytest.mark.parametrize("amount,S,P", [(0.0, 10, 3), (37.0, 3, 1), (100.0, 3, 1)])
def test_denoise_composite(ctx, oracle_mod, amount, S, P):
    """The composite can only be held to the reference's own reproducibility:
    its output moves by ~2-3e-4 between rayon pool sizes, so the GPU result
    must fall inside that envelope, and be as close to the FP64 oracle as the
    FP32 reference is."""
    O = oracle_mod
    rgb = rgb_frame(O, 400, 256, 192, iso=8.0)
    got = ctx.denoise(rgb, amount, S, P)
    want = O.denoise(rgb, amount, S, P)
    runs = thread_runs(O, rgb, amount, S, P)
    env = max(maxabs(a, b) for a in runs for b in runs)
    f64 = O.denoise(rgb.astype(np.float64), amount, S, P)
    err, e64 = maxabs(got, want), maxabs(got, f64)
    r64 = max(maxabs(a, f64) for a in runs)
    say(f"denoise amount={amount} S={S} P={P}: vs fp32 {err:.3g} (reference thread envelope {env:.3g}); "
        f"vs fp64 {e64:.3g} (reference vs fp64, worst pool size {r64:.3g})")
    assert err <= max(TOL, env)
    assert e64 <= max(TOL, 1.5 * r64)


@pytest.mark.parametrize("r", [0, 1, 4, 10, 63, 70])
def test_box_blur_exact(ctx, oracle_mod, r):
    """image.rs:269 with the reference's running sums: bit-identical."""
    rng = np.random.default_rng(r)
    p = rng.random((150, 333)).astype(np.float32)
    got, want = ctx.box_blur(p, r), oracle_mod.box_blur(p, r)
    say(f"box_blur r={r}: bit-equal {np.mean(got == want):.6f}")
    assert np.array_equal(got, want)


@pytest.mark.parametrize("w,h,r", [(336, 150, 1), (336, 150, 3), (336, 150, 23), (336, 150, 24), (336, 150, 31), (336, 150, 32),
                                   (336, 150, 63), (336, 150, 64), (336, 150, 87), (336, 150, 88), (336, 150, 95), (336, 150, 96),
                                   (4, 7, 2), (100, 3, 40), (64, 64, 5), (1000, 1, 3), (8, 500, 30), (1, 300, 7), (132, 260, 50)])
def test_box_blur_ring_paths(ctx, oracle_mod, w, h, r):
    """Widths divisible by 4 take the ring-buffered passes (both ring sizes,
    and the fallbacks past their radius limits); still bit-identical."""
    rng = np.random.default_rng(w * 1000 + r)
    p = rng.random((h, w)).astype(np.float32)
    got, want = ctx.box_blur(p, r), oracle_mod.box_blur(p, r)
    assert np.array_equal(got, want), f"bit-equal {np.mean(got == want):.6f}"


# Every ring variant on both sides of its radius limit: row passes use the
# 256-column ring to r 56 and the 512-column one to r 184; column passes over
# widths divisible by 4 move 16-byte units (256-row ring to r 60, 512 to
# r 188), other widths 4-byte ones (to r 63 / 191).  Tall and wide planes
# give many blocks per plane.
@pytest.mark.parametrize("w,h,r", [(336, 150, 3), (336, 150, 56), (336, 150, 57), (336, 150, 60), (336, 150, 61), (336, 150, 184),
                                   (336, 150, 185), (336, 150, 188), (336, 150, 189), (336, 150, 191), (336, 150, 192),
                                   (72, 16400, 1), (72, 16400, 57), (16400, 72, 2), (16400, 72, 61), (16401, 72, 30), (333, 150, 62)])
def test_box_blur_ring_limits(ctx, oracle_mod, w, h, r):
    rng = np.random.default_rng(w + 7 * r)
    p = rng.random((h, w)).astype(np.float32)
    got, want = ctx.box_blur(p, r), oracle_mod.box_blur(p, r)
    assert np.array_equal(got, want), f"bit-equal {np.mean(got == want):.6f}"


# Ring passes split the rows (columns) x planes into bands of 32..64 rows
# (columns) per block so that the busiest CU's share is smallest.  On a
# 256-CU device these shapes select row bands 64, 48 and 32 and column bands
# 64, 56, 52, 48, 44 and 40 (16-byte and 4-byte column loads); the result
# must not depend on the band.
@pytest.mark.parametrize("w,h,r", [(64, 16384, 3), (68, 12288, 17), (16384, 40, 5), (14336, 36, 21), (13312, 20, 2),
                                   (12288, 33, 9), (10240, 24, 60), (11264, 50, 7), (14334, 30, 6)])
def test_box_blur_bands(ctx, oracle_mod, w, h, r):
    rng = np.random.default_rng(w + h + r)
    p = rng.random((h, w)).astype(np.float32)
    got, want = ctx.box_blur(p, r), oracle_mod.box_blur(p, r)
    assert np.array_equal(got, want), f"bit-equal {np.mean(got == want):.6f}"


def test_guided_filter_bands(ctx, oracle_mod):
    # four-plane row passes over 3072 rows take 48-row bands
    rng = np.random.default_rng(5)
    I = rng.random((3072, 96)).astype(np.float32)
    p = (0.5 * I + 0.3 * rng.random((3072, 96))).astype(np.float32)
    assert np.array_equal(ctx.guided_filter(I, p, 6, 1e-3), oracle_mod.guided_filter(I, p, 6, 1e-3))


@pytest.mark.parametrize("sigma", [0.2, 0.5, 0.75, 1.0, 1.9, 2.5, 7.0])
def test_gaussian_blur_exact(ctx, oracle_mod, sigma):
    rng = np.random.default_rng(int(sigma * 10))
    p = rng.random((97, 210)).astype(np.float32)
    assert np.array_equal(ctx.gaussian_blur(p, sigma), oracle_mod.gaussian_blur(p, sigma))


def test_guided_filter_exact(ctx, oracle_mod):
    rng = np.random.default_rng(2)
    I = rng.random((120, 230)).astype(np.float32)
    p = (0.4 * I + 0.2 * rng.random((120, 230))).astype(np.float32)
    for r, eps in ((2, 1e-4), (10, 1e-3), (1, 0.0)):
        assert np.array_equal(ctx.guided_filter(I, p, r, eps), oracle_mod.guided_filter(I, p, r, eps))


def test_oklab_and_luminance(ctx, oracle_mod):
    O = oracle_mod
    rgb = rgb_frame(O, 600, 160, 120)
    assert np.array_equal(ctx.oklab_from_prophoto(rgb), O.oklab_from_prophoto(rgb))
    lab = O.oklab_from_prophoto(rgb)
    assert np.array_equal(ctx.oklab_to_prophoto(lab), O.oklab_to_prophoto(lab))
    assert np.array_equal(ctx.luminance(rgb), O.luminance(rgb))


@pytest.mark.parametrize("sigma,iters", [(1.0, 20), (0.75, 8), (0.5, 3), (1.9, 4), (0.2, 5), (2.5, 3), (1.0, 0)])
def test_raw_details(ctx, oracle_mod, sigma, iters):
    O = oracle_mod
    rgb = rgb_frame(O, 500, 200, 150)
    got = ctx.raw_details(rgb, sigma, iters)
    want = O.raw_details(rgb, sigma, iters)
    err = maxabs(got, want)
    say(f"raw_details sigma={sigma} iters={iters} max-abs {err:.3g} bit-equal {np.mean(got == want):.6f}")
    assert err <= TOL


def test_raw_details_first_and_last_launch(ctx, oracle_mod):
    """Bit for bit against the oracle: one iteration (a launch that is both
    the first and the last) and two, one iteration in place through the
    device entry point, and y beyond the fast division's range (> 2^64) in
    an interior and an edge tile of a 300x260 frame, after which the later
    iterations must keep the IEEE division."""
    import darkslide_amd as D
    O = oracle_mod
    w, h = 300, 260
    rgb = rgb_frame(O, 501, w, h)
    big = rgb.copy()
    big[100, 100] = (3e19, 3e19, 3e19)
    big[5, 290, 1] = 1e30
    big[150, 150] = 0.0
    for img, iters in ((rgb, 1), (rgb, 2), (big, 1), (big, 3)):
        got = ctx.raw_details(img, 1.0, iters)
        want = O.raw_details(img, 1.0, iters)
        say(f"raw_details first launch: iters={iters} huge={img is big} bit-equal {np.mean(got == want):.6f}")
        assert np.array_equal(got, want)
    L = D.lib()
    buf = ctx.alloc(w * h * 12)
    try:
        buf.upload(rgb)
        ctx._check(L.ds_raw_details_dev(ctx.handle, buf.ptr, w, h, 1.0, 1, buf.ptr))
        got = buf.download((h, w, 3), np.float32)
    finally:
        buf.free()
    assert np.array_equal(got, O.raw_details(rgb, 1.0, 1))


def test_pipeline_config1_with_fp64_error(ctx, oracle_mod):
    """Demosaic -> denoise (NLM 7x7 patch, 21x21 search, h = 0.6 sigma) ->
    RL (sigma 1, 20 iterations) at 1024x1024: each stage against the FP32
    restatement (gate) and the FP64 oracle (reported)."""
    O = oracle_mod
    raw = synth(0x5EED0000, 1024, 1024)
    rgb = ctx.demosaic_u16(raw, clip=False)
    cfa = O.cfa_normalize_u16(raw, RGGB, 512.0, list(WHITE))
    ref = O.rcd_bayer(cfa)
    ref64 = O.rcd_bayer(cfa.astype(np.float64))
    say(f"stage demosaic: vs fp32 {maxabs(rgb, ref):.3g}  vs fp64 {maxabs(rgb, ref64):.3g}")
    assert maxabs(rgb, ref) <= TOL
    dn = ctx.denoise(ref, 0.0, 10, 3)
    dref = O.denoise(ref, 0.0, 10, 3)
    dref64 = O.denoise(ref.astype(np.float64), 0.0, 10, 3)
    env = max(maxabs(O.denoise(ref, 0.0, 10, 3, nthreads=t), dref) for t in (1, 4, 16))
    say(f"stage denoise: vs fp32 {maxabs(dn, dref):.3g} (thread envelope {env:.3g})  vs fp64 {maxabs(dn, dref64):.3g} "
        f"(fp32 reference vs fp64 {maxabs(dref, dref64):.3g})")
    assert maxabs(dn, dref) <= max(TOL, 1.5 * env)
    rl = ctx.raw_details(dref, 1.0, 20)
    rref = O.raw_details(dref, 1.0, 20)
    rref64 = O.raw_details(dref.astype(np.float64), 1.0, 20)
    say(f"stage raw_details: vs fp32 {maxabs(rl, rref):.3g}  vs fp64 {maxabs(rl, rref64):.3g}")
    assert maxabs(rl, rref) <= TOL


# ------------------------------------------------------------------ 6000 x 4000

@pytest.fixture(scope="module")
def frame24(ctx):
    import darkslide_amd as D
    w, h = 6000, 4000
    buf = ctx.alloc(w * h * 2)
    ctx._check(D.lib().ds_synth_bayer_dev(ctx.handle, 0x5EED0001, w, h, 14, 1.0, D._pat(RGGB), buf.ptr))
    raw = buf.download((h, w), np.uint16)
    buf.free()
    return raw


def test_full_frame_demosaic(ctx, oracle_mod, frame24):
    O = oracle_mod
    raw = frame24
    rgb, clip = ctx.demosaic_u16(raw)
    want, wclip, _ = oracle_demosaic(O, raw, RGGB, (1, 1, 1), 0)
    err = maxabs(rgb, want)
    say(f"rcd 6000x4000 max-abs {err:.3g}, bit-equal {np.mean(rgb == want):.7f}")
    assert err <= TOL
    assert np.array_equal(clip, wclip)
    mhc = ctx.demosaic_u16(raw, algo=1, clip=False)
    assert maxabs(mhc, oracle_demosaic(O, raw, RGGB, (1, 1, 1), 1)[0]) <= TOL


CROPS = [(0, 0), (2000, 1700), (5744, 3744), (0, 3000)]  # (x0, y0) of 256x256 crops


def crop_check(full, fn, x0, y0, size, margin, w, h):
    """Runs fn on the crop grown by margin (clipped to the frame) and compares
    the crop's interior pixels with the full-frame result."""
    cx0, cy0 = max(x0 - margin, 0), max(y0 - margin, 0)
    cx1, cy1 = min(x0 + size + margin, w), min(y0 + size + margin, h)
    out = fn(cx0, cy0, cx1, cy1)
    sub = out[y0 - cy0:y0 - cy0 + size, x0 - cx0:x0 - cx0 + size]
    return maxabs(full[y0:y0 + size, x0:x0 + size], sub)


def test_full_frame_nlm_and_rl(ctx, oracle_mod, frame24):
    O = oracle_mod
    h, w = frame24.shape
    rgb = ctx.demosaic_u16(frame24, clip=False)
    L = np.ascontiguousarray(O.oklab_from_prophoto(rgb)[..., 0])
    s = ctx.estimate_noise(L)
    assert s == O.estimate_noise(L)
    hp = max(s, 1e-4) * 0.6
    lf = ctx.nlm(L, hp, 10, 3)
    for (x0, y0) in CROPS:
        err = crop_check(lf, lambda a, b, c, d: O.nlm(np.ascontiguousarray(L[b:d, a:c]), hp, 10, 3), x0, y0, 256, 14, w, h)
        say(f"nlm 6000x4000 crop ({x0},{y0}) max-abs {err:.3g}")
        assert err <= TOL
    out = ctx.raw_details(rgb, 1.0, 20)
    for (x0, y0) in CROPS:
        err = crop_check(out, lambda a, b, c, d: O.raw_details(np.ascontiguousarray(rgb[b:d, a:c]), 1.0, 20), x0, y0, 256, 124, w, h)
        say(f"raw_details 6000x4000 crop ({x0},{y0}) max-abs {err:.3g}")
        assert err <= TOL
    # RL is flux-preserving up to its clamps: the mean luminance barely moves
    y_in = rgb @ np.array([0.2880402, 0.7118741, 0.0000857], np.float32)
    y_out = out @ np.array([0.2880402, 0.7118741, 0.0000857], np.float32)
    assert abs(float(y_out.mean()) / float(y_in.mean()) - 1) < 2e-3


def test_full_frame_denoise_is_its_stages(ctx, frame24):
    """At 6000x4000 the guided pair's six- and four-plane box passes run in
    48-row / 48-column bands, the single guided filter's two- and four-plane
    passes in other bands; the running sums, and so denoise, must not depend
    on the band: bit-equal to the GPU's stages run one by one."""
    h, w = frame24.shape
    rgb = ctx.demosaic_u16(frame24, clip=False)
    lab = ctx.oklab_from_prophoto(rgb)
    L, A, B = (np.ascontiguousarray(lab[..., i]) for i in range(3))
    sigma = np.float32(max(np.float32(ctx.estimate_noise(L)), np.float32(1e-4)))
    k = np.float32(0)
    hp = np.float32(sigma * (np.float32(0.6) + np.float32(1.6) * k))
    lf = ctx.nlm(L, hp, 10, 3)
    l2 = (L + (lf - L) * (np.float32(0.5) + np.float32(0.5) * k)).astype(np.float32)
    r = int(max((np.float32(w) / np.float32(600)) * (np.float32(1) + np.float32(2) * k), np.float32(2)))
    eps = np.float32(sigma * np.float32(2)) ** 2
    a2, b2 = ctx.guided_filter(l2, A, r, eps), ctx.guided_filter(l2, B, r, eps)
    want = np.maximum(ctx.oklab_to_prophoto(np.stack([l2, a2, b2], -1).astype(np.float32)), np.float32(0))
    got = ctx.denoise(rgb, 0.0, 10, 3)
    assert np.array_equal(got, want), f"bit-equal {np.mean(got == want):.7f}"


def test_pinned_host_buffers_round_trip(ctx):
    """ds_host_alloc: a frame through pinned host memory and back, unchanged."""
    import ctypes
    import darkslide_amd as D
    L = D.lib()
    n = 1 << 20
    data = np.random.default_rng(3).random(n).astype(np.float32)
    dev = ctx.alloc(n * 4)
    dev.upload(data)
    hp = ctypes.c_void_p()
    ctx._check(L.ds_host_alloc(ctx.handle, n * 4, ctypes.byref(hp)))
    try:
        ctx._check(L.ds_memcpy_d2h(ctx.handle, hp, dev.ptr, n * 4))
        ctx.sync()
        host = np.ctypeslib.as_array((ctypes.c_float * n).from_address(hp.value))
        assert np.array_equal(host, data)
        host[:] = data * 2.0
        dev2 = ctx.alloc(n * 4)
        ctx._check(L.ds_memcpy_h2d(ctx.handle, dev2.ptr, hp, n * 4))
        ctx.sync()
        assert np.array_equal(dev2.download((n,), np.float32), data * 2.0)
        dev2.free()
    finally:
        ctx._check(L.ds_host_free(ctx.handle, hp))
    with pytest.raises(D.DarkslideError):
        ctx._check(L.ds_host_free(ctx.handle, hp))
    dev.free()
