"""The detail stages (pipeline/detail.rs) on the GPU against the oracle.

Bit-exact: Plane::resize and fast_guided_filter (the box passes and the
coefficient arithmetic are the reference's), and noise_reduction (its Oklab
forward uses the glibc-exact cube root of csrc/libm_ref.h).  sharpen uses
the device powf in its sRGB curves, within an ulp or two of glibc's, so it
is held to max-abs 2e-5 (the path's gate) and to the FP64 oracle at the
FP32 oracle's own distance.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 2e-5


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def scene(h, w, noise=0.02, seed=0):
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:h, 0:w]
    base = (0.3 + 0.2 * np.sin(xx / 9.0) * np.cos(yy / 7.0))[..., None] * np.array([1.0, 0.9, 0.7])
    return (base + rng.normal(0, noise, base.shape)).clip(0, None).astype(np.float32)


@pytest.mark.parametrize("shape,to", [((40, 60), (15, 10)), ((97, 131), (23, 17)), ((64, 80), (32, 20)), ((16, 12), (40, 33)),
                                      ((100, 100), (100, 100)), ((33, 47), (70, 51))])
def test_plane_resize_exact(ctx, oracle_mod, shape, to):
    rng = np.random.default_rng(shape[0])
    p = rng.random(shape).astype(np.float32)
    assert np.array_equal(ctx.plane_resize(p, *to), oracle_mod.plane_resize(p, *to))


@pytest.mark.parametrize("h,w,r,sub", [(64, 80, 6, 2), (97, 131, 9, 4), (120, 160, 3, 2), (20, 20, 3, 4), (90, 70, 5, 1), (256, 200, 17, 4)])
def test_fast_guided_filter_exact(ctx, oracle_mod, h, w, r, sub):
    rng = np.random.default_rng(h + w)
    I = rng.random((h, w)).astype(np.float32)
    p = (0.5 * I + 0.3 * rng.random((h, w))).astype(np.float32)
    got = ctx.fast_guided_filter(I, p, r, 3e-4, sub)
    want = oracle_mod.fast_guided_filter(I, p, r, 3e-4, sub)
    assert np.array_equal(got, want), np.abs(got - want).max()


NR_CASES = [
    (50, 50, 0, 0, 50, 50),    # luminance only
    (0, 50, 0, 25, 50, 50),    # colour only (the reference's defaults)
    (60, 30, 40, 40, 20, 80),  # both, with contrast
    (20, 80, 100, 100, 0, 100),
]


@pytest.mark.parametrize("nr", NR_CASES)
@pytest.mark.parametrize("h,w,scale", [(96, 128, 1.0), (200, 300, 0.5), (150, 100, 2.0)])
def test_noise_reduction(ctx, oracle_mod, nr, h, w, scale):
    """Every step is bit-exact, the Oklab forward included (its cube root
    returns glibc's bits), so the whole stage is.  The FP64 distance is
    reported."""
    O = oracle_mod
    img = scene(h, w, seed=h)
    got = ctx.noise_reduction(img, nr, scale=scale)
    want = O.noise_reduction(img, nr, scale=scale)
    f64 = O.noise_reduction(img.astype(np.float64), nr, scale=scale)
    err, e64 = [float(np.abs(a - b).max()) for a, b in ((got, want), (got, f64))]
    say(f"NR {nr} {w}x{h} scale {scale}: vs fp32 {err:.3g}, vs fp64 {e64:.3g}")
    assert np.abs(want - img).max() > 1e-3
    assert np.array_equal(got, want), f"bit-equal {np.mean(got == want):.6f} max-abs {err:.3g}"


def test_noise_reduction_with_field_and_identities(ctx, oracle_mod):
    O = oracle_mod
    img = scene(96, 128, seed=7)
    field = np.zeros((96, 128), np.float32)
    field[20:70, 10:90] = 60.0
    for nr in [(0, 40, 30, 0, 50, 50), (25, 40, 30, 20, 50, 50)]:
        got = ctx.noise_reduction(img, nr, field=field)
        want = O.noise_reduction(img, nr, field=field)
        assert np.array_equal(got, want), np.abs(got - want).max()
    assert np.array_equal(ctx.noise_reduction(img, (0, 50, 0, 0, 50, 50)), img)
    assert np.array_equal(ctx.noise_reduction(img, (60, 50, 0, 40, 50, 50), scale=0.3), img)


SH_CASES = [(40, 1.0, 25, 0), (80, 2.5, 10, 40), (40, 0.5, 100, 0), (150, 1.2, 0, 100)]


@pytest.mark.parametrize("sh", SH_CASES)
@pytest.mark.parametrize("h,w,scale", [(96, 128, 1.0), (200, 300, 0.6), (150, 100, 1.5)])
def test_sharpen(ctx, oracle_mod, sh, h, w, scale):
    O = oracle_mod
    img = scene(h, w, seed=w)
    got = ctx.sharpen(img, sh, scale=scale)
    want = O.sharpen(img, sh, scale=scale)
    f64 = O.sharpen(img.astype(np.float64), sh, scale=scale)
    err, e64, r64 = [float(np.abs(a - b).max()) for a, b in ((got, want), (got, f64), (want, f64))]
    say(f"sharpen {sh} {w}x{h} scale {scale}: vs fp32 {err:.3g}, vs fp64 {e64:.3g} (fp32 oracle vs fp64 {r64:.3g})")
    assert not np.array_equal(want, img)
    assert err <= TOL and e64 <= max(TOL, 1.5 * r64)


def test_sharpen_field_and_identities(ctx, oracle_mod):
    O = oracle_mod
    img = scene(64, 96, seed=3)
    field = np.zeros((64, 96), np.float32)
    field[:, 40:] = 50.0
    assert np.abs(ctx.sharpen(img, (0, 1, 25, 20), field=field) - O.sharpen(img, (0, 1, 25, 20), field=field)).max() <= TOL
    assert np.array_equal(ctx.sharpen(img, (0, 1, 25, 0)), img)
    assert np.array_equal(ctx.sharpen(img, (40, 1, 25, 0), scale=0.2), img)
    flat = np.full((32, 32, 3), 0.3, np.float32)
    assert np.array_equal(ctx.sharpen(flat, (90, 2, 25, 0)), flat)


def test_detail_rejects_bad_input(ctx):
    import ctypes
    import darkslide_amd as D
    with pytest.raises(D.DarkslideError):
        ctx._check(D.lib().ds_noise_reduction(ctx.handle, None, 4, 4, None, None, ctypes.c_float(1.0)))
    with pytest.raises(D.DarkslideError):
        ctx.plane_resize(np.zeros((0, 4), np.float32), 2, 2)


@pytest.mark.parametrize("h,w", [(96, 128), (37, 211), (300, 200), (1, 9)])
def test_super_res_refine_exact(ctx, oracle_mod, h, w):
    """sqrt, the exact gaussian (sigma 1.2), IEEE division and clamps: bit for bit."""
    img = scene(h, w, seed=h + 1)
    img[h // 2, w // 3] = 0.0025  # a pixel whose ratio clamps at 2
    got = ctx.super_res_refine(img)
    want = oracle_mod.super_res_refine(img)
    assert not np.array_equal(want, img)
    assert np.array_equal(got, want), np.abs(got - want).max()


@pytest.mark.parametrize("mode", [1, 2, 3])
@pytest.mark.parametrize("amt", [0, 1, 2])
def test_output_sharpen(ctx, oracle_mod, mode, amt):
    """The tone curves use powf (device vs glibc: an ulp or two): 2e-5."""
    O = oracle_mod
    img = scene(120, 90, seed=mode * 3 + amt)
    img[:, 45:] *= 2.5  # values above 1 take the curves' linear extension
    got = ctx.output_sharpen(img, mode, amt)
    want = O.output_sharpen(img, mode, amt)
    f64 = O.output_sharpen(img.astype(np.float64), mode, amt)
    err, e64, r64 = [float(np.abs(a - b).max()) for a, b in ((got, want), (got, f64), (want, f64))]
    say(f"output_sharpen mode {mode} amount {amt}: vs fp32 {err:.3g}, vs fp64 {e64:.3g} (fp32 oracle vs fp64 {r64:.3g})")
    assert not np.array_equal(want, img)
    assert err <= TOL and e64 <= max(TOL, 1.5 * r64)


def test_export_post_steps_identities_and_errors(ctx):
    import darkslide_amd as D
    img = scene(32, 48, seed=9)
    assert np.array_equal(ctx.output_sharpen(img, D.SHARPEN_FOR_NONE, D.SHARPEN_AMOUNT_HIGH), img)
    with pytest.raises(D.DarkslideError):
        ctx.output_sharpen(img, 4, 1)
    with pytest.raises(D.DarkslideError):
        ctx.output_sharpen(img, 1, 3)
    assert ctx.super_res_refine(np.zeros((0, 5, 3), np.float32)).shape == (0, 5, 3)
