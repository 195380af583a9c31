"""The learned and batched configurations at their own sizes, and
Engine::base_for's RL branch.

* Learned inference on 6000x4000 and 8256x5504 frames: the network's
  luminance (ds_cnn_denoise_plane, the fused BF16 MFMA kernel) against the
  oracle's bf16-emulating FP64 forward on crops at the frame's corners and
  centre, each crop widened by the network's 8-pixel receptive field on its
  inner sides, so the crop's interior equals the full-frame result; the
  learned denoise with every region sent back to NLM equals the plain
  denoise bit for bit; the trained network's region count and fallbacks.
* A batch of 12, 45 and 61 MP frames (12/14-bit, noise 1-16x) through the
  captured graph: each frame equals the single-frame chain bit for bit
  (plain and learned), and the demosaic-only and demosaic -> RL batches
  equal the oracle on crops with the stages' margins (demosaic 10 pixels,
  RL 20 iterations x 2 blurs x 3 = 120 pixels).
* reference_gating (pipeline/mod.rs:229-252): a frame whose long edge
  exceeds 3200 pixels gets raw_details(0.75, 8), a shorter one none, and
  the denoise amount is raised to 1.
* Batch graphs and the networks they use: a graph captured with a network
  that is then destroyed is never replayed for another network (even one
  at the same address); an oversized learned frame is refused before any
  capture, and the context keeps working.
"""
import os
import sys

import numpy as np
import pytest

from conftest import RGGB

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CKPT = os.path.join(ROOT, "darkslide_amd", "weights", "cnn_l32x8.dsck")
RF = 8  # receptive field of 8 conv3x3 layers


def say(*a):
    print(*a, file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def trained(ctx):
    import darkslide_amd as D
    net = D.CNN(ctx, seed=1)
    net.load(CKPT)
    yield net
    net.close()


def crop_windows(w, h, c=96):
    return [(0, 0), ((w - c) // 2, (h - c) // 2), (w - c, h - c), (w - c, 0)], c


@pytest.mark.parametrize("w,h", [(6000, 4000), (8256, 5504)])
def test_learned_plane_at_config_size(ctx, oracle_mod, trained, w, h):
    O = oracle_mod
    say(f"learned luminance {w}x{h}: synth, demosaic, Oklab")
    import darkslide_amd as D
    raw = D.synth_bayer(0x5EED0100 + w, w, h, 14, 4.0)
    rgb = ctx.demosaic_u16(raw, clip=False)
    L = np.ascontiguousarray(ctx.oklab_from_prophoto(rgb)[..., 0])
    del rgb
    sigma = ctx.estimate_noise(L)
    s = max(np.float32(sigma), np.float32(1e-4))
    lf = trained.denoise_plane(L)
    assert lf.shape == (h, w) and np.isfinite(lf).all()
    x = O.cnn_input(L, float(s))
    params = trained.params().astype(np.float64)
    wins, c = crop_windows(w, h)
    for (ox, oy) in wins:
        x0, y0 = max(ox - RF, 0), max(oy - RF, 0)
        x1, y1 = min(ox + c + RF, w), min(oy + c + RF, h)
        xc = np.ascontiguousarray(x[y0:y1, x0:x1].astype(np.float32))
        ref = O.cnn_forward(params, xc[None], bf16=True)[0][oy - y0:oy - y0 + c, ox - x0:ox - x0 + c]
        xin = x[oy:oy + c, ox:ox + c]
        got = xin + (lf[oy:oy + c, ox:ox + c].astype(np.float64) - L[oy:oy + c, ox:ox + c]) / float(s)
        scale = np.abs(ref - xin).max()
        err = np.abs(got - ref)
        say(f"  crop ({ox},{oy}) {c}x{c}: max {err.max():.3e} rms {np.sqrt((err ** 2).mean()):.3e} (residual scale {scale:.3f})")
        assert scale > 0.05
        assert err.max() <= 1e-2 * scale
        assert np.sqrt((err ** 2).mean()) <= 1e-3 * scale


@pytest.mark.parametrize("w,h", [(6000, 4000), (8256, 5504)])
def test_learned_denoise_at_config_size(ctx, trained, w, h):
    """Forced fallback (threshold 0) equals the plain denoise bit for bit at
    full size; at 0.3 dB every 256x256 region is counted and few fall back."""
    import darkslide_amd as D
    say(f"learned denoise {w}x{h}")
    raw = D.synth_bayer(0x5EED0200 + w, w, h, 14, 8.0)
    rgb = ctx.demosaic_u16(raw, clip=False)
    plain = ctx.denoise(rgb, 0.0, 10, 3)
    forced, st0 = ctx.denoise_learned(trained, rgb, 0.0, 10, 3, threshold_db=0.0)
    regions = ((w + 255) // 256) * ((h + 255) // 256)
    assert st0["regions"] == regions and st0["fallback"] == regions
    assert np.array_equal(forced, plain)
    learned, st = ctx.denoise_learned(trained, rgb, 0.0, 10, 3, threshold_db=0.3)
    say(f"  regions {st['regions']}, fallback {st['fallback']}, mean gap {st['mean_gap_db']:.4f} dB, max {st['max_gap_db']:.4f} dB")
    assert st["regions"] == regions
    assert st["fallback"] <= 0.05 * regions
    assert np.isfinite(learned).all()
    assert 0.0 <= st["mean_gap_db"] <= 0.3


# ---------------------------------------------------------------- batch at config sizes

C5_FRAMES = [  # w, h, bits, iso, pattern, seed
    (4256, 2848, 12, 16.0, (0, 1, 1, 2), 0x5EED0300),
    (8256, 5504, 14, 4.0, (2, 1, 1, 0), 0x5EED0301),
    (9504, 6336, 12, 1.0, (0, 1, 1, 2), 0x5EED0302),
]


@pytest.fixture(scope="module")
def big_batch(ctx):
    from test_gpu_batch import Batch
    say("allocating the 12 / 45 / 61 MP batch")
    b = Batch(ctx, C5_FRAMES)
    yield b
    b.free()


def test_config_batch_matches_single_frames(ctx, big_batch):
    from test_gpu_batch import single
    say("batch 12/45/61 MP, plain")
    big_batch.clear_outputs()
    info = ctx.pipeline_batch(big_batch.frames, streams=2)
    say(f"  {info}")
    ref = single(ctx, big_batch)
    for k, (got, r) in enumerate(zip(big_batch.outputs(), ref)):
        assert np.array_equal(got, r), k


def test_config_batch_learned_matches_single_frames(ctx, big_batch, trained):
    import darkslide_amd as D
    from test_gpu_batch import single
    say("batch 12/45/61 MP, learned")
    big_batch.clear_outputs()
    info = ctx.pipeline_batch(big_batch.frames, denoise=D.DENOISE_LEARNED, net=trained, threshold_db=0.3, streams=2)
    say(f"  {info}")
    regions = sum(((w + 255) // 256) * ((h + 255) // 256) for (w, h, *_) in C5_FRAMES)
    assert info["regions"] == regions
    ref = single(ctx, big_batch, denoise="learned", net=trained)
    for k, (got, r) in enumerate(zip(big_batch.outputs(), ref)):
        assert np.array_equal(got, r), k


def _raw_of(big_batch, k):
    raw_buf, _, w, h, black, white, pat = big_batch.bufs[k]
    return raw_buf.download((h, w), np.uint16), w, h, black, white, pat


@pytest.mark.parametrize("rl", [False, True])
def test_config_batch_against_oracle_crops(ctx, oracle_mod, big_batch, rl):
    """demosaic-only (and demosaic -> RL 1.0 x 20) batches: the oracle on
    crops with the stages' margins reproduces every crop bit for bit."""
    import darkslide_amd as D
    O = oracle_mod
    say(f"batch 12/45/61 MP against oracle crops, rl={rl}")
    big_batch.clear_outputs()
    ctx.pipeline_batch(big_batch.frames, denoise=D.DENOISE_NONE, raw_details=rl, rl_sigma=1.0, rl_iters=20, streams=2)
    outs = big_batch.outputs()
    margin = 10 + (120 if rl else 0)
    for k in range(len(C5_FRAMES)):
        raw, w, h, black, white, pat = _raw_of(big_batch, k)
        wins, c = crop_windows(w, h, 128)
        for (ox, oy) in wins:
            x0, y0 = max(ox - margin, 0), max(oy - margin, 0)
            x1, y1 = min(ox + c + margin, w), min(oy + c + margin, h)
            if (x0 % 2) != 0 or (y0 % 2) != 0:  # keep the crop on the frame's CFA phase
                x0 -= x0 % 2
                y0 -= y0 % 2
            cfa = O.cfa_normalize_u16(np.ascontiguousarray(raw[y0:y1, x0:x1]), pat, black, [white] * 3)
            ref = O.rcd_bayer(cfa, pat)
            if rl:
                ref = O.raw_details(ref, 1.0, 20)
            ref = ref[oy - y0:oy - y0 + c, ox - x0:ox - x0 + c]
            got = outs[k][oy:oy + c, ox:ox + c]
            assert np.array_equal(got, ref), (k, ox, oy, float(np.abs(got - ref).max()))


# ---------------------------------------------------------------- Engine::base_for gating

def test_reference_gating_rl_branch(ctx, oracle_mod):
    """A 3300 x 24 frame (long edge > 3200) gets raw_details(0.75, 8); a
    200 x 150 one none (pipeline/mod.rs:235, 244-246)."""
    import darkslide_amd as D
    from test_gpu_batch import Batch, single
    O = oracle_mod
    specs = [(3300, 24, 14, 4.0, (0, 1, 1, 2), 0x5EED0400), (200, 150, 14, 4.0, (0, 1, 1, 2), 0x5EED0401)]
    b = Batch(ctx, specs)
    try:
        # without denoise: the oracle's demosaic -> RL (0.75, 8) for the long frame, demosaic alone for the short one
        b.clear_outputs()
        ctx.pipeline_batch(b.frames, denoise=D.DENOISE_NONE, raw_details=True, rl_sigma=1.0, rl_iters=20, reference_gating=True,
                           streams=2)
        outs = b.outputs()
        for k, rl in ((0, True), (1, False)):
            raw_buf, _, w, h, black, white, pat = b.bufs[k]
            cfa = O.cfa_normalize_u16(raw_buf.download((h, w), np.uint16), pat, black, [white] * 3)
            ref = O.rcd_bayer(cfa, pat)
            if rl:
                ref = O.raw_details(ref, 0.75, 8)
            assert np.array_equal(outs[k], ref), k
        # with NLM: denoise at amount max(0, 1) = 1, then the gated RL
        b.clear_outputs()
        ctx.pipeline_batch(b.frames, amount=0.0, raw_details=True, reference_gating=True, streams=2)
        outs = b.outputs()
        long_ref = single(ctx, b, amount=1.0, rl=(0.75, 8))[0]
        short_ref = single(ctx, b, amount=1.0, rl=None)[1]
        assert np.array_equal(outs[0], long_ref)
        assert np.array_equal(outs[1], short_ref)
    finally:
        b.free()


# ---------------------------------------------------------------- graphs and networks

def test_graph_not_replayed_for_a_new_network(ctx):
    import darkslide_amd as D
    from test_gpu_batch import FRAMES, Batch, single
    b = Batch(ctx, FRAMES[:3])
    try:
        outs = []
        for seed in (5, 6):
            net = D.CNN(ctx, seed=seed)
            b.clear_outputs()
            # an infinite threshold keeps the network's result in every region
            ctx.pipeline_batch(b.frames, denoise=D.DENOISE_LEARNED, net=net, threshold_db=float("inf"), streams=2)
            got = b.outputs()
            ref = single(ctx, b, denoise="learned", net=net, thr=float("inf"))
            for k, (g, r) in enumerate(zip(got, ref)):
                assert np.array_equal(g, r), (seed, k)
            outs.append(got)
            net.close()
        assert not all(np.array_equal(a, c) for a, c in zip(*outs))
    finally:
        b.free()


def test_oversized_learned_frame_refused_before_capture(ctx):
    import darkslide_amd as D
    from test_gpu_batch import FRAMES, Batch, single
    b = Batch(ctx, FRAMES[:2])
    big = Batch(ctx, [(65537, 16, 14, 1.0, (0, 1, 1, 2), 77)])
    net = D.CNN(ctx, seed=3)
    try:
        with pytest.raises(D.DarkslideError, match="frame 1"):
            ctx.pipeline_batch([b.frames[0], big.frames[0]], denoise=D.DENOISE_LEARNED, net=net, streams=2)
        b.clear_outputs()
        ctx.pipeline_batch(b.frames, streams=2)
        for k, (g, r) in enumerate(zip(b.outputs(), single(ctx, b))):
            assert np.array_equal(g, r), k
    finally:
        net.close()
        big.free()
        b.free()
