"""Batched export through one captured graph, the tile autotuner and the
FP64 white-balance statistics, on the GPU.

The batch must reproduce, bit for bit, what the single-frame entry points
give for each frame (demosaic -> denoise -> RL), for any branch count, on
graph replay, after workspace growth, with autotuned tiles and with the
learned denoiser; every tile variant must give the same result.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FRAMES = [  # w, h, bits, iso, pattern, seed
    (200, 150, 14, 1.0, (0, 1, 1, 2), 11),
    (333, 257, 12, 4.0, (2, 1, 1, 0), 12),
    (130, 99, 14, 16.0, (1, 0, 2, 1), 13),
    (64, 64, 12, 2.0, (1, 2, 0, 1), 14),
    (520, 300, 14, 8.0, (0, 1, 1, 2), 15),
]


def say(*a):
    print(*a, file=sys.stderr, flush=True)


class Batch:
    def __init__(self, ctx, specs):
        import darkslide_amd as D
        self.ctx = ctx
        self.frames = []
        self.bufs = []
        L = D.lib()
        for (w, h, bits, iso, pat, seed) in specs:
            raw = ctx.alloc(w * h * 2)
            out = ctx.alloc(w * h * 12)
            ctx._check(L.ds_synth_bayer_dev(ctx.handle, seed, w, h, bits, iso, D._pat(pat), raw.ptr))
            black, white = (512.0, 16383.0) if bits == 14 else (128.0, 4095.0)
            f = D.Frame(raw.ptr.value, w, h, w, (ctypes_u8(pat)), black, (white,) * 3, (1.0, 1.0, 1.0), out.ptr.value)
            self.frames.append(f)
            self.bufs.append((raw, out, w, h, black, white, pat))

    def outputs(self):
        return [out.download((h, w, 3), np.float32) for (_, out, w, h, _, _, _) in self.bufs]

    def clear_outputs(self):
        for (_, out, w, h, _, _, _) in self.bufs:
            out.upload(np.full((h, w, 3), np.nan, np.float32))

    def free(self):
        for (raw, out, *_) in self.bufs:
            raw.free()
            out.free()


def ctypes_u8(pat):
    import ctypes
    return (ctypes.c_uint8 * 4)(*pat)


def single(ctx, b, denoise="nlm", net=None, amount=0.0, search=10, patch=3, rl=(1.0, 20), thr=0.3):
    """The frame's result through the single-frame entry points."""
    import darkslide_amd as D
    L = D.lib()
    res = []
    for (raw, _, w, h, black, white, pat) in b.bufs:
        rgb, dn, out = ctx.alloc(w * h * 12), ctx.alloc(w * h * 12), ctx.alloc(w * h * 12)
        ctx._check(L.ds_demosaic_u16_dev(ctx.handle, raw.ptr, w, h, w, D._pat(pat), black, D._f3((white,) * 3), D._f3((1, 1, 1)),
                                         D.ALGO_RCD, rgb.ptr, None))
        cur = rgb
        if denoise == "nlm":
            ctx._check(L.ds_denoise_ex_dev(ctx.handle, rgb.ptr, w, h, amount, search, patch, dn.ptr))
            cur = dn
        elif denoise == "learned":
            ctx._check(L.ds_denoise_learned_dev(ctx.handle, net.handle, rgb.ptr, w, h, amount, search, patch, thr, dn.ptr, None))
            cur = dn
        if rl is not None:
            ctx._check(L.ds_raw_details_dev(ctx.handle, cur.ptr, w, h, rl[0], rl[1], out.ptr))
            cur = out
        res.append(cur.download((h, w, 3), np.float32))
        for bb in (rgb, dn, out):
            bb.free()
    return res


@pytest.fixture(scope="module")
def batch(ctx):
    b = Batch(ctx, FRAMES)
    yield b
    b.free()


@pytest.fixture(scope="module")
def expected_plain(ctx, batch):
    return single(ctx, batch)


@pytest.mark.parametrize("streams", [1, 2, 3])
def test_batch_matches_single_frames(ctx, batch, expected_plain, streams):
    say("batch plain, streams", streams)
    batch.clear_outputs()
    info = ctx.pipeline_batch(batch.frames, streams=streams)
    say(f"  {info}")
    assert info["frames"] == len(FRAMES) and info["branches"] == streams
    for k, (got, ref) in enumerate(zip(batch.outputs(), expected_plain)):
        assert np.array_equal(got, ref), k


def test_batch_graph_replay(ctx, batch, expected_plain):
    """Replays of the cached graph, with eager work between them, give the same result."""
    say("batch replay")
    ctx.pipeline_batch(batch.frames, streams=2)
    for rep in range(3):
        batch.clear_outputs()
        ctx.nlm(np.random.default_rng(rep).uniform(0, 1, (40, 50)).astype(np.float32), 0.05, 3, 1)  # eager work between
        info = ctx.pipeline_batch(batch.frames, streams=2)
        assert info["graph_reused"] == 1
        for k, (got, ref) in enumerate(zip(batch.outputs(), expected_plain)):
            assert np.array_equal(got, ref), (rep, k)


@pytest.mark.parametrize("streams", [1, 3])
def test_batch_host_frames(ctx, batch, expected_plain, streams):
    """host_frames: the same frames in pinned host memory, uploaded, run and
    downloaded as a pipeline over three streams (no graph; `streams` has no
    effect), give the device-frame results bit for bit, again in reverse
    order, so that each staging buffer holds frames of other sizes; pageable
    host memory is rejected."""
    import darkslide_amd as D
    say("batch host frames, streams", streams)
    hb, frames = [], []
    try:
        for f, (raw, _, w, h, *_rest) in zip(batch.frames, batch.bufs):
            hin, hout = ctx.host_alloc(w * h * 2), ctx.host_alloc(w * h * 12)
            hin.array((h, w), np.uint16)[:] = raw.download((h, w), np.uint16)
            g = D.Frame.from_buffer_copy(f)
            g.raw, g.rgb_out = hin.ptr.value, hout.ptr.value
            frames.append(g)
            hb.append((hin, hout, w, h))
        for rep in range(2):
            for (_, hout, w, h) in hb:
                hout.array((h, w, 3), np.float32)[:] = np.nan
            order = list(range(len(frames)))[::(-1 if rep else 1)]
            info = ctx.pipeline_batch([frames[k] for k in order], streams=streams, host_frames=True)
            assert info["frames"] == len(frames) and info["branches"] == 1 and info["graph_reused"] == 0
            for k, ((_, hout, w, h), ref) in enumerate(zip(hb, expected_plain)):
                assert np.array_equal(hout.array((h, w, 3), np.float32), ref), (rep, k)
        pageable = np.zeros(FRAMES[0][0] * FRAMES[0][1], np.uint16)
        g = D.Frame.from_buffer_copy(frames[0])
        g.raw = pageable.ctypes.data
        with pytest.raises(D.DarkslideError):
            ctx.pipeline_batch([g], streams=1, host_frames=True)
    finally:
        for (hin, hout, *_r) in hb:
            hin.free()
            hout.free()


def test_batch_after_workspace_growth(ctx, batch, expected_plain):
    """A larger batch grows the branch workspaces; the first batch's graph is captured again, not replayed stale."""
    say("batch growth")
    big = Batch(ctx, [(900, 700, 14, 2.0, (0, 1, 1, 2), 21)])
    try:
        ctx.pipeline_batch(big.frames, streams=2)
        ctx.pipeline_batch(big.frames, streams=1)
        batch.clear_outputs()
        info = ctx.pipeline_batch(batch.frames, streams=2)
        assert info["graph_reused"] == 0
        for k, (got, ref) in enumerate(zip(batch.outputs(), expected_plain)):
            assert np.array_equal(got, ref), k
        assert np.array_equal(big.outputs()[0], single(ctx, big)[0])
    finally:
        big.free()


def test_batch_without_stages_and_gating(ctx, batch):
    say("batch gating")
    import darkslide_amd as D
    # demosaic only
    batch.clear_outputs()
    ctx.pipeline_batch(batch.frames, denoise=D.DENOISE_NONE, raw_details=False, streams=2)
    ref = single(ctx, batch, denoise=None, rl=None)
    for got, r in zip(batch.outputs(), ref):
        assert np.array_equal(got, r)
    # Engine::base_for: amount.max(1); no RL below a 3200-pixel long edge
    batch.clear_outputs()
    ctx.pipeline_batch(batch.frames, amount=0.0, reference_gating=True, streams=2)
    ref = single(ctx, batch, amount=1.0, rl=None)
    for got, r in zip(batch.outputs(), ref):
        assert np.array_equal(got, r)


def test_batch_learned_matches_single_frames(ctx, batch):
    say("batch learned")
    import darkslide_amd as D
    net = D.CNN(ctx, seed=5)
    try:
        batch.clear_outputs()
        info = ctx.pipeline_batch(batch.frames, denoise=D.DENOISE_LEARNED, net=net, threshold_db=0.3, streams=2)
        say(f"  {info}")
        ref = single(ctx, batch, denoise="learned", net=net)
        for k, (got, r) in enumerate(zip(batch.outputs(), ref)):
            assert np.array_equal(got, r), k
        regions = sum(((w + 255) // 256) * ((h + 255) // 256) for (w, h, *_) in FRAMES)
        assert info["regions"] == regions
        # the untrained network is far from NLM: the spot check sends every region back
        assert info["fallback_regions"] == regions
        # the same from pinned host memory, through the three-stream pipeline
        hb, frames = [], []
        try:
            for f, (raw, _, w, h, *_rest) in zip(batch.frames, batch.bufs):
                hin, hout = ctx.host_alloc(w * h * 2), ctx.host_alloc(w * h * 12)
                hb.append((hin, hout, w, h))
                hin.array((h, w), np.uint16)[:] = raw.download((h, w), np.uint16)
                hout.array((h, w, 3), np.float32)[:] = np.nan
                g = D.Frame.from_buffer_copy(f)
                g.raw, g.rgb_out = hin.ptr.value, hout.ptr.value
                frames.append(g)
            info = ctx.pipeline_batch(frames, denoise=D.DENOISE_LEARNED, net=net, threshold_db=0.3, host_frames=True)
            assert info["regions"] == regions and info["fallback_regions"] == regions
            for k, ((_, hout, w, h), r) in enumerate(zip(hb, ref)):
                assert np.array_equal(hout.array((h, w, 3), np.float32), r), k
        finally:
            for (hin, hout, *_r) in hb:
                hin.free()
                hout.free()
    finally:
        net.close()


def test_tile_variants_are_identical(ctx):
    say("tile variants")
    import darkslide_amd as D
    w, h = 517, 389
    raw = D.synth_bayer(31, w, h, 14, 4.0)
    base = None
    for v in range(3):
        ctx.set_tuning(w, h, rcd=v, mhc=v, rl=v)
        assert ctx.tuning(w, h) == (v, v, v)
        rgb, clip = ctx.demosaic_u16(raw)
        mhc, _ = ctx.demosaic_u16(raw, algo=D.ALGO_MHC)
        rl = ctx.raw_details(rgb, 1.0, 5)
        cur = (rgb, clip, mhc, rl)
        if base is None:
            base = cur
        else:
            for a, b in zip(cur, base):
                assert np.array_equal(a, b), v
    t = ctx.autotune(w, h)
    say(f"  autotune {t}")
    assert all(x > 0 for x in t["rcd_ms"]) and all(x > 0 for x in t["rl_ms"])
    assert ctx.tuning(w, h) == (t["rcd_variant"], t["mhc_variant"], t["rl_variant"])
    assert t["rcd_ms"][t["rcd_variant"]] == min(t["rcd_ms"])
    ctx.set_tuning(w, h)


def test_batch_autotuned(ctx, batch, expected_plain):
    say("batch autotuned")
    batch.clear_outputs()
    info = ctx.pipeline_batch(batch.frames, streams=0, autotune=True)
    say(f"  {info}")
    assert info["branches"] in (1, 2, 4)
    for k, (got, ref) in enumerate(zip(batch.outputs(), expected_plain)):
        assert np.array_equal(got, ref), k


def test_batch_rejects_bad_input(ctx, batch):
    import darkslide_amd as D
    f = batch.frames[0]
    bad = D.Frame(f.raw, f.w, f.h, f.w - 1, f.pattern, f.black, f.white, f.mul, f.rgb_out)
    with pytest.raises(D.DarkslideError):
        ctx.pipeline_batch([bad])
    with pytest.raises(D.DarkslideError):
        ctx.pipeline_batch(batch.frames, denoise=D.DENOISE_LEARNED, net=None)
    with pytest.raises(D.DarkslideError):
        ctx.pipeline_batch(batch.frames, streams=99)


def test_auto_wb_stats_matches_oracle(ctx, oracle_mod):
    say("wb stats")
    import darkslide_amd as D
    raw = D.synth_bayer(41, 640, 480, 14, 2.0)
    rgb, _ = ctx.demosaic_u16(raw)
    nas = (0.55, 1.0, 0.8)
    got = ctx.auto_wb_stats(rgb, nas)
    acc, n, neutral = oracle_mod.auto_wb_stats(rgb, nas)
    say(f"  gpu {got['neutral']} oracle {neutral}")
    assert got["valid"] == 1 and n > 0
    assert abs(got["n"] - n) <= 1e-12 * n
    assert np.allclose(got["acc"], acc, rtol=1e-12, atol=1e-9)
    assert np.allclose(got["neutral"], neutral, rtol=1e-6)
    none = ctx.auto_wb_stats(np.full((8, 8, 3), 0.95, np.float32), nas)
    assert none["valid"] == 0 and none["n"] == 0.0
