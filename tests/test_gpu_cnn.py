"""GPU parity of the learned denoiser against its FP64 oracle.

* forward (BF16 MFMA implicit GEMM, FP32 accumulation) against the oracle's
  bf16-emulating FP64 forward;
* loss and gradient of every layer against the oracle's FP64 backward;
* Adam steps against the oracle's FP64 Adam on the same gradient;
* the training-pair generator against the plain path's stages;
* the learned denoise's NLM spot check and fallback (forced by thresholds
  and by a corrupted network);
* training on the GPU lowers the loss, and the trained network gains PSNR
  over the untrained one against the noise-free frame;
* the shipped checkpoint against NLM and against no denoise at noise levels
  1x, 4x and 16x (luminance PSNR against the noise-free frame).
"""
import sys
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RGGB = (0, 1, 1, 2)
WHITE = (16383.0, 16383.0, 16383.0)
NP = 56097


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _params(seed, out_scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NP)
    p[0:288] = rng.normal(0, np.sqrt(2 / 9), 288)
    p[288:320] = rng.normal(0, 0.05, 32)
    o = 320
    for _ in range(6):
        p[o:o + 9216] = rng.normal(0, np.sqrt(2 / 288), 9216)
        p[o + 9216:o + 9248] = rng.normal(0, 0.05, 32)
        o += 9248
    p[o:o + 288] = rng.normal(0, np.sqrt(2 / 288), 288) * out_scale
    p[o + 288] = 0.01
    return p.astype(np.float32)


# layer slices of the flat parameter vector: (name, start, stop)
LAYERS = [("in.w", 0, 288), ("in.b", 288, 320)]
for _l in range(6):
    _o = 320 + _l * 9248
    LAYERS += [(f"mid{_l}.w", _o, _o + 9216), (f"mid{_l}.b", _o + 9216, _o + 9248)]
LAYERS += [("out.w", 55808, 56096), ("out.b", 56096, 56097)]


@pytest.fixture(scope="module")
def net(ctx):
    import darkslide_amd as D
    n = D.CNN(ctx, seed=7)
    yield n
    n.close()


@pytest.mark.parametrize("shape", [(2, 19, 70), (1, 64, 64), (3, 5, 7), (1, 33, 130)])
def test_forward_matches_oracle(net, oracle_mod, shape):
    say("cnn forward", shape)
    p = _params(11)
    net.set_params(p)
    x = np.random.default_rng(12).normal(0, 1, shape).astype(np.float32)
    y = net.forward(x)
    ref = oracle_mod.cnn_forward(p.astype(np.float64), x, bf16=True)
    scale = np.abs(ref - x).max()
    err = np.abs(y - ref)
    say(f"  max {err.max():.3e} rms {np.sqrt((err ** 2).mean()):.3e} (residual scale {scale:.3f})")
    assert scale > 0.1
    assert err.max() <= 1e-2 * scale
    assert np.sqrt((err ** 2).mean()) <= 1e-3 * scale


@pytest.mark.parametrize("shape", [(1, 64, 64), (2, 19, 70), (1, 300, 170), (3, 5, 7), (1, 1, 1), (1, 517, 81), (2, 260, 161)])
def test_fused_forward_equals_layer_by_layer(net, shape):
    """The fused inference kernel (one block per 80-column strip, rings of
    rows in LDS) reproduces the layer-by-layer forward bit for bit, across
    strip and row-block boundaries and image borders."""
    say("cnn fused forward", shape)
    net.set_params(_params(21))
    x = np.random.default_rng(22).normal(0, 2, shape).astype(np.float32)
    y = net.forward(x)
    yf = net.infer(x)
    assert np.array_equal(yf, y), np.abs(yf - y).max()


@pytest.mark.parametrize("shape", [(1, 600, 700), (2, 400, 181)])
def test_fused_forward_repeats_across_row_blocks(net, shape):
    """Repeated fused inference over several 192-row blocks gives the same
    bits every time, equal to the layer-by-layer forward: each block's first
    step reads the input rows its prologue committed only after a barrier."""
    say("cnn fused forward repeated", shape)
    net.set_params(_params(23))
    x = np.random.default_rng(24).normal(0, 2, shape).astype(np.float32)
    y = net.forward(x)
    for _ in range(4):
        yf = net.infer(x)
        assert np.array_equal(yf, y), np.argwhere(yf != y)[:4]


@pytest.mark.parametrize("shape,margin", [((2, 19, 70), 2), ((1, 40, 64), 0)])
def test_gradient_matches_oracle(net, oracle_mod, shape, margin):
    say("cnn gradient", shape, margin)
    p = _params(13)
    net.set_params(p)
    rng = np.random.default_rng(14)
    x = rng.normal(0, 1, shape).astype(np.float32)
    t = (x + rng.normal(0, 0.5, shape)).astype(np.float32)
    loss, g = net.loss_grad(x, t, margin)
    ref_loss, ref_g = oracle_mod.cnn_loss_grad(p.astype(np.float64), x, t, margin, bf16=True)
    say(f"  loss {loss:.6f} oracle {ref_loss:.6f}")
    assert abs(loss - ref_loss) <= 1e-3 * ref_loss
    for name, a, b in LAYERS:
        gg, rr = g[a:b].astype(np.float64), ref_g[a:b]
        rel = np.linalg.norm(gg - rr) / max(np.linalg.norm(rr), 1e-30)
        cos = gg @ rr / max(np.linalg.norm(gg) * np.linalg.norm(rr), 1e-30)
        say(f"  {name:8s} rel {rel:.2e} cos {cos:.6f}")
        assert rel <= 3e-2, name
        assert cos >= 0.999, name


def test_adam_steps_match_oracle(net, oracle_mod, ctx):
    say("cnn adam")
    p = _params(15)
    net.set_params(p)
    rng = np.random.default_rng(16)
    shape = (2, 16, 24)
    x = rng.normal(0, 1, shape).astype(np.float32)
    t = (x + rng.normal(0, 0.5, shape)).astype(np.float32)
    bx, bt = ctx.alloc(x.nbytes), ctx.alloc(t.nbytes)
    bx.upload(x)
    bt.upload(t)
    w = p.astype(np.float64)
    m = np.zeros(NP)
    v = np.zeros(NP)
    lr = 1e-3
    for step in (1, 2, 3):
        _, g = net.loss_grad(x, t, 1)
        net.train_step_dev(bx.ptr, bt.ptr, *shape, 1, lr)
        oracle_mod.cnn_adam(w, g.astype(np.float64), m, v, step, lr)
        got = net.params()
        err = np.abs(got - w).max()
        say(f"  step {step}: max |gpu - oracle| {err:.3e}")
        assert net.step == step
        assert err <= 2e-6
    # the steps moved the weights by about lr each
    assert 2e-3 < np.abs(net.params() - p).max() <= 3.01e-3
    bx.free()
    bt.free()


def _frame_L(ctx, seed, w, h, iso):
    """Noisy or (iso 0) noise-free frame -> demosaic (RCD) -> Oklab L on the GPU."""
    import darkslide_amd as D
    buf = ctx.alloc(w * h * 2)
    ctx._check(D.lib().ds_synth_bayer_dev(ctx.handle, seed, w, h, 14, float(iso), D._pat(RGGB), buf.ptr))
    raw = buf.download((h, w), np.uint16)
    buf.free()
    rgb, _ = ctx.demosaic_u16(raw, RGGB, 512.0, WHITE)
    return rgb, ctx.oklab_from_prophoto(rgb)[..., 0].copy()


def test_make_pairs_matches_plain_stages(ctx, oracle_mod):
    say("cnn make_pairs")
    seed, n, iso = 0xABC, 96, 4.0
    x, t = ctx.make_pairs(seed, n, n, iso, crop=n, n_crops=1)
    _, L = _frame_L(ctx, seed, n, n, iso)
    sigma = ctx.estimate_noise(L)
    s = max(np.float32(sigma), np.float32(1e-4))
    xr = oracle_mod.cnn_input(L, float(s))
    assert np.abs(x[0] - xr).max() <= 1e-4 * max(1.0, np.abs(xr).max())
    lf = ctx.nlm(L, np.float32(s) * np.float32(0.6), 10, 3)
    tr = x[0] + (lf - L) / s
    assert np.abs(t[0] - tr).max() <= 1e-5
    # seeded crops: deterministic, and inside the frame's values
    x2, t2 = ctx.make_pairs(seed, n, n, iso, crop=32, n_crops=5)
    x3, _ = ctx.make_pairs(seed, n, n, iso, crop=32, n_crops=5)
    assert np.array_equal(x2, x3)
    for k in range(5):
        hits = [(oy, ox) for oy in range(n - 31) for ox in range(n - 31)
                if x2[k, 0, 0] == x[0, oy, ox] and np.array_equal(x2[k], x[0, oy:oy + 32, ox:ox + 32])]
        assert hits, k


def test_denoise_plane_is_forward_plus_output_map(ctx, net):
    say("cnn denoise_plane")
    net.set_params(_params(17, out_scale=0.2))
    n = 128
    x, _ = ctx.make_pairs(0x77, n, n, 8.0, crop=n, n_crops=1)
    _, L = _frame_L(ctx, 0x77, n, n, 8.0)
    s = max(np.float32(ctx.estimate_noise(L)), np.float32(1e-4))
    lf = net.denoise_plane(L)
    y = net.forward(x)[0]
    ref = L + s * (y - x[0])
    assert np.abs(lf - ref).max() <= 1e-6
    assert np.array_equal(net.infer(x)[0], y)


def _noisy_rgb(ctx, seed, w, h, iso):
    rgb, _ = _frame_L(ctx, seed, w, h, iso)
    return rgb


def test_learned_fallback_forced(ctx, net):
    """threshold <= 0 sends every region to NLM: the result is the plain path's, bit for bit."""
    say("learned fallback forced")
    import darkslide_amd as D
    net.set_params(_params(18, out_scale=0.2))
    rgb = _noisy_rgb(ctx, 0x1234, 300, 260, 4.0)
    plain = ctx.denoise(rgb, 0.0, 10, 3)
    out, st = ctx.denoise_learned(net, rgb, 0.0, 10, 3, threshold_db=0.0)
    assert st["regions"] == 4 and st["fallback"] == 4
    assert np.array_equal(out, plain)
    # no fallback at an infinite threshold: the network's result stands
    out2, st2 = ctx.denoise_learned(net, rgb, 0.0, 10, 3, threshold_db=float("inf"))
    assert st2["fallback"] == 0 and st2["regions"] == 4
    assert np.abs(out2 - plain).max() > 1e-4
    assert st2["mean_gap_db"] > 0.0
    D.lib()


def test_learned_fallback_on_corrupted_network(ctx, net):
    """A network far from NLM is caught by the spot check in every region."""
    say("learned fallback corrupted")
    p = _params(19, out_scale=0.2)
    p[55808:56097] *= 200.0
    net.set_params(p)
    rgb = _noisy_rgb(ctx, 0x4321, 520, 300, 8.0)
    plain = ctx.denoise(rgb, 0.0, 10, 3)
    out, st = ctx.denoise_learned(net, rgb, 0.0, 10, 3, threshold_db=0.3)
    say(f"  stats {st}")
    assert st["regions"] == 6 and st["fallback"] == 6
    assert st["max_gap_db"] > 3.0
    assert np.array_equal(out, plain)


def _psnr(a, ref):
    mse = float(np.mean((a.astype(np.float64) - ref.astype(np.float64)) ** 2))
    return 10 * np.log10(1.0 / max(mse, 1e-30))


@pytest.mark.slow
def test_training_gains_over_untrained(ctx):
    """Trained on the GPU for a short schedule, the network's luminance is
    closer to the noise-free frame than the untrained network's, and its
    loss on the training targets has dropped."""
    import darkslide_amd as D
    from darkslide_amd import train_cnn
    net = D.CNN(ctx, seed=3)
    try:
        wL = 384
        seeds = [0x900D + i for i in range(3)]
        evals = []
        for sd in seeds:
            _, Ln = _frame_L(ctx, sd, wL, wL, 8.0)
            _, Lc = _frame_L(ctx, sd, wL, wL, 0.0)
            evals.append((Ln, Lc))

        def score():
            return float(np.mean([_psnr(net.denoise_plane(Ln), Lc) - _psnr(Ln, Lc) for Ln, Lc in evals]))

        g0 = score()
        t0 = time.time()
        hist = train_cnn.train(ctx, net, steps=400, batch=16, crop=64, frame=256, pool_frames=64, lr=1e-3, seed=5, log=say)
        say(f"  400 steps in {time.time() - t0:.1f} s; loss {hist[0]:.4f} -> {hist[-1]:.4f}")
        g1 = score()
        nlm_gain = float(np.mean([
            _psnr(ctx.nlm(Ln, np.float32(max(np.float32(ctx.estimate_noise(Ln)), np.float32(1e-4))) * np.float32(0.6), 10, 3), Lc)
            - _psnr(Ln, Lc) for Ln, Lc in evals]))
        say(f"  PSNR gain over the noisy input: untrained {g0:.3f} dB, trained {g1:.3f} dB, NLM {nlm_gain:.3f} dB")
        assert hist[-1] < 0.8 * hist[0]
        assert g1 > g0 + 0.2
        assert g1 > 0.0
    finally:
        net.close()


@pytest.mark.parametrize("iso", [1.0, 4.0, 16.0])
def test_trained_network_against_nlm_and_no_denoise(ctx, iso):
    """The shipped checkpoint against NLM (h = 0.6 sigma, 21 x 21 search, 7 x 7
    patch, the plain path's denoise) and against no denoise: luminance PSNR
    against the noise-free frame, four seeded 512 x 512 synthetic frames per
    noise level.  The network must improve on the noisy input and stay
    within 0.3 dB of NLM at every level."""
    import os
    import darkslide_amd as D
    ck = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "darkslide_amd", "weights", "cnn_l32x8.dsck")
    net = D.CNN(ctx, seed=1)
    try:
        net.load(ck)
        rows = []
        for sd in range(4):
            _, Ln = _frame_L(ctx, 0xC0 + sd, 512, 512, iso)
            _, Lc = _frame_L(ctx, 0xC0 + sd, 512, 512, 0.0)
            s = max(np.float32(ctx.estimate_noise(Ln)), np.float32(1e-4))
            rows.append((_psnr(Ln, Lc), _psnr(net.denoise_plane(Ln), Lc), _psnr(ctx.nlm(Ln, s * np.float32(0.6), 10, 3), Lc)))
        noisy, cnn, nlm = (float(np.mean(c)) for c in zip(*rows))
        say(f"  iso {iso:g}: luminance PSNR noisy {noisy:.3f} dB, trained CNN {cnn:.3f} dB ({cnn - noisy:+.3f}), "
            f"NLM {nlm:.3f} dB ({nlm - noisy:+.3f})")
        assert cnn > noisy
        assert cnn >= nlm - 0.3
    finally:
        net.close()
