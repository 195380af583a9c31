"""Pins of the learned denoiser's FP64 oracle (oracle/cnn_oracle.cpp).

The network has no counterpart in the reference, so the oracle is pinned by
independent library routines and by calculus:
  * forward against torch.nn.functional.conv2d in FP64 on the CPU;
  * gradient against torch autograd in FP64, and against central finite
    differences of the oracle's own loss;
  * the bf16 rounding against torch's float32 -> bfloat16 conversion;
  * the bf16-emulating forward against torch with the same roundings
    (weights, the input layer's x, activations);
  * Adam against torch.optim.Adam;
  * the network input against a numpy restatement with edge padding.
"""
import numpy as np
import pytest
import torch

import oracle as O

NP = 56097


def _params(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    p = np.zeros(NP)
    p[0:288] = rng.normal(0, np.sqrt(2 / 9), 288)
    p[288:320] = rng.normal(0, 0.05, 32)
    o = 320
    for _ in range(6):
        p[o:o + 9216] = rng.normal(0, np.sqrt(2 / 288), 9216)
        p[o + 9216:o + 9248] = rng.normal(0, 0.05, 32)
        o += 9248
    p[o:o + 288] = rng.normal(0, np.sqrt(2 / 288), 288) * scale
    p[o + 288] = 0.01
    return p


def _torch_forward(P, X, bf16=False):
    F = torch.nn.functional

    def rnd(t):
        return t.to(torch.float32).to(torch.bfloat16).to(torch.float64) if bf16 else t

    Win = P[0:288].view(32, 1, 3, 3)
    a = rnd(torch.relu(F.conv2d(rnd(X)[:, None], rnd(Win), P[288:320], padding=1)))
    o = 320
    for _ in range(6):
        W = P[o:o + 9216].view(32, 3, 3, 32).permute(0, 3, 1, 2).contiguous()
        if bf16:
            W = rnd(W)
        a = rnd(torch.relu(F.conv2d(a, W, P[o + 9216:o + 9248], padding=1)))
        o += 9248
    Wout = P[o:o + 288].view(3, 3, 32).permute(2, 0, 1)[None].contiguous()
    return X + F.conv2d(a, rnd(Wout), P[o + 288:o + 289], padding=1)[:, 0]


def test_param_count():
    assert O.cnn_param_count() == NP == 9 * (32 + 6 * 32 * 32 + 32) + 32 + 6 * 32 + 1


def test_forward_matches_torch_conv2d():
    p = _params(1)
    x = np.random.default_rng(2).normal(0, 1, (2, 11, 13)).astype(np.float32)
    out = O.cnn_forward(p, x)
    ref = _torch_forward(torch.tensor(p), torch.tensor(x, dtype=torch.float64)).numpy()
    assert np.abs(out - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    # the network is not the identity and not trivially zero inside
    assert np.abs(out - x).max() > 1e-3


@pytest.mark.parametrize("margin", [0, 2])
def test_gradient_matches_torch_autograd(margin):
    p = _params(3)
    rng = np.random.default_rng(4)
    x = rng.normal(0, 1, (2, 9, 12)).astype(np.float32)
    t = rng.normal(0, 1, (2, 9, 12)).astype(np.float32)
    loss, g = O.cnn_loss_grad(p, x, t, margin)
    P = torch.tensor(p, requires_grad=True)
    out = _torch_forward(P, torch.tensor(x, dtype=torch.float64))
    d = out - torch.tensor(t, dtype=torch.float64)
    if margin:
        d = d[:, margin:-margin, margin:-margin]
    tl = (d ** 2).mean()
    tl.backward()
    assert abs(loss - tl.item()) <= 1e-12 * abs(tl.item())
    assert np.abs(g - P.grad.numpy()).max() <= 1e-12 * np.abs(g).max()


def test_gradient_matches_finite_differences():
    p = _params(5)
    rng = np.random.default_rng(6)
    x = rng.normal(0, 1, (1, 8, 9)).astype(np.float32)
    t = rng.normal(0, 1, (1, 8, 9)).astype(np.float32)
    _, g = O.cnn_loss_grad(p, x, t, 1)
    # a few parameters of every layer, biases included
    idx = [0, 100, 300, 320 + 5, 320 + 9216 + 3, 320 + 3 * 9248 + 777, 320 + 5 * 9248 + 9000, 320 + 6 * 9248 + 17, NP - 1]
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (O.cnn_loss_grad(pp, x, t, 1, grad=False) - O.cnn_loss_grad(pm, x, t, 1, grad=False)) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(7)
    v = np.concatenate([rng.normal(0, 1, 2000), rng.normal(0, 1e-3, 500), [1 + 2 ** -8, 1 + 3 * 2 ** -8, -1 - 2 ** -8, 0.0, 3.0]])
    v32 = v.astype(np.float32)
    ref = torch.tensor(v32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(O.bf16_round(v32.astype(np.float64)), ref)


def test_bf16_forward_matches_torch_with_roundings():
    p = _params(8)
    x = np.random.default_rng(9).normal(0, 1, (1, 10, 10)).astype(np.float32)
    out = O.cnn_forward(p, x, bf16=True)
    ref = _torch_forward(torch.tensor(p), torch.tensor(x, dtype=torch.float64), bf16=True).numpy()
    # equal up to the summation order of FP64 sums (which can move a value
    # across a bf16 rounding boundary in rare cases)
    assert np.median(np.abs(out - ref)) <= 1e-12
    assert np.abs(out - ref).max() <= 2e-2 * np.abs(ref - x).max()
    # and the roundings matter: the bf16 forward is not the FP64 one
    assert np.abs(out - O.cnn_forward(p, x)).max() > 1e-6


def test_adam_matches_torch():
    rng = np.random.default_rng(10)
    w0 = rng.normal(0, 1, 64)
    w = w0.copy()
    m = np.zeros(64)
    v = np.zeros(64)
    tw = torch.tensor(w0, requires_grad=True)
    opt = torch.optim.Adam([tw], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    for step in range(1, 6):
        g = rng.normal(0, 1, 64)
        O.cnn_adam(w, g, m, v, step, 1e-3)
        opt.zero_grad()
        tw.grad = torch.tensor(g)
        opt.step()
    assert np.abs(w - tw.detach().numpy()).max() <= 1e-12


def test_input_normalisation_closed_form():
    rng = np.random.default_rng(11)
    L = rng.uniform(0, 1, (13, 17)).astype(np.float32)
    sigma = 0.005
    x = O.cnn_input(L, sigma)
    pad = np.pad(L.astype(np.float64), 2, mode="edge")
    mean = sum(pad[2 + dy:2 + dy + 13, 2 + dx:2 + dx + 17] for dy in range(-2, 3) for dx in range(-2, 3)) / 25.0
    ref = np.clip((L - mean) / sigma, -64, 64)
    assert np.abs(x - ref).max() <= 1e-9
    assert np.abs(ref).max() == 64.0  # the clamp is exercised
    # sigma floor 1e-4
    assert np.allclose(O.cnn_input(L, 0.0), np.clip((L - mean) / 1e-4, -64, 64))
