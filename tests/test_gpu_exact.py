"""Bit-exact parity where the reference's rounding depends on libm and on
its thread pool.

* The device builds of ds_cbrtf / ds_expf (csrc/libm_ref.h) against the
  host glibc on all 2^32 inputs, by an order-free hash of the outputs.
* Oklab::from_prophoto (color.rs:408-412) bit for bit, including inputs
  spread over the whole float range.
* nlm (enhance.rs:34-72) in the reference's summation order for a rayon
  pool of T threads, and denoise (enhance.rs:74-92) around it: bit for bit
  the oracle run with T threads, for several T.
* noise_reduction (pipeline/detail.rs:13-92), whose only inexact step was
  the Oklab cube root: bit for bit.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

from conftest import RGGB, synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def rgb_frame(O, seed, w, h, iso=4.0):
    raw = synth(seed, w, h, 14, iso)
    return O.rcd_bayer(O.cfa_normalize_u16(raw, RGGB, 512.0, [16383.0] * 3))


def test_libm_device_hash_equals_glibc(ctx):
    exe = os.path.join(ROOT, "tests", "cpp", "libm_exhaustive")
    assert os.path.exists(exe), "tests/cpp/libm_exhaustive is built by __graft_entry__.build()"
    say("hashing glibc cbrtf / expf over 2^32 inputs on the host ...")
    r = subprocess.run([exe, "hash"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    host = {k: int(v, 16) for k, v in (line.split() for line in r.stdout.strip().splitlines())}
    dev = {"cbrtf": ctx.libm_hash(0), "expf": ctx.libm_hash(1)}
    say(f"glibc {host}  device {dev}")
    assert dev == host


def test_oklab_forward_bit_exact(ctx, oracle_mod):
    O = oracle_mod
    rgb = rgb_frame(O, 600, 160, 120)
    assert np.array_equal(ctx.oklab_from_prophoto(rgb), O.oklab_from_prophoto(rgb))
    rng = np.random.default_rng(7)
    # linear data, wide-range magnitudes (lms spans subnormal to large) and exact zeros
    v = np.concatenate([rng.random(3 * 400_000, np.float32),
                        (rng.random(3 * 200_000) * 2.0 ** rng.integers(-140, 60, 3 * 200_000)).astype(np.float32),
                        np.zeros(30, np.float32)]).reshape(-1, 3)
    got, want = ctx.oklab_from_prophoto(v), O.oklab_from_prophoto(v)
    assert np.array_equal(got, want, equal_nan=True), f"bit-equal {np.mean(got == want):.7f}"


@pytest.mark.parametrize("S,P,T", [(3, 1, 1), (3, 1, 8), (10, 3, 16), (10, 3, 5), (2, 0, 3), (0, 2, 4), (4, 2, 100)])
def test_nlm_reference_order_bit_exact(ctx, oracle_mod, S, P, T):
    O = oracle_mod
    rgb = rgb_frame(O, 700 + S, 97, 61, iso=8.0)
    L = np.ascontiguousarray(O.oklab_from_prophoto(rgb)[..., 0])
    hp = np.float32(1.5) * np.float32(max(O.estimate_noise(L), 1e-4))
    got = ctx.nlm_reference_order(L, hp, S, P, T)
    want = O.nlm(L, hp, S, P, nthreads=T)
    assert np.array_equal(got, want), f"bit-equal {np.mean(got == want):.6f} max-abs {np.abs(got - want).max():.3g}"


@pytest.mark.parametrize("amount,S,P,T", [(0.0, 10, 3, 8), (0.0, 10, 3, 16), (37.0, 3, 1, 4), (100.0, 3, 1, 1)])
def test_denoise_reference_order_bit_exact(ctx, oracle_mod, amount, S, P, T):
    """Oklab split, noise estimate, reference-order NLM, both guided filters,
    Oklab back: the composite equals the oracle's for the same pool size."""
    O = oracle_mod
    rgb = rgb_frame(O, 400, 256, 192, iso=8.0)
    got = ctx.denoise_reference_order(rgb, amount, S, P, T)
    want = O.denoise(rgb, amount, S, P, nthreads=T)
    fast = ctx.denoise(rgb, amount, S, P)
    say(f"denoise amount={amount} S={S} P={P} T={T}: reference order bit-equal {np.array_equal(got, want)}; "
        f"fast path vs oracle {np.abs(fast - want).max():.3g}")
    assert np.array_equal(got, want), f"bit-equal {np.mean(got == want):.6f} max-abs {np.abs(got - want).max():.3g}"


def test_denoise_reference_order_1024(ctx, oracle_mod):
    """Config-1 frame (1024x1024), the configured NLM (search 10, patch 3),
    for the oracle's default pool of 16 threads."""
    O = oracle_mod
    raw = synth(0x5EED0000, 1024, 1024)
    rgb = O.rcd_bayer(O.cfa_normalize_u16(raw, RGGB, 512.0, [16383.0] * 3))
    got = ctx.denoise_reference_order(rgb, 0.0, 10, 3, 16)
    want = O.denoise(rgb, 0.0, 10, 3, nthreads=16)
    assert np.array_equal(got, want), f"bit-equal {np.mean(got == want):.6f} max-abs {np.abs(got - want).max():.3g}"
