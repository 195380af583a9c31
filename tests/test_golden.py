"""Golden fixtures (tests/golden/make_fixtures.py): the oracle reproduces
them bit-for-bit on the CPU, and the GPU path matches them."""
import os

import numpy as np
import pytest

from conftest import ROOT, maxabs

GOLDEN = os.path.join(ROOT, "tests", "golden")
FILES = sorted(f for f in os.listdir(GOLDEN) if f.endswith(".npz"))


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.mark.parametrize("name", FILES)
def test_oracle_reproduces_golden(oracle_mod, name):
    O = oracle_mod
    g = load(name)
    cfa = O.cfa_normalize_u16(g["raw"], (0, 1, 1, 2), 512.0, [16383.0] * 3)
    assert np.array_equal(cfa, g["cfa"])
    assert np.array_equal(O.rcd_bayer(cfa), g["rcd"])
    assert np.array_equal(O.mhc_bayer(cfa), g["mhc"])
    assert np.array_equal(O.clip_plane(cfa), g["clip"])
    assert np.float32(O.estimate_noise(g["L"])) == g["sigma"]
    assert np.array_equal(O.nlm(g["L"], g["sigma"] * np.float32(0.6), 10, 3, nthreads=1), g["nlm_10_3"])
    assert np.array_equal(O.raw_details(g["rcd"], 1.0, 20), g["raw_details_1_20"])
    assert np.array_equal(O.denoise(g["rcd"], 0.0, 10, 3, nthreads=1), g["denoise_10_3"])


@pytest.mark.gpu
@pytest.mark.parametrize("name", FILES)
def test_gpu_matches_golden(ctx, name):
    g = load(name)
    rgb, clip = ctx.demosaic_u16(g["raw"])
    assert np.array_equal(rgb, g["rcd"]) and np.array_equal(clip, g["clip"])
    assert np.array_equal(ctx.demosaic_u16(g["raw"], algo=1, clip=False), g["mhc"])
    assert np.float32(ctx.estimate_noise(g["L"])) == g["sigma"]
    assert maxabs(ctx.nlm(g["L"], g["sigma"] * np.float32(1.5), 3, 1), g["nlm_3_1"]) <= 2e-5
    assert maxabs(ctx.nlm(g["L"], g["sigma"] * np.float32(0.6), 10, 3), g["nlm_10_3"]) <= 2e-5
    assert np.array_equal(ctx.raw_details(g["rcd"], 1.0, 20), g["raw_details_1_20"])
    # the fixtures' NLM and denoise were written by the oracle with one thread:
    # the reference-order mode reproduces them bit for bit
    assert np.array_equal(ctx.nlm_reference_order(g["L"], g["sigma"] * np.float32(1.5), 3, 1, 1), g["nlm_3_1"])
    assert np.array_equal(ctx.nlm_reference_order(g["L"], g["sigma"] * np.float32(0.6), 10, 3, 1), g["nlm_10_3"])
    assert np.array_equal(ctx.denoise_reference_order(g["rcd"], 0.0, 10, 3, 1), g["denoise_10_3"])
