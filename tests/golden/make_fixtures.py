"""Writes the committed golden fixtures (tests/golden/*.npz).

The reference ships no golden vectors (its tests are invariants only), so
these are generated by the oracle's FP32 restatement from seeded synthetic
RGGB frames; the CPU suite checks the oracle still reproduces them
bit-for-bit and the GPU suite checks the HIP path against them.

    python3 tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

CASES = [("rggb_64x64_iso4", 0x5EED0100, 64, 64, 4.0), ("rggb_48x40_iso16", 0x5EED0101, 48, 40, 16.0)]


def make(name, seed, w, h, iso):
    import darkslide_amd as D
    import oracle as O
    raw = D.synth_bayer(seed, w, h, 14, iso)
    cfa = O.cfa_normalize_u16(raw, D.RGGB, 512.0, [16383.0] * 3)
    rcd = O.rcd_bayer(cfa)
    L = np.ascontiguousarray(O.oklab_from_prophoto(rcd)[..., 0])
    sigma = np.float32(O.estimate_noise(L))
    out = dict(
        raw=raw, cfa=cfa, rcd=rcd, mhc=O.mhc_bayer(cfa), clip=O.clip_plane(cfa), L=L, sigma=sigma,
        nlm_3_1=O.nlm(L, sigma * np.float32(1.5), 3, 1, nthreads=1),
        nlm_10_3=O.nlm(L, sigma * np.float32(0.6), 10, 3, nthreads=1),
        denoise_10_3=O.denoise(rcd, 0.0, 10, 3, nthreads=1),
        raw_details_1_20=O.raw_details(rcd, 1.0, 20),
    )
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)


if __name__ == "__main__":
    for c in CASES:
        make(*c)
        print("wrote", c[0])
