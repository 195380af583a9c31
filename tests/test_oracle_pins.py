"""Pins of the CPU oracle (oracle/oracle.cpp) against the reference's own
tests, closed forms, textbook restatements and brute force.

The reference ships no golden vectors; its only pins are the invariant
tests re-expressed here (demosaic.rs:514-537, enhance.rs:137-151,
image.rs:455-471, color.rs:502-512).  Everything else is checked against an
independent numpy implementation of the published method, written here in
float64 with explicit clamped indexing, so that a dropped term, a wrong sign
or index, or a transposed operand in the oracle fails a test.
"""
import numpy as np
import pytest

from conftest import PATTERNS, RGGB, maxabs, synth

F32 = np.float32


def bayer_colors(h, w, pat):
    yy, xx = np.mgrid[0:h, 0:w]
    return np.asarray(pat)[(yy % 2) * 2 + (xx % 2)]


def clamp_idx(n, lo, hi):
    return np.clip(np.arange(lo, hi), 0, n - 1)


def lcg_noise_plane():
    """The 64x64 step edge + LCG noise of enhance.rs:139-145."""
    seed = 1
    vals = []
    for _ in range(64 * 64):
        seed = (seed * 1664525 + 1013904223) & 0xFFFFFFFF
        vals.append(F32(seed >> 8) / F32(1 << 24) - F32(0.5))
    p = np.where(np.arange(64)[None, :] < 32, F32(0.3), F32(0.7)) * np.ones((64, 1), F32)
    noise = np.array(vals, F32).reshape(64, 64)
    noisy = (p + noise * F32(0.1)).astype(F32)
    return p.astype(F32), noisy


def frame_cfa(seed, w, h, iso=1.0, pattern=RGGB):
    import oracle as O
    raw = synth(seed, w, h, 14, iso, pattern)
    return O.cfa_normalize_u16(raw, pattern, 512.0, [16383.0] * 3)


# ---------------------------------------------------------------- reference tests

def test_flat_field_is_preserved(oracle_mod):
    """demosaic.rs:514-522 (Bayer half): flat CFA -> flat RGB at the centre within 1e-4."""
    O = oracle_mod
    v = np.array([0.2, 0.4, 0.3], F32)
    cfa = v[bayer_colors(32, 32, RGGB)].astype(F32)
    rgb, _ = O.demosaic(cfa, RGGB, (1, 1, 1))
    assert maxabs(rgb[16, 16], v) < 1e-4
    assert maxabs(O.mhc_bayer(cfa)[16, 16], v) < 1e-4


def test_nlm_reduces_noise(oracle_mod):
    """enhance.rs:137-151."""
    O = oracle_mod
    p, noisy = lcg_noise_plane()
    s = O.estimate_noise(noisy)
    assert 0.01 < s < 0.06, s
    d = O.nlm(noisy, s * 1.5, 3, 1)
    err = lambda q: float(((q.astype(np.float64) - p) ** 2).sum())
    assert err(d) < err(noisy) * 0.5


def test_box_blur_preserves_constant(oracle_mod):
    """image.rs:455-462."""
    O = oracle_mod
    p = np.full((23, 37), 0.5, F32)
    assert maxabs(O.box_blur(p, 4), 0.5) < 1e-5
    assert maxabs(O.gaussian_blur(p, 7.0), 0.5) < 1e-4


def test_guided_preserves_edge(oracle_mod):
    """image.rs:464-471."""
    O = oracle_mod
    p = np.where(np.arange(40)[None, :] < 20, F32(0.1), F32(0.9)) * np.ones((30, 1), F32)
    g = O.guided_filter(p.astype(F32), p.astype(F32), 4, 1e-4)
    assert abs(g[5, 5] - 0.1) < 0.02 and abs(g[5, 35] - 0.9) < 0.02


def test_oklab_roundtrip_and_white(oracle_mod):
    """color.rs:502-512."""
    O = oracle_mod
    w = O.oklab_from_prophoto(np.array([[1, 1, 1]], F32))[0]
    assert abs(w[0] - 1) < 1e-3 and abs(w[1]) < 1e-3 and abs(w[2]) < 1e-3
    c = np.array([[0.3, 0.1, 0.6]], F32)
    assert maxabs(O.oklab_to_prophoto(O.oklab_from_prophoto(c)), c) < 1e-4


# ---------------------------------------------------------------- closed forms

def test_cfa_normalize_closed_form(oracle_mod):
    """decode.rs:302-303 on chosen DNs, below-black values stay negative."""
    O = oracle_mod
    raw = np.array([[0, 512, 513, 16383], [1000, 511, 8000, 65535]], np.uint16)
    out = O.cfa_normalize_u16(raw, RGGB, 512.0, [16383.0, 16000.0, 12000.0])
    cols = bayer_colors(2, 4, RGGB)
    white = np.array([16383.0, 16000.0, 12000.0], F32)[cols]
    expect = (raw.astype(F32) - F32(512)) / (white - F32(512))
    assert np.array_equal(out, expect.astype(F32))
    assert out[1, 1] < 0


def test_clip_plane_brute_force(oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(3)
    cfa = rng.uniform(0.9, 1.0, (17, 23)).astype(F32)
    got = O.clip_plane(cfa)
    raw = (cfa >= F32(0.995)).astype(F32)
    yi, xi = clamp_idx(17, -1, 18), clamp_idx(23, -1, 24)
    pad = raw[yi][:, xi]
    want = np.zeros_like(raw)
    for dy in range(3):
        for dx in range(3):
            want = np.maximum(want, pad[dy:dy + 17, dx:dx + 23])
    assert np.array_equal(got, want)


def test_box_blur_brute_force(oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(4)
    p = rng.random((29, 41)).astype(F32)
    for r in (1, 3, 7, 25):
        yi, xi = clamp_idx(29, -r, 29 + r), clamp_idx(41, -r, 41 + r)
        pad = p.astype(np.float64)[yi][:, xi]
        want = np.zeros((29, 41))
        for dy in range(2 * r + 1):
            for dx in range(2 * r + 1):
                want += pad[dy:dy + 29, dx:dx + 41]
        want /= (2 * r + 1) ** 2
        assert maxabs(O.box_blur(p, r), want) < 2e-6, r
    assert np.array_equal(O.box_blur(p, 0), p)


def gauss_f64(p, sigma):
    """Separable clamped gaussian, taps exp(-i^2 / 2 sigma^2) / sum, float64."""
    r = int(np.ceil(np.float32(sigma) * np.float32(3)))
    k = np.exp(-np.arange(-r, r + 1) ** 2 / (2.0 * sigma * sigma))
    k /= k.sum()
    h, w = p.shape
    xi = clamp_idx(w, -r, w + r)
    tmp = sum(k[j] * p[:, xi[j:j + w]] for j in range(2 * r + 1))
    yi = clamp_idx(h, -r, h + r)
    return sum(k[j] * tmp[yi[j:j + h], :] for j in range(2 * r + 1))


def test_gaussian_taps_and_blur(oracle_mod):
    O = oracle_mod
    for sigma in (0.5, 0.75, 1.0, 1.9):
        taps, r = O.gaussian_taps(sigma)
        assert r == int(np.ceil(sigma * 3))
        k = np.exp(-np.arange(-r, r + 1) ** 2 / (2.0 * sigma * sigma))
        assert maxabs(taps, k / k.sum()) < 1e-7
    rng = np.random.default_rng(5)
    p = rng.random((31, 45)).astype(F32)
    for sigma in (0.5, 1.0, 1.7):
        assert maxabs(O.gaussian_blur(p, sigma), gauss_f64(p.astype(np.float64), sigma)) < 2e-6
    assert np.array_equal(O.gaussian_blur(p, 0.2), p)
    # delta response of the exact path is the separable kernel
    d = np.zeros((21, 21), F32)
    d[10, 10] = 1
    taps, r = O.gaussian_taps(1.0)
    g = O.gaussian_blur(d, 1.0)
    assert maxabs(g[10 - r:10 + r + 1, 10 - r:10 + r + 1], np.outer(taps, taps)) < 1e-7


def test_three_box_gaussian(oracle_mod):
    """sigma >= 2: Wells' radii give a total variance within a box width of sigma^2,
    and the blur equals three brute-force clamped boxes."""
    O = oracle_mod
    for sigma in (2.0, 3.3, 7.0, 20.0):
        radii = O.box_radii_for_gauss(sigma)
        var = sum(((2 * r + 1) ** 2 - 1) / 12.0 for r in radii)
        assert abs(var - sigma * sigma) < 2 * max(radii) + 2, (sigma, radii)
        assert max(radii) - min(radii) <= 1
    rng = np.random.default_rng(6)
    p = rng.random((40, 50)).astype(F32)
    radii = O.box_radii_for_gauss(3.3)
    q = p
    for r in radii:
        q = O.box_blur(q, r)
    assert np.array_equal(O.gaussian_blur(p, 3.3), q)


def test_oklab_matrices_independent(oracle_mod):
    """Oklab from ProPhoto rebuilt in float64 from the published constants:
    ROMM primaries / D50 (color.rs:129), Bradford D50->D65 (color.rs:79-87),
    Ottosson's M1 / M2 (color.rs:383, :387)."""
    O = oracle_mod
    def xyz(x, y):
        return np.array([x / y, 1.0, (1 - x - y) / y])
    d50, d65 = xyz(0.34567, 0.35850), xyz(0.31271, 0.32902)
    prim = np.stack([xyz(0.734699, 0.265301), xyz(0.159597, 0.840403), xyz(0.036598, 0.000105)], axis=1)
    to_xyz = prim * np.linalg.solve(prim, d50)[None, :]
    B = np.array([[0.8951, 0.2664, -0.1614], [-0.7502, 1.7135, 0.0367], [0.0389, -0.0685, 1.0296]])
    brad = np.linalg.inv(B) @ np.diag((B @ d65) / (B @ d50)) @ B
    M1 = np.array([[0.818933, 0.36186674, -0.12885971], [0.032984544, 0.9293119, 0.03614564], [0.0482003, 0.26436627, 0.6338517]])
    to_lms = M1 @ brad @ to_xyz
    m = O.oklab_matrices(np.float32)
    assert maxabs(m["to_lms"], to_lms) < 2e-6
    assert maxabs(m["from_lms"], np.linalg.inv(to_lms)) < 2e-5
    M2 = np.array([[0.21045426, 0.7936178, -0.004072047], [1.9779985, -2.4285922, 0.4505937], [0.025904037, 0.78277177, -0.80867577]])
    assert maxabs(m["m2_inv"], np.linalg.inv(M2)) < 2e-6
    rng = np.random.default_rng(7)
    rgb = rng.uniform(0.01, 1.0, (200, 3))
    lab = (M2 @ np.cbrt(to_lms @ rgb.T)).T
    assert maxabs(O.oklab_from_prophoto(rgb.astype(F32)), lab) < 2e-6
    assert maxabs(O.oklab_from_prophoto(rgb), lab) < 1e-6


def test_luminance(oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(8)
    rgb = rng.random((50, 3)).astype(F32)
    want = rgb.astype(np.float64) @ np.array([0.2880402, 0.7118741, 0.0000857])
    assert maxabs(O.luminance(rgb), want) < 1e-7


# ---------------------------------------------------------------- MHC: textbook

MHC_K = {
    "g_rb": np.array([[0, 0, -1, 0, 0], [0, 0, 2, 0, 0], [-1, 2, 4, 2, -1], [0, 0, 2, 0, 0], [0, 0, -1, 0, 0]], float),
    "row": np.array([[0, 0, .5, 0, 0], [0, -1, 0, -1, 0], [-1, 4, 5, 4, -1], [0, -1, 0, -1, 0], [0, 0, .5, 0, 0]], float),
    "diag": np.array([[0, 0, -1.5, 0, 0], [0, 2, 0, 2, 0], [-1.5, 0, 6, 0, -1.5], [0, 2, 0, 2, 0], [0, 0, -1.5, 0, 0]], float),
}
MHC_K["col"] = MHC_K["row"].T


def mhc_textbook(cfa, pat, mul):
    """Malvar, He, Cutler (ICASSP 2004) 5x5 filters / 8 on cfa*mul with
    clamp-to-edge, outputs floored at 0."""
    h, w = cfa.shape
    cols = bayer_colors(h, w, pat)
    v = cfa.astype(np.float64) * np.asarray(mul, np.float64)[cols]
    pad = v[clamp_idx(h, -2, h + 2)][:, clamp_idx(w, -2, w + 2)]
    f = {}
    for name, k in MHC_K.items():
        acc = np.zeros((h, w))
        for dy in range(5):
            for dx in range(5):
                if k[dy, dx]:
                    acc += k[dy, dx] * pad[dy:dy + h, dx:dx + w]
        f[name] = acc / 8.0
    out = np.zeros((h, w, 3))
    xs = np.arange(w)[None, :].repeat(h, 0)
    ys = np.arange(h)[:, None].repeat(w, 1)
    row_c = np.asarray(pat)[(ys % 2) * 2 + ((xs + 1) % 2)]
    for c in range(3):
        out[..., c] = np.where(cols == c, v, 0)
    g = cols == 1
    rb = ~g
    out[..., 1] = np.where(rb, f["g_rb"], out[..., 1])
    for c in (0, 2):
        other = 2 - c
        out[..., c] = np.where(cols == other, f["diag"], out[..., c])
        out[..., c] = np.where(g & (row_c == c), f["row"], out[..., c])
        out[..., c] = np.where(g & (row_c != c), f["col"], out[..., c])
    return np.maximum(out, 0)


@pytest.mark.parametrize("pname", list(PATTERNS))
def test_mhc_matches_textbook(oracle_mod, pname):
    O = oracle_mod
    pat = PATTERNS[pname]
    rng = np.random.default_rng(9)
    cfa = rng.uniform(-0.05, 1.0, (21, 30)).astype(F32)
    mul = (1.7, 1.0, 1.3)
    assert maxabs(O.mhc_bayer(cfa, pat, mul), mhc_textbook(cfa, pat, mul)) < 5e-6


# ---------------------------------------------------------------- RCD

def test_rcd_small_is_mhc(oracle_mod):
    """demosaic.rs:149-151: below 16 pixels RCD is MHC."""
    O = oracle_mod
    cfa = frame_cfa(11, 15, 40)
    assert np.array_equal(O.rcd_bayer(cfa), O.mhc_bayer(cfa))


def test_rcd_border_is_mhc(oracle_mod):
    """Pixels within 4 px of the frame edge keep their MHC values: every RCD
    stage writes only inside [4, w-4) x [4, h-4) (demosaic.rs:158), and own-colour
    samples are max(cfa*mul, 0) in both."""
    O = oracle_mod
    cfa = frame_cfa(12, 48, 40)
    rcd, mhc = O.rcd_bayer(cfa), O.mhc_bayer(cfa)
    band = np.ones(cfa.shape, bool)
    band[4:-4, 4:-4] = False
    assert np.array_equal(rcd[band], mhc[band])
    assert not np.array_equal(rcd[~band], mhc[~band])


def test_rcd_transpose_equivariant(oracle_mod):
    """RGGB is its own transpose; RCD treats V and H (and the two diagonals)
    symmetrically, so rcd(cfa.T) == rcd(cfa).T up to float rounding.  A wrong
    index in one direction breaks this."""
    O = oracle_mod
    cfa = frame_cfa(13, 96, 96, iso=4.0)
    a = O.rcd_bayer(cfa)
    b = O.rcd_bayer(np.ascontiguousarray(cfa.T)).transpose(1, 0, 2)
    d = np.abs(a.astype(np.float64) - b)
    assert np.quantile(d, 0.99) < 1e-6
    assert float((d > 1e-4).mean()) < 0.002


def test_rcd_rotate180_equivariant(oracle_mod):
    """Rotating an even-sized RGGB frame by 180 degrees gives BGGR; N/S, E/W,
    NW/SE and NE/SW swap, so the output rotates with it."""
    O = oracle_mod
    cfa = frame_cfa(14, 96, 80, iso=4.0)
    a = O.rcd_bayer(cfa, RGGB)
    b = O.rcd_bayer(np.ascontiguousarray(cfa[::-1, ::-1]), PATTERNS["BGGR"])[::-1, ::-1]
    d = np.abs(a.astype(np.float64) - b)
    assert np.quantile(d, 0.99) < 1e-6
    assert float((d > 1e-4).mean()) < 0.002


def test_rcd_linear_ramp(oracle_mod):
    """A per-channel linear ramp: MHC is exact for linear signals (its
    corrections are Laplacians), and RCD's ratio/colour-difference estimates
    stay close to the truth in the interior."""
    O = oracle_mod
    h, w = 48, 64
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    truth = np.stack([0.2 + 0.004 * xx + 0.002 * yy, 0.3 + 0.003 * xx + 0.003 * yy, 0.25 + 0.001 * xx + 0.005 * yy], -1)
    cols = bayer_colors(h, w, RGGB)
    cfa = np.take_along_axis(truth, cols[..., None], 2)[..., 0].astype(F32)
    mhc = O.mhc_bayer(cfa)
    assert maxabs(mhc[2:-2, 2:-2], truth[2:-2, 2:-2]) < 1e-5
    rcd = O.rcd_bayer(cfa)
    assert maxabs(rcd[6:-6, 6:-6], truth[6:-6, 6:-6]) < 5e-3


def test_rcd_f32_matches_f64(oracle_mod):
    O = oracle_mod
    cfa = frame_cfa(15, 80, 64, iso=2.0)
    a = O.rcd_bayer(cfa)
    b = O.rcd_bayer(cfa.astype(np.float64))
    d = np.abs(a - b)
    assert np.quantile(d, 0.999) < 1e-5


# ---------------------------------------------------------------- noise + NLM

def test_estimate_noise_brute_force(oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(10)
    for (h, w) in ((64, 64), (480, 500), (3, 3), (2, 9)):
        p = rng.random((h, w)).astype(F32)
        step = max(w * h // 200000, 1)
        flat = p.ravel()
        vals = []
        for i in range(0, w * h, step):
            x, y = i % w, i // w
            if x < 1 or y < 1 or x + 1 >= w or y + 1 >= h:
                continue
            lap = F32(4) * flat[i] - flat[i - 1] - flat[i + 1] - flat[i - w] - flat[i + w]
            vals.append(abs(F32(lap)))
        if not vals:
            assert O.estimate_noise(p) == 0.0
            continue
        v = np.sort(np.array(vals, F32))[len(vals) // 2]
        want = F32(F32(v / F32(0.6745)) / np.sqrt(F32(20)))
        assert O.estimate_noise(p) == float(want)


def nlm_brute(p, hp, S, P):
    """Direct patch-distance NLM in float64 with the reference's two-level clamp."""
    p = p.astype(np.float64)
    h, w = p.shape
    h2 = max(hp * hp, 1e-12)
    ys, xs = np.arange(h), np.arange(w)
    acc = np.zeros((h, w))
    ws = np.zeros((h, w))
    for dy in range(-S, S + 1):
        for dx in range(-S, S + 1):
            q = p[np.clip(ys + dy, 0, h - 1)][:, np.clip(xs + dx, 0, w - 1)]
            diff = (p - q) ** 2
            pad = diff[clamp_idx(h, -P, h + P)][:, clamp_idx(w, -P, w + P)]
            dist = np.zeros((h, w))
            for j in range(2 * P + 1):
                for i in range(2 * P + 1):
                    dist += pad[j:j + h, i:i + w]
            dist /= (2 * P + 1) ** 2
            wt = np.exp(-dist / h2)
            acc += wt * q
            ws += wt
    return acc / np.maximum(ws, 1e-12)


@pytest.mark.parametrize("S,P", [(1, 0), (3, 1), (2, 2), (4, 3)])
def test_nlm_brute_force(oracle_mod, S, P):
    O = oracle_mod
    _, noisy = lcg_noise_plane()
    p = np.ascontiguousarray(noisy[:24, 20:52])
    hp = 0.05
    assert maxabs(O.nlm(p, hp, S, P), nlm_brute(p, hp, S, P)) < 2e-6
    assert maxabs(O.nlm(p.astype(np.float64), hp, S, P), nlm_brute(p, hp, S, P)) < 1e-12


def test_nlm_closed_forms(oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(11)
    p = rng.random((20, 27)).astype(F32)
    S = 2
    # h -> inf: the mean of the clamped (2S+1)^2 window
    h, w = p.shape
    win = np.zeros((h, w))
    for dy in range(-S, S + 1):
        for dx in range(-S, S + 1):
            win += p.astype(np.float64)[np.clip(np.arange(h) + dy, 0, h - 1)][:, np.clip(np.arange(w) + dx, 0, w - 1)]
    assert maxabs(O.nlm(p, 1e6, S, 1), win / (2 * S + 1) ** 2) < 2e-6
    # h -> 0: identity where every other patch differs
    assert maxabs(O.nlm(p, 1e-6, S, 1), p) < 1e-7
    # constant in, constant out
    c = np.full((13, 17), 0.37, F32)
    assert maxabs(O.nlm(c, 0.01, 3, 1), c) < 1e-6
    # search < 0: no offsets at all, acc / max(0, 1e-12) = 0
    assert np.array_equal(O.nlm(p, 0.1, -1, 1), np.zeros_like(p))


def test_nlm_thread_chunking_is_rounding_only(oracle_mod):
    """enhance.rs:43-70: the chunk count only changes summation order."""
    O = oracle_mod
    cfa = frame_cfa(16, 96, 64)
    L = O.oklab_from_prophoto(O.rcd_bayer(cfa))[..., 0].copy()
    s = O.estimate_noise(L)
    a = O.nlm(L, 0.6 * s, 10, 3, nthreads=1)
    b = O.nlm(L, 0.6 * s, 10, 3, nthreads=7)
    assert 0 < maxabs(a, b) < 1e-5 or np.array_equal(a, b)


# ---------------------------------------------------------------- guided / denoise / RL

def box_f64(p, r):
    h, w = p.shape
    pad = p[clamp_idx(h, -r, h + r)][:, clamp_idx(w, -r, w + r)]
    acc = np.zeros((h, w))
    for dy in range(2 * r + 1):
        for dx in range(2 * r + 1):
            acc += pad[dy:dy + h, dx:dx + w]
    return acc / (2 * r + 1) ** 2


def guided_textbook(I, p, r, eps):
    """He, Sun, Tang, "Guided Image Filtering" (ECCV 2010), Algorithm 1."""
    I = I.astype(np.float64)
    p = p.astype(np.float64)
    mI, mp = box_f64(I, r), box_f64(p, r)
    cov = box_f64(I * p, r) - mI * mp
    var = box_f64(I * I, r) - mI * mI
    a = cov / (var + eps)
    b = mp - a * mI
    return box_f64(a, r) * I + box_f64(b, r)


def test_guided_filter_textbook(oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(12)
    I = rng.random((33, 41)).astype(F32)
    p = (0.5 * I + 0.1 * rng.random((33, 41))).astype(F32)
    for r, eps in ((2, 1e-3), (5, 1e-2)):
        assert maxabs(O.guided_filter(I, p, r, eps), guided_textbook(I, p, r, eps)) < 2e-5


def test_denoise_composition(oracle_mod):
    """enhance.rs:74-92 rebuilt from the pinned primitives."""
    O = oracle_mod
    cfa = frame_cfa(17, 72, 56, iso=8.0)
    rgb = O.rcd_bayer(cfa)
    for amount, S, P in ((0.0, 10, 3), (37.0, 3, 1)):
        lab = O.oklab_from_prophoto(rgb)
        L, A, B = (np.ascontiguousarray(lab[..., i]) for i in range(3))
        sigma = max(O.estimate_noise(L), F32(1e-4))
        k = F32(amount) / F32(100)
        lf = O.nlm(L, F32(sigma) * (F32(0.6) + F32(1.6) * k), S, P)
        l2 = (L + (lf - L) * (F32(0.5) + F32(0.5) * k)).astype(F32)
        r = int(max((F32(72) / F32(600)) * (F32(1) + F32(2) * k), F32(2)))
        eps = F32(F32(sigma) * F32(2)) ** 2
        a2 = O.guided_filter(l2, A, r, eps)
        b2 = O.guided_filter(l2, B, r, eps)
        want = np.maximum(O.oklab_to_prophoto(np.stack([l2, a2, b2], -1).astype(F32)), 0)
        assert np.array_equal(O.denoise(rgb, amount, S, P), want)


def rl_textbook(rgb, sigma, iters):
    """Richardson (1972) / Lucy (1974) with the reference's clamps
    (enhance.rs:96-119); the PSF is symmetric, so correlation = convolution."""
    rgb = rgb.astype(np.float64)
    y = np.maximum(rgb @ np.array([0.2880402, 0.7118741, 0.0000857]), 1e-6)
    est = y.copy()
    for _ in range(iters):
        conv = gauss_f64(est, sigma)
        corr = gauss_f64(y / np.maximum(conv, 1e-6), sigma)
        est = np.clip(est * corr, 0, 64)
    return rgb * np.clip(est / y, 0.6, 1.6)[..., None]


def test_raw_details_textbook(oracle_mod):
    O = oracle_mod
    cfa = frame_cfa(18, 64, 48, iso=2.0)
    rgb = O.rcd_bayer(cfa)
    for sigma, iters in ((1.0, 20), (0.75, 8)):
        assert maxabs(O.raw_details(rgb, sigma, iters), rl_textbook(rgb, sigma, iters)) < 2e-5
    assert np.array_equal(O.raw_details(rgb, 1.0, 0), rgb)
    flat = np.full((20, 30, 3), 0.25, F32)
    assert maxabs(O.raw_details(flat, 1.0, 20), flat) < 1e-6


def test_fp64_twin_agrees(oracle_mod):
    """The FP64 oracle is the same algorithm: per-stage FP32-vs-FP64 error stays small."""
    O = oracle_mod
    cfa = frame_cfa(19, 64, 64, iso=4.0)
    rgb32 = O.rcd_bayer(cfa)
    rgb64 = rgb32.astype(np.float64)
    assert maxabs(O.raw_details(rgb32, 1.0, 20), O.raw_details(rgb64, 1.0, 20)) < 2e-5
    assert maxabs(O.denoise(rgb32, 0.0, 10, 3), O.denoise(rgb64, 0.0, 10, 3)) < 5e-5


def test_reference_denoise_reproducibility_envelope(oracle_mod):
    """The reference's denoise is not reproducible to 2e-5 across rayon pool
    sizes: NLM's chunk order (enhance.rs:43-70) moves L by ~1e-6 and the
    guided filter (image.rs:396-398, var = E[I^2] - E[I]^2) amplifies that by
    two orders of magnitude.  Its sub-stages on identical inputs are exact."""
    O = oracle_mod
    cfa = frame_cfa(400, 256, 192, iso=8.0)
    rgb = O.rcd_bayer(cfa)
    d1 = O.denoise(rgb, 0.0, 10, 3, nthreads=1)
    d8 = O.denoise(rgb, 0.0, 10, 3, nthreads=8)
    assert maxabs(d1, d8) > 2e-5
    L = np.ascontiguousarray(O.oklab_from_prophoto(rgb)[..., 0])
    s = O.estimate_noise(L)
    assert maxabs(O.nlm(L, 0.6 * s, 10, 3, nthreads=1), O.nlm(L, 0.6 * s, 10, 3, nthreads=8)) < 5e-6
    A = np.ascontiguousarray(O.oklab_from_prophoto(rgb)[..., 1])
    assert np.array_equal(O.guided_filter(L, A, 2, 1e-4), O.guided_filter(L, A, 2, 1e-4))


# ------------------------------------------------------------------ auto_wb_xy statistics (pipeline/mod.rs:421-441)

def test_auto_wb_stats_uniform_closed_form(oracle_mod):
    """A uniform image: every sampled pixel adds ln(raw_c) sqrt(max); the neutral is raw_c / raw_g."""
    O = oracle_mod
    h, w = 7, 11
    v = np.array([0.3, 0.5, 0.2], np.float32)
    nas = np.array([0.8, 1.0, 1.25], np.float32)
    rgb = np.broadcast_to(v, (h, w, 3)).copy()
    acc, n, neutral = O.auto_wb_stats(rgb, nas)
    cnt = len(range(0, h * w, 3))
    raw = (v * nas).astype(np.float32).astype(np.float64)
    wt = np.sqrt(raw.max())
    assert abs(n - cnt * wt) <= 1e-12 * n
    assert np.allclose(acc, cnt * np.log(raw) * wt, rtol=1e-12, atol=0)
    assert np.allclose(neutral, [raw[0] / raw[1], 1.0, raw[2] / raw[1]], rtol=2e-7)


def test_auto_wb_stats_sampling_and_thresholds(oracle_mod):
    """Only every third pixel counts, and pixels with max > 0.9 or min < 0.002 are skipped."""
    O = oracle_mod
    rng = np.random.default_rng(3)
    rgb = rng.uniform(0.05, 0.6, (9, 10, 3)).astype(np.float32)
    nas = np.ones(3, np.float32)
    base = O.auto_wb_stats(rgb, nas)
    flat = rgb.reshape(-1, 3)
    other = rgb.copy().reshape(-1, 3)
    other[1::3] = rng.uniform(0.05, 0.6, other[1::3].shape)  # unsampled pixels do not matter
    other[2::3] = 0.95
    o2 = O.auto_wb_stats(other.reshape(rgb.shape), nas)
    assert np.array_equal(base[0], o2[0]) and base[1] == o2[1]
    # a sampled pixel pushed over 0.9 (or under 0.002) drops out: the sums equal those without it
    cut = flat.copy()
    cut[0] = [0.95, 0.5, 0.5]
    cut[3] = [0.5, 0.001, 0.5]
    acc_c, n_c, _ = O.auto_wb_stats(cut.reshape(rgb.shape), nas)
    keep = flat[6::3]
    m = keep.max(axis=1).astype(np.float64)
    assert abs(n_c - np.sum(np.sqrt(m))) <= 1e-12 * n_c
    # nothing valid -> no neutral
    assert O.auto_wb_stats(np.full((4, 4, 3), 0.95, np.float32), nas)[2] is None
