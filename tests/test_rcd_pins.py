"""Closed-form pins of the RCD restatement (demosaic.rs:145-344) that the
plausible transcription slips fail, with each slip planted in a copy of the
oracle to show that it does.

RCD's discrete choices are where a slip hides: which of the vertical and
horizontal estimates to trust (vh, demosaic.rs:173-184, and disc,
:216-220), which of the two diagonals (pq, :262-273), and which neighbour's
gradient weights which estimate (:226-241, :285-296, :318-333).  A slip that
is symmetric under transpose and 180-degree rotation (swapping vd and
1 - vd, or both gradient pairs) passes the equivariance pins of
test_oracle_pins.py.  The scenes here decide each choice and have a known
answer:

* horizontal stripes (every channel constant along x): bh = 0 exactly, so
  vh -> 1 and RCD must interpolate along rows.  G at R/B sites is the row's
  green (the W/E estimates coincide, g * 2 lpf / (eps + 2 lpf) with lpf the
  same at x and x +- 2), R at the G sites of R rows is the row's red, B at
  the G sites of B rows the row's blue.  Vertical stripes: the transpose,
  vh -> 0.
* diagonal stripes with R = B (functions of x + y, or of x - y): along the
  stripe bq (or bp) is 0 exactly, so pq -> 1 (or 0), and R at B / B at R
  comes from the two diagonal neighbours on the same stripe.  Every
  non-green site of a stripe sees the same neighbourhood (R = B), so their
  green estimates are equal and the colour difference gives the stripe's
  value exactly.
* strong vertical stripes with a weak green step across rows y0 (y >= y0
  gets +delta): at R/B sites in rows y0-3, y0-2, y0+1, y0+2 one vertical
  side is flat (n_grad or s_grad = eps) and its ratio-corrected estimate
  is exact, so G must come from that side; at the G sites of rows y0-2 ..
  y0+1 whose vertical neighbours are true R (or B) samples, one of n_g / s_g
  is eps and R (B) must come from that side.  Transposed: the W/E roles.

* diagonal stripes with R = B and a green step across them: four to six
  diagonal steps from the step one diagonal side's gradient is eps, so
  R at B / B at R must come from the undisturbed side (ne_g/sw_g and
  nw_g/se_g roles).
* R and G constant, B varying along one axis: vh is exactly 0.5 at R sites
  and decided at their diagonal B neighbours, so G is exact only if disc
  takes the decided value.

Tolerances: the opposite-direction weights (eps / energy; ~1e-4 of vh next
to the weak green step) and the eps terms leave at most ~2e-4 (measured:
<= 2.8e-6 on every pin but the two with a weak step, 1.8e-4 there); every
planted slip moves some pinned value by at least 5x its tolerance (measured
>= 1.7e-2).  disc's tie (|0.5 - central| == |0.5 - neigh| with
central != neigh) needs mirror-image values that no scene here forces, so a
< / <= slip there is not pinned.
"""
import concurrent.futures as cf
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import RGGB

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SRC = os.path.join(ROOT, "oracle", "oracle.cpp")

H = W = 64
M = 12  # pinned sites keep this far from the frame (RCD reads up to +-5 around a site, MHC borders are 4 wide)


def colours(h, w):
    yy, xx = np.mgrid[0:h, 0:w]
    return np.array(RGGB)[(yy & 1) * 2 + (xx & 1)]


def mosaic(truth):
    cols = colours(*truth.shape[:2])
    return np.ascontiguousarray(np.take_along_axis(truth, cols[..., None], 2)[..., 0].astype(np.float32))


def interior(h, w, m=M):
    k = np.zeros((h, w), bool)
    k[m:h - m, m:w - m] = True
    return k


def site_masks(h, w):
    yy, xx = np.mgrid[0:h, 0:w]
    return {
        "R": (yy % 2 == 0) & (xx % 2 == 0),
        "B": (yy % 2 == 1) & (xx % 2 == 1),
        "G_Rrow": (yy % 2 == 0) & (xx % 2 == 1),  # G between R samples horizontally, B vertically
        "G_Brow": (yy % 2 == 1) & (xx % 2 == 0),  # G between B samples horizontally, R vertically
    }


def stripes_horizontal(seed):
    """Every row its own (r, g, b): checks along rows."""
    rng = np.random.default_rng(seed)
    rows = rng.uniform(0.05, 0.95, (H, 3))
    truth = np.broadcast_to(rows[:, None, :], (H, W, 3)).copy()
    s, k = site_masks(H, W), interior(H, W)
    checks = [((s["R"] | s["B"]) & k, 1), (s["G_Rrow"] & k, 0), (s["G_Brow"] & k, 2)]
    return truth, checks


def transpose_pin(pin):
    truth, checks = pin
    return np.ascontiguousarray(truth.transpose(1, 0, 2)), [(m.T.copy(), ch) for m, ch in checks]


def stripes_diagonal(seed, anti):
    """R = B = f(s), G = g(s), s = x + y (or x - y): R at B and B at R."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W]
    s = (xx - yy + H) if anti else (xx + yy)
    f = rng.uniform(0.05, 0.95, 2 * (H + W))
    g = rng.uniform(0.05, 0.95, 2 * (H + W))
    truth = np.stack([f[s], g[s], f[s]], -1)
    m, k = site_masks(H, W), interior(H, W)
    return truth, [(m["B"] & k, 0), (m["R"] & k, 2)]


def stripes_with_step(seed, delta=0.03, y0=32):
    """Vertical stripes alternating high / low every two columns (strong
    horizontal energy), green raised by delta from row y0 on."""
    rng = np.random.default_rng(seed)
    x = np.arange(W)
    hi_lo = lambda ph: np.where(((x + ph) // 2) % 2 == 0, 0.85, 0.1) + rng.uniform(-0.02, 0.02, W)
    cols = np.stack([hi_lo(0), hi_lo(1), hi_lo(0)], -1)
    truth = np.broadcast_to(cols[None, :, :], (H, W, 3)).copy()
    truth[y0:, :, 1] += delta
    yy, xx = np.mgrid[0:H, 0:W]
    s, k = site_masks(H, W), interior(H, W)
    rows = lambda *ys: np.isin(yy, [y0 + d for d in ys])
    checks = [((s["R"] | s["B"]) & rows(-3, -2, 1, 2) & k, 1),  # G from the flat vertical side
              (s["G_Brow"] & rows(-1, 1) & k, 0),              # R at G between true R samples vertically
              (s["G_Rrow"] & rows(-2, 0) & k, 2)]              # B at G between true B samples vertically
    return truth, checks


def diagonal_stripes_with_step(seed, anti, delta=0.2):
    """R = B = f(s) along one diagonal (s = x + y, or x - y), G = g(s) raised
    by delta where the other diagonal coordinate d reaches d0.  The green
    step leaves bq (bp) at R/B sites at 0 (the stripe's samples there are
    all R/B), so pq stays decided; four to six diagonal steps from it, the
    green estimates of the centre and of the far-side neighbours are those
    of the undisturbed stripe, while the near-side ones are not, so one of
    the two gradients (ne_g / sw_g, or nw_g / se_g) is eps and R at B / B at R
    must come from the undisturbed side, giving f(s)."""
    rng = np.random.default_rng(seed)
    yy, xx = np.mgrid[0:H, 0:W]
    s = (xx - yy + H) if anti else (xx + yy)
    d = (xx + yy) if anti else (xx - yy + H)
    d0 = H
    f = rng.uniform(0.05, 0.95, 2 * (H + W))
    g = rng.uniform(0.05, 0.7, 2 * (H + W))
    truth = np.stack([f[s], g[s] + delta * (d >= d0), f[s]], -1)
    m, k = site_masks(H, W), interior(H, W)
    near = np.isin(d, [d0 - 6, d0 - 4, d0 + 4, d0 + 6])
    return truth, [(m["B"] & near & k, 0), (m["R"] & near & k, 2)]


def blue_stripes(seed):
    """R and G constant, B a function of y.  An R site's own row and column
    hold only R and G samples, so bv = bh = 0 there and vh = eps / 2 eps =
    0.5 exactly: undecided.  Its diagonal neighbours are B sites, whose
    columns vary and rows do not, so their vh -> 1.  disc (demosaic.rs:
    216-220) must take the decided neighbour average and interpolate G along
    the row, where the ratio correction cancels (lpf is the same at x and
    x +- 2): G = g0.  At B sites disc keeps their own decided vh."""
    rng = np.random.default_rng(seed)
    b = rng.uniform(0.05, 0.95, H)
    truth = np.empty((H, W, 3))
    truth[..., 0], truth[..., 1] = 0.4, 0.5
    truth[..., 2] = b[:, None]
    m, k = site_masks(H, W), interior(H, W)
    return truth, [((m["R"] | m["B"]) & k, 1)]


def pins():
    p = {
        "h-stripes": stripes_horizontal(1),
        "v-stripes": transpose_pin(stripes_horizontal(2)),
        "diag(x+y)": stripes_diagonal(3, anti=False),
        "diag(x-y)": stripes_diagonal(4, anti=True),
        "v-stripes+h-step": stripes_with_step(5),
        "h-stripes+v-step": transpose_pin(stripes_with_step(6)),
        "diag(x+y)+step": diagonal_stripes_with_step(7, anti=False),
        "diag(x-y)+step": diagonal_stripes_with_step(8, anti=True),
        "B-rows (disc)": blue_stripes(9),
        "B-columns (disc)": transpose_pin(blue_stripes(10)),
    }
    return p


TOL = {"h-stripes": 1e-4, "v-stripes": 1e-4, "diag(x+y)": 1e-4, "diag(x-y)": 1e-4, "v-stripes+h-step": 1e-3,
       "h-stripes+v-step": 1e-3, "diag(x+y)+step": 1e-4, "diag(x-y)+step": 1e-4,
       "B-rows (disc)": 1e-4, "B-columns (disc)": 1e-4}


def pin_errors(rcd):
    out = {}
    for name, (truth, checks) in pins().items():
        rgb = rcd(mosaic(truth)).astype(np.float64)
        out[name] = max(float(np.abs(rgb[m, ch] - truth[m, ch]).max()) for m, ch in checks)
    return out


def test_pin_scenes_are_well_posed():
    for name, (truth, checks) in pins().items():
        assert truth.min() >= 0.0 and truth.max() <= 1.0, name
        assert all(m.sum() >= 20 for m, _ in checks), name


def test_rcd_pins(oracle_mod):
    errs = pin_errors(oracle_mod.rcd_bayer)
    print(errs)
    for name, e in errs.items():
        assert e <= TOL[name], (name, e)


# ---------------------------------------------------------------- planted slips

# (name, the oracle's text, the slip) -- each a plausible misreading of the
# reference's expression at the cited line
MUTATIONS = [
    ("step3 vd <-> 1-vd (:244)", "g[i] = vmax(vd * h_est + ((T)1 - vd) * v_est, (T)0);",
     "g[i] = vmax(((T)1 - vd) * h_est + vd * v_est, (T)0);"),
    ("vh = hs / (vs + hs) (:182)", "vh[i] = vs / (vs + hs);", "vh[i] = hs / (vs + hs);"),
    ("step4.3 pd <-> 1-pd (:304)", "target[i] = vmax(g[i] + ((T)1 - pd) * p_est + pd * q_est, (T)0);",
     "target[i] = vmax(g[i] + pd * p_est + ((T)1 - pd) * q_est, (T)0);"),
    ("pq = qs / (ps + qs) (:271)", "pq[i] = ps / (ps + qs);", "pq[i] = qs / (ps + qs);"),
    ("step3 N/S gradient roles (:238)", "T v_est = (s_grad * n_est + n_grad * s_est) / (n_grad + s_grad);",
     "T v_est = (n_grad * n_est + s_grad * s_est) / (n_grad + s_grad);"),
    ("step3 W/E gradient roles (:239)", "T h_est = (w_grad * e_est + e_grad * w_est) / (e_grad + w_grad);",
     "T h_est = (e_grad * e_est + w_grad * w_est) / (e_grad + w_grad);"),
    ("step3 n_est ratio from the south lpf (:234)", "T n_est = c[i - w1] * (T)2 * lpf[i] / (eps + lpf[i] + lpf[i - w2]);",
     "T n_est = c[i - w1] * (T)2 * lpf[i] / (eps + lpf[i] + lpf[i + w2]);"),
    ("step4.4 vd <-> 1-vd (:337)", "target[i] = vmax(g[i] + ((T)1 - vd) * v_est + vd * h_est, (T)0);",
     "target[i] = vmax(g[i] + vd * v_est + ((T)1 - vd) * h_est, (T)0);"),
    ("step4.4 N/S gradient roles (:333)", "T v_est = (n_g * s_e + s_g * n_e) / (n_g + s_g);",
     "T v_est = (s_g * s_e + n_g * n_e) / (n_g + s_g);"),
    ("step4.4 W/E gradient roles (:334)", "T h_est = (e_g * w_e + w_g * e_e) / (e_g + w_g);",
     "T h_est = (w_g * w_e + e_g * e_e) / (e_g + w_g);"),
    ("step4.3 NW/SE gradient roles (:300)", "T p_est = (nw_g * se_e + se_g * nw_e) / (nw_g + se_g);",
     "T p_est = (se_g * se_e + nw_g * nw_e) / (nw_g + se_g);"),
    ("step4.3 NE/SW gradient roles (:301)", "T q_est = (ne_g * sw_e + sw_g * ne_e) / (ne_g + sw_g);",
     "T q_est = (sw_g * sw_e + ne_g * ne_e) / (ne_g + sw_g);"),
    ("step4.3 q taken along the P diagonal (:297)", "T ne_e = cc[i - w1 + 1] - g[i - w1 + 1];",
     "T ne_e = cc[i - w1 - 1] - g[i - w1 - 1];"),
    ("disc keeps the less decided value (:219)",
     "return std::fabs((T)0.5 - central) < std::fabs((T)0.5 - neigh) ? neigh : central;",
     "return std::fabs((T)0.5 - central) < std::fabs((T)0.5 - neigh) ? central : neigh;"),
    ("step4.4 north colour difference from the south (:329)", "T n_e = cc[i - w1] - g[i - w1];", "T n_e = cc[i + w1] - g[i + w1];"),
]


def build_mutant(idx, out_dir):
    name, old, new = MUTATIONS[idx]
    src = open(ORACLE_SRC).read()
    assert src.count(old) == 1, f"{name}: the oracle text to mutate was not found exactly once"
    path = os.path.join(out_dir, f"mut{idx}.cpp")
    with open(path, "w") as f:
        f.write(src.replace(old, new))
    so = os.path.join(out_dir, f"mut{idx}.so")
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared", path, "-o", so, "-lm"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return so


@pytest.fixture(scope="module")
def mutants(tmp_path_factory):
    d = str(tmp_path_factory.mktemp("rcd_mutants"))
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        return list(ex.map(lambda i: build_mutant(i, d), range(len(MUTATIONS))))


def rcd_from(so_path):
    lib = ctypes.CDLL(so_path)
    f = lib.or_rcd_bayer_f32
    f.restype = None

    def rcd(cfa):
        cfa = np.ascontiguousarray(cfa, np.float32)
        h, w = cfa.shape
        out = np.empty((h, w, 3), np.float32)
        f(ctypes.c_void_p(cfa.ctypes.data), ctypes.c_size_t(w), ctypes.c_size_t(h), (ctypes.c_uint8 * 4)(*RGGB),
          ctypes.c_void_p(np.ones(3, np.float32).ctypes.data), ctypes.c_void_p(out.ctypes.data))
        return out

    return rcd


@pytest.mark.slow
@pytest.mark.parametrize("idx", range(len(MUTATIONS)), ids=[m[0] for m in MUTATIONS])
def test_pins_catch_planted_slip(mutants, idx):
    errs = pin_errors(rcd_from(mutants[idx]))
    caught = {k: e for k, e in errs.items() if e > 5 * TOL[k]}
    print(MUTATIONS[idx][0], errs)
    assert caught, f"no pin moved by more than 5x its tolerance: {errs}"


@pytest.mark.gpu
def test_rcd_pins_on_gpu(ctx, oracle_mod):
    """The HIP RCD kernel on the pin scenes: the closed forms hold, and the
    output equals the oracle's bit for bit."""
    for name, (truth, checks) in pins().items():
        cfa = mosaic(truth)
        got = ctx.demosaic(cfa, clip=False)
        assert np.array_equal(got, oracle_mod.rcd_bayer(cfa)), name
    errs = pin_errors(lambda cfa: ctx.demosaic(cfa, clip=False))
    for name, e in errs.items():
        assert e <= TOL[name], (name, e)
