"""Training of the learned luminance denoiser on the GPU.

Pairs (noisy network input, NLM target) come from seeded synthetic Bayer
frames that the pipeline itself renders, demosaics, converts to Oklab and
denoises with NLM (ds_cnn_make_pairs_dev).  The forward, backward and Adam
kernels are the library's own (ds_cnn_train_step_dev).

    python -m darkslide_amd.train_cnn --crops 200000 --out darkslide_amd/weights/cnn_l32x8.dsck

trains the configured schedule: 128 x 128 crops from 512 x 512 frames at
noise levels (iso) drawn log-uniformly from [1, 16], Adam with lr 1e-3.
"""
import argparse
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_CHECKPOINT = os.path.join(HERE, "weights", "cnn_l32x8.dsck")
MARGIN = 8  # receptive-field radius of the 8 conv3x3 layers: the loss ignores the zero-padded border


class PairPool:
    """Device-resident pool of training crops, refilled frame by frame."""

    def __init__(self, ctx, frames, crops_per_frame, crop, frame, rng, iso_range=(1.0, 16.0), search=10, patch=3):
        self.ctx = ctx
        self.frames = frames
        self.cpf = crops_per_frame
        self.crop = crop
        self.frame = frame
        self.rng = rng
        self.iso_range = iso_range
        self.search = search
        self.patch = patch
        self.n = frames * crops_per_frame
        self.px = crop * crop
        self.bx = ctx.alloc(self.n * self.px * 4)
        self.bt = ctx.alloc(self.n * self.px * 4)
        self.generated = 0

    def refill(self):
        import darkslide_amd as D
        L = D.lib()
        lo, hi = math.log(self.iso_range[0]), math.log(self.iso_range[1])
        for f in range(self.frames):
            seed = int(self.rng.integers(1, 2 ** 62))
            iso = float(math.exp(self.rng.uniform(lo, hi)))
            off = f * self.cpf * self.px * 4
            self.ctx._check(L.ds_cnn_make_pairs_dev(self.ctx.handle, seed, self.frame, self.frame, iso, self.search, self.patch, self.crop,
                                                    self.cpf, self.bx.at(off), self.bt.at(off)))
        self.generated += self.n

    def batch_ptrs(self, start):
        off = start * self.px * 4
        return self.bx.at(off), self.bt.at(off)

    def free(self):
        self.bx.free()
        self.bt.free()


def train(ctx, net, steps, batch=16, crop=128, frame=512, pool_frames=256, crops_per_frame=8, lr=1e-3, seed=1, log=None,
          log_every=50, iso_range=(1.0, 16.0)):
    """Runs `steps` Adam steps; returns the loss history (one value per log point)."""
    rng = np.random.default_rng(seed)
    if batch % crops_per_frame and crops_per_frame % batch:
        raise ValueError("batch and crops per frame must divide one another")
    pool = PairPool(ctx, pool_frames, crops_per_frame, crop, frame, rng, iso_range)
    hist = []
    try:
        steps_per_pool = max(1, pool.n // batch)
        order = []
        t0 = time.time()
        for s in range(steps):
            if not order:
                pool.refill()
                order = list(rng.permutation(steps_per_pool))
            start = order.pop() * batch
            px, pt = pool.batch_ptrs(start)
            want = (s % log_every == 0) or s == steps - 1
            loss = net.train_step_dev(px, pt, batch, crop, crop, MARGIN, lr, want_loss=want)
            if want:
                hist.append(loss)
                if log:
                    log(f"step {s + 1}/{steps} loss {loss:.5f} crops {pool.generated} {time.time() - t0:.1f} s")
    finally:
        pool.free()
    return hist


def evaluate(ctx, net, seeds, size=512, iso=8.0):
    """Mean PSNR gains (dB) over the noisy luminance against the noise-free
    frame's luminance, for the network and for NLM (h = 0.6 sigma, 21 x 21
    search, 7 x 7 patch)."""
    import darkslide_amd as D
    L = D.lib()
    res = {"cnn": [], "nlm": [], "noisy_psnr": []}
    for sd in seeds:
        planes = []
        for lvl in (iso, 0.0):
            buf = ctx.alloc(size * size * 2)
            ctx._check(L.ds_synth_bayer_dev(ctx.handle, sd, size, size, 14, float(lvl), D._pat(D.RGGB), buf.ptr))
            raw = buf.download((size, size), np.uint16)
            buf.free()
            rgb, _ = ctx.demosaic_u16(raw)
            planes.append(ctx.oklab_from_prophoto(rgb)[..., 0].copy())
        Ln, Lc = planes
        s = max(np.float32(ctx.estimate_noise(Ln)), np.float32(1e-4))
        p0 = psnr(Ln, Lc)
        res["noisy_psnr"].append(p0)
        res["cnn"].append(psnr(net.denoise_plane(Ln), Lc) - p0)
        res["nlm"].append(psnr(ctx.nlm(Ln, s * np.float32(0.6), 10, 3), Lc) - p0)
    return {k: float(np.mean(v)) for k, v in res.items()}


def psnr(a, ref):
    mse = float(np.mean((np.asarray(a, np.float64) - np.asarray(ref, np.float64)) ** 2))
    return 10.0 * math.log10(1.0 / max(mse, 1e-30))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--crops", type=int, default=200000, help="training crops generated in total")
    ap.add_argument("--crop", type=int, default=128)
    ap.add_argument("--frame", type=int, default=512)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--epochs-per-pool", type=int, default=1)
    ap.add_argument("--pool-frames", type=int, default=256)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--init-seed", type=int, default=1)
    ap.add_argument("--resume", default=None)
    ap.add_argument("--out", default=DEFAULT_CHECKPOINT)
    args = ap.parse_args()

    sys.path.insert(0, os.path.dirname(HERE))
    import darkslide_amd as D

    def log(*a):
        print(*a, flush=True)

    ctx = D.Context(0)
    net = D.CNN(ctx, seed=args.init_seed)
    if args.resume:
        net.load(args.resume)
    cpf = 8
    crops_per_pool = args.pool_frames * cpf
    pools = max(1, args.crops // crops_per_pool)
    steps = pools * args.epochs_per_pool * (crops_per_pool // args.batch)
    log(f"training {steps} steps: {pools} pools of {crops_per_pool} crops {args.crop}x{args.crop}, batch {args.batch}")
    before = evaluate(ctx, net, [0xE0 + i for i in range(4)])
    log(f"before: {before}")
    t0 = time.time()
    hist = train(ctx, net, steps, batch=args.batch, crop=args.crop, frame=args.frame, pool_frames=args.pool_frames,
                 crops_per_frame=cpf, lr=args.lr, seed=args.seed, log=log, log_every=200)
    log(f"trained in {time.time() - t0:.1f} s; loss {hist[0]:.5f} -> {hist[-1]:.5f}")
    after = evaluate(ctx, net, [0xE0 + i for i in range(4)])
    log(f"after: {after}")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    net.save(args.out)
    log(f"saved {args.out} (step {net.step})")
    net.close()
    ctx.close()


if __name__ == "__main__":
    main()
