#!/usr/bin/env python3
"""Throughput of the raw detail path on one MI355X.

A step is one 6000x4000 (24 MP) RGGB uint16 14-bit synthetic frame, already
resident in HBM, through the plain pipeline: RCD demosaic (with the clip
plane) -> denoise (Oklab, noise estimate, NLM with 7x7 patches over a 21x21
search at h = 0.6 sigma, guided chroma filters) -> RL raw details (sigma 1.0,
20 iterations), output RGB left in HBM.  W untimed warm-up steps, then K
timed steps between two device synchronisations.  Per-kernel HIP events on
the library's stream give the dominant kernel's average duration for the
roofline.  The CPU baseline is the C++ restatement of the reference path
(oracle/) on a bounded 1024x1024 sample, on all host cores.

Prints one JSON line on stdout.
"""
import argparse
import ctypes
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MP/s demosaic+denoise+RL at 24 MP; learned vs plain NLM at ≤0.3 dB PSNR gap"
PEAK_HBM_GBS = 8000.0
PEAK_FP32_TFLOPS = 157.3
NLM_FLOP_PER_PX = 441 * 12      # per (pixel, offset): diff 2, two window updates 4, scale 1, /h^2 1, exp 1, fma 2, add 1
DEMOSAIC_BYTES_PER_PX = 18      # u16 in + f32 RGB out + f32 clip out
# HBM bytes per pixel that k_nlm_tile moves per launch in denoise (its output
# stage forms the guided pair's four guide planes from L, a and b), from
# rocprofv3 FETCH_SIZE / WRITE_SIZE runs calibrated on a 1 GiB copy
# (profiles/hbm_traffic_counters.md)
NLM_TRAFFIC_BYTES_PER_PX = 41.09
MHC_BYTES_PER_PX = 14
# reconstruct_highlights, algorithmic HBM bytes per pixel when something is
# clipped: flag scan 4, near-clip box 16, weights 32, three-box gaussian of
# three planes 3 x 2 x 3 x 8 = 144, apply 40
HIGHLIGHTS_BYTES_PER_PX = 236
HL_MUL = (2.0, 1.0, 1.5)
# Algorithmic HBM bytes per pixel of the detail and export stages: each
# elementwise pass reads and writes its planes once, each separable gaussian
# pass reads and writes one plane.  Sharpen (masking 0), super_res_refine and
# output_sharpen: luminance 16, gaussian 2 x 8, ratio 32 (RGB in and out, two
# planes in).  Noise reduction at the reference's defaults (colour 25, so the
# chroma fast guided filters at sub 4): Oklab split 24, the downscale of L, a
# and b 12, the upsampled a I + b mixed into a and b 20 (L, a, b in; a, b out),
# the recomposition 24; the guided filters on the 1/16-size planes are left out.
DETAIL_BYTES_PER_PX = {"sharpen": 64, "super_res_refine": 64, "output_sharpen": 64, "noise_reduction": 80}


def saturate(raw, seed, discs=12):
    """Plants saturated discs in an RGGB uint16 frame: every site, green
    sites only, red sites only, in turn."""
    import numpy as np
    h, w = raw.shape
    rng = np.random.default_rng(seed)
    rad = max(3, min(w, h) // 16)
    for k in range(discs):
        cx, cy = int(rng.integers(0, w)), int(rng.integers(0, h))
        x0, x1, y0, y1 = max(cx - rad, 0), min(cx + rad + 1, w), max(cy - rad, 0), min(cy + rad + 1, h)
        yy, xx = np.mgrid[y0:y1, x0:x1]
        m = (xx - cx) ** 2 + (yy - cy) ** 2 <= rad * rad
        if k % 3 == 1:
            m &= ((xx + yy) & 1) == 1  # green sites of RGGB
        elif k % 3 == 2:
            m &= ((xx & 1) == 0) & ((yy & 1) == 0)  # red sites
        raw[y0:y1, x0:x1][m] = 16383
    return raw


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(size, reps):
    """The oracle's FP32 restatement of the same pipeline on a size x size frame."""
    import numpy as np
    import darkslide_amd as D
    import oracle as O
    O.build()
    threads = O.threads()
    raw = D.synth_bayer(0x5EED0000, size, size)

    def run():
        cfa = O.cfa_normalize_u16(raw, D.RGGB, 512.0, [16383.0] * 3)
        rgb, _ = O.demosaic(cfa)
        dn = O.denoise(rgb, 0.0, 10, 3, nthreads=threads)
        return O.raw_details(dn, 1.0, 20)

    run()
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        run()
        times.append(time.perf_counter() - t0)
    t = statistics.median(times)
    return {
        "value": round(size * size / 1e6 / t, 4),
        "unit": "MP/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{size}x{size} synthetic RGGB frame through the same demosaic->denoise->RL pipeline "
                  f"(C++ FP32 restatement of the reference, OpenMP), median of {reps} after 1 warm-up; "
                  f"{t:.2f} s per frame",
    }


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_stages(w, h):
    """Each stage of the pipeline once on the CPU at the full frame size
    (the oracle's FP32 restatement, OpenMP over all host cores, NLM offsets
    chunked as rayon chunks them): RCD, MHC, denoise, RL."""
    import numpy as np
    import darkslide_amd as D
    import oracle as O
    threads = O.threads()
    raw = D.synth_bayer(0x5EED0001, w, h)
    out = {}
    t0 = time.perf_counter()
    cfa = O.cfa_normalize_u16(raw, D.RGGB, 512.0, [16383.0] * 3)
    rgb = O.rcd_bayer(cfa)
    O.clip_plane(cfa)
    out["demosaic_rcd"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.mhc_bayer(O.cfa_normalize_u16(raw, D.RGGB, 512.0, [16383.0] * 3))
    out["demosaic_mhc"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    dn = O.denoise(rgb, 0.0, 10, 3, nthreads=threads)
    out["denoise"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.raw_details(dn, 1.0, 20)
    out["raw_details"] = time.perf_counter() - t0
    n = w * h
    return {k: {"s": round(v, 3), "mp_s": round(n / 1e6 / v, 3)} for k, v in out.items()}


def rcd_config1(ctx, size=1024, reps=20):
    """Configuration 1: RCD demosaic of one 1024x1024 frame, GPU kernel time
    (HIP events, device-resident) beside the CPU restatement's stage time
    (median of 3 after 1 warm-up)."""
    import numpy as np
    import darkslide_amd as D
    import oracle as O
    L = D.lib()
    n = size * size
    raw_h = D.synth_bayer(0x5EED0000, size, size)
    raw, rgb, clip = ctx.alloc(n * 2), ctx.alloc(n * 12), ctx.alloc(n * 4)
    raw.upload(raw_h)
    pat, white, ones = D._pat(D.RGGB), D._f3((16383.0,) * 3), D._f3((1.0, 1.0, 1.0))
    go = lambda: ctx._check(L.ds_demosaic_u16_dev(ctx.handle, raw.ptr, size, size, size, pat, 512.0, white, ones, D.ALGO_RCD,
                                                  rgb.ptr, clip.ptr))
    for _ in range(3):
        go()
    ctx.sync()
    ctx.prof_enable(True)
    ctx.prof_reset()
    for _ in range(reps):
        go()
    gpu_ms = ctx.prof()["demosaic_rcd"][0] / reps
    ctx.prof_enable(False)
    for b in (raw, rgb, clip):
        b.free()
    times = []
    for k in range(4):
        t0 = time.perf_counter()
        cfa = O.cfa_normalize_u16(raw_h, D.RGGB, 512.0, [16383.0] * 3)
        O.rcd_bayer(cfa)
        O.clip_plane(cfa)
        if k:
            times.append(time.perf_counter() - t0)
    cpu_s = statistics.median(times)
    return {"workload": f"RCD demosaic + clip plane, {size}x{size} RGGB 14-bit", "gpu_ms": round(gpu_ms, 4),
            "gpu_mp_s": round(n / 1e6 / (gpu_ms * 1e-3), 1),
            "gpu_hbm_gbs": round(DEMOSAIC_BYTES_PER_PX * n / (gpu_ms * 1e-3) / 1e9, 1),
            "cpu_ms": round(cpu_s * 1e3, 3), "cpu_mp_s": round(n / 1e6 / cpu_s, 2), "cpu_threads": O.threads()}


def _psnr(a, ref):
    import numpy as np
    d = a.astype(np.float64) - ref.astype(np.float64)
    mse = float(np.mean(d * d))
    return 10.0 * math.log10(1.0 / max(mse, 1e-30))


def learned_leg(ctx, args, raw, rgb, den, out, w, h, plain_ms):
    """Learned pipeline: RCD -> CNN luminance with the per-region NLM spot
    check and fallback -> guided chroma filters -> RL 1.0 x 20, timed like the
    plain one, plus its PSNR against the noise-free frame next to the plain
    path's (both after denoise)."""
    import ctypes
    import numpy as np
    import darkslide_amd as D
    L = D.lib()
    if not os.path.exists(args.checkpoint):
        return {"error": f"no trained checkpoint at {os.path.relpath(args.checkpoint, ROOT)}"}
    net = D.CNN(ctx, seed=1)
    net.load(args.checkpoint)
    pat = D._pat(D.RGGB)
    white = D._f3((16383.0,) * 3)
    ones = D._f3((1.0, 1.0, 1.0))
    res = {"checkpoint": os.path.relpath(args.checkpoint, ROOT), "train_steps": net.step, "threshold_db": args.threshold_db,
           "sizes": {}}
    for spec in args.learned_sizes.split(","):
        fw, fh = (int(v) for v in spec.split("x"))
        n = fw * fh
        if (fw, fh) == (w, h):
            b_raw, b_rgb, b_den, b_out = raw, rgb, den, out
            own = []
        else:
            b_raw, b_rgb, b_den, b_out = ctx.alloc(n * 2), ctx.alloc(n * 12), ctx.alloc(n * 12), ctx.alloc(n * 12)
            own = [b_raw, b_rgb, b_den, b_out]
            ctx._check(L.ds_synth_bayer_dev(ctx.handle, 0x5EED0002, fw, fh, 14, 1.0, pat, b_raw.ptr))
        st = D.LearnedStats()

        def lstep(stats=None):
            ctx._check(L.ds_demosaic_u16_dev(ctx.handle, b_raw.ptr, fw, fh, fw, pat, 512.0, white, ones, D.ALGO_RCD, b_rgb.ptr, None))
            ctx._check(L.ds_denoise_learned_dev(ctx.handle, net.handle, b_rgb.ptr, fw, fh, 0.0, 10, 3, float(args.threshold_db), b_den.ptr,
                                                stats))
            ctx._check(L.ds_raw_details_dev(ctx.handle, b_den.ptr, fw, fh, 1.0, 20, b_out.ptr))

        lstep(ctypes.byref(st))
        ctx.sync()
        for _ in range(max(args.warmup - 1, 0)):
            lstep()
        ctx.prof_enable(True)
        ctx.prof_reset()
        ctx.sync()
        steps = args.steps
        t0 = time.perf_counter()
        for _ in range(steps):
            lstep()
        ctx.sync()
        t1 = time.perf_counter()
        prof = ctx.prof()
        ctx.prof_enable(False)
        ms = (t1 - t0) * 1e3 / steps
        entry = {
            "value": round(n / 1e6 / (ms * 1e-3), 2), "unit": "MP/s", "ms_per_step": round(ms, 3), "steps": steps,
            "fallback_regions": int(st.fallback), "regions": int(st.regions),
            "mean_spot_deviation_db": round(float(st.mean_gap_db), 4), "max_spot_deviation_db": round(float(st.max_gap_db), 4),
            "stages_ms": {k: round(v[0] / max(v[1], 1), 4) for k, v in prof.items()},
        }
        log(f"learned {fw}x{fh}: {ms:.3f} ms/step, {entry['value']} MP/s, fallback {st.fallback}/{st.regions}")
        if (fw, fh) == (w, h):
            entry["vs_plain_pipeline"] = round(plain_ms / ms, 3)
            # quality against the noise-free frame, after denoise
            clean_raw = ctx.alloc(n * 2)
            ctx._check(L.ds_synth_bayer_dev(ctx.handle, 0x5EED0001, fw, fh, 14, 0.0, pat, clean_raw.ptr))
            ctx._check(L.ds_demosaic_u16_dev(ctx.handle, clean_raw.ptr, fw, fh, fw, pat, 512.0, white, ones, D.ALGO_RCD, b_out.ptr, None))
            clean = b_out.download((fh, fw, 3), np.float32)
            clean_raw.free()
            ctx._check(L.ds_demosaic_u16_dev(ctx.handle, b_raw.ptr, fw, fh, fw, pat, 512.0, white, ones, D.ALGO_RCD, b_rgb.ptr, None))
            noisy = b_rgb.download((fh, fw, 3), np.float32)
            ctx._check(L.ds_denoise_ex_dev(ctx.handle, b_rgb.ptr, fw, fh, 0.0, 10, 3, b_den.ptr))
            plain = b_den.download((fh, fw, 3), np.float32)
            lstep()
            learned_rgb = b_den.download((fh, fw, 3), np.float32)
            p_noisy, p_plain, p_learned = _psnr(noisy, clean), _psnr(plain, clean), _psnr(learned_rgb, clean)
            entry.update({"psnr_noisy_db": round(p_noisy, 4), "psnr_plain_db": round(p_plain, 4), "psnr_learned_db": round(p_learned, 4),
                          "psnr_gap_db": round(p_plain - p_learned, 4)})
            log(f"PSNR vs noise-free: noisy {p_noisy:.3f}  plain {p_plain:.3f}  learned {p_learned:.3f} dB")
            del clean, noisy, plain, learned_rgb
        for b in own:
            b.free()
        res["sizes"][f"{fw}x{fh}"] = entry
    net.close()
    return res


BATCH_SIZES = [(4256, 2848), (6000, 4000), (7008, 4672), (8256, 5504), (9504, 6336)]  # 12, 24, 33, 45, 61 MP


def batched_leg(ctx, args):
    """Batched export: frames cycling 12/24/33/45/61 MP, alternating 12/14-bit,
    noise level seeded in [1, 16], all resident in HBM, through one captured
    graph (tile shapes and branch count autotuned on the first call); plain
    and learned denoise.  Timed over whole graph launches."""
    import numpy as np
    import darkslide_amd as D
    L = D.lib()
    rng = np.random.default_rng(0x5EED)
    frames, bufs = [], []
    total = 0
    for i in range(args.batch_frames):
        fw, fh = BATCH_SIZES[i % len(BATCH_SIZES)]
        bits = 14 if i % 2 == 0 else 12
        iso = float(np.exp(rng.uniform(0.0, np.log(16.0))))
        raw, out = ctx.alloc(fw * fh * 2), ctx.alloc(fw * fh * 12)
        ctx._check(L.ds_synth_bayer_dev(ctx.handle, 0x5EED0000 + i, fw, fh, bits, iso, D._pat(D.RGGB), raw.ptr))
        black, white = (512.0, 16383.0) if bits == 14 else (128.0, 4095.0)
        frames.append(D.Frame(raw.ptr.value, fw, fh, fw, (0, 1, 1, 2), black, (white,) * 3, (1.0, 1.0, 1.0), out.ptr.value))
        bufs += [raw, out]
        total += fw * fh
    log(f"batched: {len(frames)} frames, {total / 1e6:.0f} MP")
    res = {"frames": len(frames), "megapixels": round(total / 1e6, 1),
           "sizes": "12/24/33/45/61 MP cycling, 12/14-bit alternating, noise 1-16x", "legs": {}}
    net = None
    if os.path.exists(args.checkpoint):
        net = D.CNN(ctx, seed=1)
        net.load(args.checkpoint)
    legs = [("plain", D.DENOISE_NLM, None)] + ([("learned", D.DENOISE_LEARNED, net)] if net else [])
    for name, mode, nn in legs:
        kw = dict(denoise=mode, amount=0.0, search=10, patch=3, raw_details=True, rl_sigma=1.0, rl_iters=20, net=nn,
                  threshold_db=args.threshold_db)
        t0 = time.perf_counter()
        info = ctx.pipeline_batch(frames, streams=0, autotune=True, **kw)  # capture (+ tuning) and one run
        setup_s = time.perf_counter() - t0
        ctx.sync()
        times = []
        for _ in range(args.batch_reps):
            t0 = time.perf_counter()
            ctx.pipeline_batch(frames, streams=0, autotune=True, info=False, **kw)
            ctx.sync()
            times.append(time.perf_counter() - t0)
        t = statistics.median(times)
        last = ctx.pipeline_batch(frames, streams=0, autotune=True, **kw)
        entry = {"value": round(total / 1e6 / t, 2), "unit": "MP/s", "s_per_batch": round(t, 4), "reps": args.batch_reps,
                 "branches": info["branches"], "graph_reused_on_rerun": bool(last["graph_reused"]),
                 "first_call_s_incl_capture_and_tuning": round(setup_s, 3)}
        if mode == D.DENOISE_LEARNED:
            entry["fallback_regions"] = last["fallback_regions"]
            entry["regions"] = last["regions"]
        res["legs"][name] = entry
        log(f"batched {name}: {entry['value']} MP/s ({t:.3f} s per batch, {info['branches']} branches)")
    # the first frames again from pinned host memory: uploads and downloads
    # run on copy streams while other frames' kernels run; the same frames
    # device-resident beside it
    nh = min(args.batch_host_frames, len(frames))
    if nh:
        sub, hbufs, mp = [], [], 0
        for f in frames[:nh]:
            hin, hout = ctx.host_alloc(f.w * f.h * 2), ctx.host_alloc(f.w * f.h * 12)
            ctx._check(L.ds_memcpy_d2h(ctx.handle, hin.ptr, ctypes.c_void_p(f.raw), f.w * f.h * 2))
            g = D.Frame.from_buffer_copy(f)
            g.raw, g.rgb_out = hin.ptr.value, hout.ptr.value
            sub.append(g)
            hbufs += [hin, hout]
            mp += f.w * f.h
        ctx.sync()
        kw = dict(denoise=D.DENOISE_NLM, amount=0.0, search=10, patch=3, raw_details=True, rl_sigma=1.0, rl_iters=20)
        out = {"frames": nh, "megapixels": round(mp / 1e6, 1)}
        for name, fr, host in (("device_resident", frames[:nh], False), ("pinned_host", sub, True)):
            info = ctx.pipeline_batch(fr, streams=0, autotune=True, host_frames=host, **kw)
            times = []
            for _ in range(args.batch_reps):
                t0 = time.perf_counter()
                ctx.pipeline_batch(fr, streams=0, autotune=True, host_frames=host, info=False, **kw)
                ctx.sync()
                times.append(time.perf_counter() - t0)
            t = statistics.median(times)
            out[name] = {"value": round(mp / 1e6 / t, 2), "unit": "MP/s", "s_per_batch": round(t, 4), "branches": info["branches"]}
            log(f"batched subset {name}: {out[name]['value']} MP/s")
        res["legs"]["plain_host_frames"] = out
        for b in hbufs:
            b.free()
    tuning = {}
    for fw, fh in BATCH_SIZES:
        tuning[f"{fw}x{fh}"] = ctx.tuning(fw, fh)
    res["tile_variants_rcd_mhc_rl"] = tuning
    if net:
        net.close()
    for b in bufs:
        b.free()
    return res


def highlights_leg(ctx, args, raw, rgb, clip, w, h):
    """Engine::base's highlight reconstruction (pipeline/mod.rs:186-187) on
    the bench frame with saturated discs: RCD demosaic with the clip plane,
    cfa_clip_levels, reconstruct_highlights; and the same on the unsaturated
    frame, where the device flag skips the work."""
    import numpy as np
    import darkslide_amd as D
    L = D.lib()
    n = w * h
    host = saturate(raw.download((h, w), np.uint16), 7)
    sat = ctx.alloc(n * 2)
    sat.upload(host)
    lv = ctx.alloc(64)
    pat = D._pat(D.RGGB)
    white = D._f3((16383.0,) * 3)
    mul = D._f3(HL_MUL)

    def run(src):
        ctx._check(L.ds_demosaic_u16_dev(ctx.handle, src.ptr, w, h, w, pat, 512.0, white, mul, D.ALGO_RCD, rgb.ptr, clip.ptr))
        ctx._check(L.ds_cfa_clip_levels_u16_dev(ctx.handle, src.ptr, w, h, w, pat, 512.0, white, lv.ptr))
        ctx._check(L.ds_reconstruct_highlights_dev(ctx.handle, rgb.ptr, clip.ptr, w, h, mul, lv.ptr))

    res = {}
    for name, src in (("saturated", sat), ("unsaturated", raw)):
        run(src)
        ctx.sync()
        ctx.prof_enable(True)
        ctx.prof_reset()
        for _ in range(5):
            run(src)
        ctx.sync()
        pr = ctx.prof()
        ctx.prof_enable(False)
        st = {k: round(v[0] / max(v[1], 1), 4) for k, v in pr.items() if k in ("cfa_clip_levels", "reconstruct_highlights")}
        res[name] = {"stage_ms": st}
        if name == "saturated":
            cp = clip.download((h, w), np.float32)
            res[name]["clipped_fraction"] = round(float((cp > 0).mean()), 5)
            res[name]["levels"] = [float(v) for v in lv.download((3,), np.float32)]
            ms = st["reconstruct_highlights"]
            res[name]["hbm_gbs"] = round(HIGHLIGHTS_BYTES_PER_PX * n / (ms * 1e-3) / 1e9, 1)
            res[name]["hbm_frac"] = round(res[name]["hbm_gbs"] / PEAK_HBM_GBS, 4)
    res["algorithmic"] = f"{HIGHLIGHTS_BYTES_PER_PX} B/px when any sample is clipped; as-shot mul {list(HL_MUL)}"
    sat.free()
    lv.free()
    log(f"highlights: {res}")
    return res


def transfer_leg(ctx, step, raw, out, w, h, reps):
    """The same step from host memory to host memory: the u16 frame copied in
    from pinned host memory, the device pipeline, the RGB result copied back,
    one synchronisation per frame; and each copy on its own."""
    import darkslide_amd as D
    L = D.lib()
    n = w * h
    hin, hout = ctypes.c_void_p(), ctypes.c_void_p()
    ctx._check(L.ds_host_alloc(ctx.handle, n * 2, ctypes.byref(hin)))
    ctx._check(L.ds_host_alloc(ctx.handle, n * 12, ctypes.byref(hout)))
    p_in, p_out, pinned = hin.value, hout.value, True
    src, dst = ctypes.c_void_p(p_in), ctypes.c_void_p(p_out)
    ctx._check(L.ds_memcpy_d2h(ctx.handle, src, raw.ptr, n * 2))
    ctx.sync()

    def timed(fn):
        fn()
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
            ctx.sync()
        return (time.perf_counter() - t0) * 1e3 / reps

    def whole():
        ctx._check(L.ds_memcpy_h2d(ctx.handle, raw.ptr, src, n * 2))
        step()
        ctx._check(L.ds_memcpy_d2h(ctx.handle, dst, out.ptr, n * 12))

    ms = timed(whole)
    h2d = timed(lambda: ctx._check(L.ds_memcpy_h2d(ctx.handle, raw.ptr, src, n * 2)))
    d2h = timed(lambda: ctx._check(L.ds_memcpy_d2h(ctx.handle, dst, out.ptr, n * 12)))
    res = {"ms_per_frame": round(ms, 3), "mp_s": round(n / 1e6 / (ms * 1e-3), 1), "pinned_host_memory": pinned,
           "h2d_bytes": n * 2, "h2d_ms": round(h2d, 3), "h2d_gbs": round(n * 2 / (h2d * 1e-3) / 1e9, 1),
           "d2h_bytes": n * 12, "d2h_ms": round(d2h, 3), "d2h_gbs": round(n * 12 / (d2h * 1e-3) / 1e9, 1),
           "note": "u16 frame in, RGB f32 out, synchronised per frame; not the headline value"}
    ctx._check(L.ds_host_free(ctx.handle, hin))
    ctx._check(L.ds_host_free(ctx.handle, hout))
    log(f"host to host: {res}")
    return res


def detail_leg(ctx, rgb, w, h):
    """The render detail stages (pipeline/detail.rs) and the export post-steps
    (export.rs:244-247), in place on the bench's resident 24 MP image, timed
    by the library's per-call events: five calls each after one warm-up."""
    import darkslide_amd as D
    L = D.lib()
    cases = [
        ("noise_reduction", "defaults (colour 25)", lambda: L.ds_noise_reduction_dev(
            ctx.handle, rgb.ptr, w, h, ctypes.byref(D.NrParams(0, 50, 0, 25, 50, 50)), None, 1.0)),
        ("noise_reduction", "luminance 50, colour 25", lambda: L.ds_noise_reduction_dev(
            ctx.handle, rgb.ptr, w, h, ctypes.byref(D.NrParams(50, 50, 0, 25, 50, 50)), None, 1.0)),
        ("sharpen", "defaults (40, 1.0, 25, 0)", lambda: L.ds_sharpen_dev(
            ctx.handle, rgb.ptr, w, h, ctypes.byref(D.SharpenParams(40, 1.0, 25, 0)), None, 1.0)),
        ("sharpen", "masking 40", lambda: L.ds_sharpen_dev(
            ctx.handle, rgb.ptr, w, h, ctypes.byref(D.SharpenParams(40, 1.0, 25, 40)), None, 1.0)),
        ("super_res_refine", "", lambda: L.ds_super_res_refine_dev(ctx.handle, rgb.ptr, w, h)),
        ("output_sharpen", "Screen, Standard", lambda: L.ds_output_sharpen_dev(
            ctx.handle, rgb.ptr, w, h, D.SHARPEN_FOR_SCREEN, D.SHARPEN_AMOUNT_STANDARD)),
    ]
    n = w * h
    res = []
    for stage, label, call in cases:
        ctx._check(call())
        ctx.sync()
        ctx.prof_enable(True)
        ctx.prof_reset()
        for _ in range(5):
            ctx._check(call())
        ctx.sync()
        pr = ctx.prof()
        ctx.prof_enable(False)
        t, k = pr[stage]
        ms = t / max(k, 1)
        row = {"stage": stage, "settings": label, "ms": round(ms, 4), "mp_s": round(n / 1e6 / (ms * 1e-3), 1)}
        if label.startswith("defaults") or stage in ("super_res_refine", "output_sharpen"):
            bpp = DETAIL_BYTES_PER_PX[stage]
            row["algorithmic_bytes_per_px"] = bpp
            row["hbm_gbs"] = round(bpp * n / (ms * 1e-3) / 1e9, 1)
            row["hbm_frac"] = round(row["hbm_gbs"] / PEAK_HBM_GBS, 4)
        res.append(row)
    log(f"detail: {res}")
    return res


def stage_errors(ctx, args):
    """Max-abs and RMS error of every stage against the FP32 restatement of
    the reference and against its FP64 twin, on a size x size crop of the
    bench's synthetic scene: each stage's GPU input is also the oracle's
    input.  The learned CNN's forward is compared with the oracle's
    bf16-emulating FP64 forward."""
    import numpy as np
    import darkslide_amd as D
    import oracle as O
    O.build()
    n = args.error_size
    raw = D.synth_bayer(0x5EED0003, n, n, 14, 4.0)
    out = {}

    def err(name, gpu, ref32, ref64):
        g = np.asarray(gpu, np.float64)
        d32 = g - np.asarray(ref32, np.float64)
        d64 = g - np.asarray(ref64, np.float64)
        out[name] = {"max_abs_vs_fp32": float(np.abs(d32).max()), "rms_vs_fp32": float(np.sqrt(np.mean(d32 * d32))),
                     "max_abs_vs_fp64": float(np.abs(d64).max()), "rms_vs_fp64": float(np.sqrt(np.mean(d64 * d64)))}

    cfa32 = O.cfa_normalize_u16(raw, D.RGGB, 512.0, [16383.0] * 3)
    cfa64 = O.cfa_normalize_u16(raw, D.RGGB, 512.0, [16383.0] * 3, dtype=np.float64)
    rgb, clip = ctx.demosaic_u16(raw)
    r32, c32 = O.demosaic(cfa32)
    r64, c64 = O.demosaic(cfa64)
    err("demosaic_rcd", rgb, r32, r64)
    err("clip_plane", clip, c32, c64)
    mhc, _ = ctx.demosaic_u16(raw, algo=D.ALGO_MHC)
    err("demosaic_mhc", mhc, O.mhc_bayer(cfa32), O.mhc_bayer(cfa64))
    lab = ctx.oklab_from_prophoto(rgb)
    err("oklab", lab, O.oklab_from_prophoto(rgb), O.oklab_from_prophoto(rgb.astype(np.float64)))
    Lp = np.ascontiguousarray(lab[..., 0])
    sg = ctx.estimate_noise(Lp)
    err("estimate_noise", [sg], [O.estimate_noise(Lp)], [O.estimate_noise(Lp.astype(np.float64))])
    hp = np.float32(max(np.float32(sg), np.float32(1e-4))) * np.float32(0.6)
    err("nlm", ctx.nlm(Lp, hp, 10, 3), O.nlm(Lp, hp, 10, 3), O.nlm(Lp.astype(np.float64), float(hp), 10, 3))
    dn = ctx.denoise(rgb, 0.0, 10, 3)
    err("denoise", dn, O.denoise(rgb, 0.0, 10, 3), O.denoise(rgb.astype(np.float64), 0.0, 10, 3))
    rl = ctx.raw_details(dn, 1.0, 20)
    err("raw_details", rl, O.raw_details(dn, 1.0, 20), O.raw_details(dn.astype(np.float64), 1.0, 20))
    sat = saturate(raw.copy(), 8, discs=6)
    srgb, sclip = ctx.demosaic_u16(sat, mul=HL_MUL)
    lv = ctx.cfa_clip_levels_u16(sat)
    cfa_s = O.cfa_normalize_u16(sat, D.RGGB, 512.0, [16383.0] * 3)
    err("cfa_clip_levels", lv, O.cfa_clip_levels(cfa_s), O.cfa_clip_levels(cfa_s.astype(np.float64)))
    hl = ctx.reconstruct_highlights(srgb, sclip, HL_MUL, lv)
    err("reconstruct_highlights", hl, O.reconstruct_highlights(srgb, sclip, HL_MUL, lv),
        O.reconstruct_highlights(srgb.astype(np.float64), sclip.astype(np.float64), HL_MUL, lv.astype(np.float64)))
    out["reconstruct_highlights"]["note"] = ("saturated discs; near-clip is box_blur(clip, 3) > 0, whose FP32 running sums "
                                             "leave residues that FP64 does not, so FP64 classifies some pixels differently")
    r64d = rl.astype(np.float64)
    for nr in [(0, 50, 0, 25, 50, 50), (50, 50, 0, 25, 50, 50)]:
        err(f"noise_reduction{nr}", ctx.noise_reduction(rl, nr), O.noise_reduction(rl, nr), O.noise_reduction(r64d, nr))
    err("sharpen(40,1,25,0)", ctx.sharpen(rl, (40, 1.0, 25, 0)), O.sharpen(rl, (40, 1.0, 25, 0)), O.sharpen(r64d, (40, 1.0, 25, 0)))
    err("super_res_refine", ctx.super_res_refine(rl), O.super_res_refine(rl), O.super_res_refine(r64d))
    err("output_sharpen(Screen,Standard)", ctx.output_sharpen(rl, 1, 1), O.output_sharpen(rl, 1, 1), O.output_sharpen(r64d, 1, 1))
    if os.path.exists(args.checkpoint):
        net = D.CNN(ctx, seed=1)
        net.load(args.checkpoint)
        s = max(np.float32(sg), np.float32(1e-4))
        x = O.cnn_input(Lp, float(s)).astype(np.float32)[None, :128, :128].copy()
        y = net.infer(x)
        p = net.params().astype(np.float64)
        ref = O.cnn_forward(p, x, bf16=True)
        err("cnn_forward_bf16_vs_fp64_emulation", y, ref, ref)
        out["cnn_forward_bf16_vs_fp64_emulation"]["note"] = "crop 128x128; both columns against the FP64 forward with bf16 roundings"
        net.close()
    out["_input"] = f"{n}x{n} RGGB 14-bit synthetic frame, noise 4x; NLM search 10 patch 3 h=0.6 sigma; RL 1.0 x 20"
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    # five untimed frames (~40 ms of work) bring the GPU to its full clock
    # before the timed ones
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--width", type=int, default=6000)
    ap.add_argument("--height", type=int, default=4000)
    ap.add_argument("--cpu-size", type=int, default=1024)
    ap.add_argument("--cpu-reps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cpu-stages", action="store_true", help="skip the full-frame per-stage CPU timings")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="also run one plain device copy of this size (known-size reference for HBM counters)")
    ap.add_argument("--checkpoint", default=os.path.join(ROOT, "darkslide_amd", "weights", "cnn_l32x8.dsck"))
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--threshold-db", type=float, default=0.3)
    ap.add_argument("--learned-sizes", default="6000x4000,8256x5504")
    ap.add_argument("--batch-frames", type=int, default=48, help="frames of the batched-export leg (0: skip)")
    ap.add_argument("--batch-reps", type=int, default=3)
    ap.add_argument("--batch-host-frames", type=int, default=10, help="frames of the batch also run from pinned host memory")
    ap.add_argument("--no-errors", action="store_true", help="skip the per-stage error report against the oracle")
    ap.add_argument("--error-size", type=int, default=384)
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on exactly one GPU (asked for --gpus {args.gpus}, WORLD_SIZE {world}); "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)

    import ctypes
    import darkslide_amd as D

    L = D.lib()
    ctx = D.Context(0)
    w, h = args.width, args.height
    n = w * h
    raw = ctx.alloc(n * 2)
    rgb = ctx.alloc(n * 12)
    clip = ctx.alloc(n * 4)
    den = ctx.alloc(n * 12)
    out = ctx.alloc(n * 12)
    pat = D._pat(D.RGGB)
    white = D._f3((16383.0,) * 3)
    ones = D._f3((1.0, 1.0, 1.0))
    ctx._check(L.ds_synth_bayer_dev(ctx.handle, 0x5EED0001, w, h, 14, 1.0, pat, raw.ptr))
    ctx.sync()

    def step():
        ctx._check(L.ds_demosaic_u16_dev(ctx.handle, raw.ptr, w, h, w, pat, 512.0, white, ones, D.ALGO_RCD, rgb.ptr, clip.ptr))
        ctx._check(L.ds_denoise_ex_dev(ctx.handle, rgb.ptr, w, h, 0.0, 10, 3, den.ptr))
        ctx._check(L.ds_raw_details_dev(ctx.handle, den.ptr, w, h, 1.0, 20, out.ptr))

    if args.copy_calibration_mib:
        nb = args.copy_calibration_mib << 20
        a, b = ctx.alloc(nb), ctx.alloc(nb)
        ctx._check(L.ds_dev_copy(ctx.handle, b.ptr, a.ptr, nb))
        ctx.sync()
        a.free()
        b.free()
        log(f"copy calibration: {args.copy_calibration_mib} MiB read + written")
    for i in range(args.warmup):
        step()
        ctx.sync()
        log(f"warmup step {i + 1}/{args.warmup} done")
    ctx.prof_enable(True)
    ctx.prof_reset()
    ctx.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.sync()
    t1 = time.perf_counter()
    prof = ctx.prof()
    ctx.prof_enable(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = n / 1e6 * args.steps / (t1 - t0)
    log(f"timed {args.steps} steps: {ms_per_step:.3f} ms/step, {value:.1f} MP/s")

    stages = {k: {"avg_ms": round(v[0] / max(v[1], 1), 4), "launches": v[1]} for k, v in prof.items()}
    # MHC-only demosaic, timed the same way (not part of the step)
    ctx.prof_enable(True)
    ctx.prof_reset()
    for _ in range(5):
        ctx._check(L.ds_demosaic_u16_dev(ctx.handle, raw.ptr, w, h, w, pat, 512.0, white, ones, D.ALGO_MHC, rgb.ptr, clip.ptr))
    mhc = ctx.prof().get("demosaic_mhc", (0.0, 1))
    ctx.prof_enable(False)
    stages["demosaic_mhc"] = {"avg_ms": round(mhc[0] / max(mhc[1], 1), 4), "launches": mhc[1]}

    nlm_ms = stages["nlm"]["avg_ms"]
    nlm_tflops = NLM_FLOP_PER_PX * n / (nlm_ms * 1e-3) / 1e12
    rcd_ms = stages["demosaic_rcd"]["avg_ms"]
    stages["demosaic_rcd"]["hbm_gbs"] = round(DEMOSAIC_BYTES_PER_PX * n / (rcd_ms * 1e-3) / 1e9, 1)
    stages["demosaic_rcd"]["hbm_frac"] = round(stages["demosaic_rcd"]["hbm_gbs"] / PEAK_HBM_GBS, 4)
    mhc_ms = stages["demosaic_mhc"]["avg_ms"]
    if mhc_ms > 0:
        stages["demosaic_mhc"]["hbm_gbs"] = round(MHC_BYTES_PER_PX * n / (mhc_ms * 1e-3) / 1e9, 1)
    dominant = max(prof.items(), key=lambda kv: kv[1][0])[0]
    roofline = {
        "kernel": "k_nlm_tile (nlm)",
        "bound": "valu",
        "bound_note": "compute-bound: FP32 arithmetic on the vector ALUs (no MFMA instructions), priced against "
                      "MI355X's published FP32 peak of 157.3 TFLOP/s",
        "achieved": round(nlm_tflops, 3),
        "peak": PEAK_FP32_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(nlm_tflops / PEAK_FP32_TFLOPS, 4),
        "traffic": round(NLM_TRAFFIC_BYTES_PER_PX * n),
        "traffic_note": "HBM bytes per launch from rocprofv3 FETCH_SIZE + WRITE_SIZE (calibrated on a 1 GiB copy); "
                        "algorithmic minimum 28 B/px (L, a, b in; the four guide planes out)",
        "algorithmic": f"{NLM_FLOP_PER_PX} flop/px x {n} px per launch",
        "dominant_stage": dominant,
    }
    learned = None
    if not args.no_learned:
        learned = learned_leg(ctx, args, raw, rgb, den, out, w, h, ms_per_step)
    host_to_host = transfer_leg(ctx, step, raw, out, w, h, 5)
    highlights = highlights_leg(ctx, args, raw, rgb, clip, w, h)
    detail = detail_leg(ctx, rgb, w, h)
    batched = None
    if args.batch_frames > 0:
        batched = batched_leg(ctx, args)
    errors = None
    if not args.no_errors:
        log("per-stage errors against the oracle ...")
        errors = stage_errors(ctx, args)
    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(args.cpu_size, args.cpu_reps)
        cpu["cpu_model"] = cpu_model()
        cpu["logical_cpus"] = os.cpu_count()
        log(f"cpu baseline {cpu['value']} MP/s on {cpu['cores']} threads ({cpu['cpu_model']})")
        if not args.no_cpu_stages:
            log(f"cpu stages at {w}x{h} ...")
            cs = cpu_stages(w, h)
            gpu_of = {"demosaic_rcd": stages["demosaic_rcd"]["avg_ms"], "demosaic_mhc": stages["demosaic_mhc"]["avg_ms"],
                      "denoise": sum(stages[k]["avg_ms"] for k in ("oklab_fwd", "estimate_noise", "nlm", "guided_filter") if k in stages),
                      "raw_details": stages["raw_details"]["avg_ms"]}
            for k, v in cs.items():
                v["gpu_ms"] = gpu_of[k]
                v["gpu_speedup"] = round(v["s"] * 1e3 / gpu_of[k], 1) if gpu_of[k] > 0 else None
            cpu["stages_full_frame"] = {"frame": f"{w}x{h}", "threads": cpu["cores"], "stages": cs,
                                        "note": "each stage once on the CPU (s); gpu_ms: the same stage's kernels, averaged over the timed steps"}
            log(f"cpu stages: {cs}")
        cpu["config1_rcd"] = rcd_config1(ctx)
        log(f"config 1 RCD: {cpu['config1_rcd']}")
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "MP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded RGGB uint16 14-bit scene (gradients, polygons, sinusoids), Poisson-Gaussian noise gain 0.5, read 4 DN",
        "config": {
            "workload": "plain pipeline, 6000x4000 RGGB: RCD demosaic -> denoise (NLM patch 7x7, search 21x21, h=0.6 sigma) -> RL sigma 1.0 x20",
            "width": w,
            "height": h,
            "megapixels": round(n / 1e6, 3),
            "denoiser": "nlm",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "stages_ms": stages,
        "learned": learned,
        "batched": batched,
        "host_to_host": host_to_host,
        "highlights": highlights,
        "detail": detail,
        "stage_errors": errors,
    }
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
