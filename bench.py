#!/usr/bin/env python3
"""Throughput of two-step BM3D on one MI355X: 1920x1080 RGB, sigma = 25.

Workload: the ISP_MW Frame path (OPP colour, step 1 hard threshold matched on
Y, step 2 Wiener matched on the basic Y, OPP inverse with clip) on a seeded
synthetic 1920x1080 RGB image (isp_mw_amd.synth, rectangles + gradient +
gratings, channels mixed, AWGN sigma 25/255).  Parameters: BlockSize 8,
BlockStep 3 (both steps), BMrange 12, BMstep 1, GroupSize 16 / 32, lambda 2.7,
thMSE from thMSE_Default (profile "np" otherwise).

One step = two-step BM3D of a batch of --frames-per-step frames (default 24)
with inputs and outputs resident in HBM, processed as a frame sequence is:
--streams of them in flight (default 2: one context and one default-priority
stream each), without instrumentation.  The same frames one after another,
and once more with per-stage HIP events (stage times and the roofline),
follow.  Prints one JSON line (rank 0, one GPU).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

W, H, SIGMA, SEED = 1920, 1080, 25.0, 1
ROOT = os.path.dirname(os.path.abspath(__file__))
B = 8
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
PEAK_F16_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0


def make_para():
    from isp_mw_amd import BM3D_Para
    p = BM3D_Para("np")
    for st in (p.basic, p.final):
        st.BlockStep = 3
        st.BMrange = 12
        st.BMstep = 1
    p.basic.GroupSize = 16
    p.final.GroupSize = 32
    p.basic.lambda_ = 2.7
    p.set_sigma(SIGMA)
    return p


def grid(n: int, s: int) -> np.ndarray:
    last = n - B
    return np.minimum(np.arange(last // s + 1 + (1 if last % s else 0)) * s, last)


def window_len(pos: np.ndarray, n: int, R: int) -> np.ndarray:
    # clipped search window with BMstep 1 (include/Block.h:716-751)
    return np.minimum(pos + R, n - B) - np.maximum(pos - R, 0) + 1


def transform_flops(n) -> np.ndarray:
    # dense separable N x 8 x 8 DCT: 2 B^2 N (2B + N)
    n = np.asarray(n, np.float64)
    return 2.0 * B * B * n * (2 * B + n)


def cpu_baseline(para, clean_rgb_noisy, threads: int, target_s: float, runs: int = 3):
    """Time the oracle (FP64 C restatement, row bands over `threads` threads)
    on a crop of the frame sized to about target_s per run; median of `runs`."""
    import oracle
    fast = oracle.host_supports_fast()
    oracle.set_threads(threads, fast=fast)
    noisy = clean_rgb_noisy
    # calibrate on a small crop, then size the timed crop to ~target_s per run
    hc = 32 if threads == 1 else 64
    t0 = time.perf_counter()
    oracle.bm3d_rgb(para.basic, para.final, noisy[:, :hc, :], fast=fast)
    rate = W * hc / max(time.perf_counter() - t0, 1e-6)
    hc = int(min(H, max(32, rate * target_s / W)))
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        oracle.bm3d_rgb(para.basic, para.final, noisy[:, :hc, :], fast=fast)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    oracle.set_threads(1, fast=fast)
    return {
        "value": round(W * hc / dt / 1e6, 4),
        "unit": "Mpix/s",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle (FP64 C restatement, {'x86-64-v3' if fast else 'baseline x86-64'} build, "
                  f"{threads} thread{'s, row bands' if threads > 1 else ''}) two-step BM3D on the top "
                  f"{W}x{hc} crop of the same frame, median of {runs} runs "
                  f"({', '.join(f'{t:.2f}' for t in times)} s)",
    }


TRAFFIC_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                            "bm3d_1080p_rgb_{prec}_hbm_traffic.txt")


def measured_traffic(prec: str, kernel: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary
    (FETCH_SIZE and WRITE_SIZE runs, each scaled by a copy kernel's reading).
    Lines of the form `traffic <stage> <fetch_bytes> <write_bytes>`."""
    path = TRAFFIC_FILE.format(prec=prec)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            f = line.split()
            if len(f) == 4 and f[0] == "traffic" and f[1] == kernel:
                return {"bytes": float(f[2]) + float(f[3]), "fetch_bytes": float(f[2]),
                        "write_bytes": float(f[3]), "source": os.path.relpath(path, os.path.dirname(TRAFFIC_FILE))}
    return None


STATS_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                          "bm3d_1080p_rgb_{prec}_kernel_stats.txt")
# kernel-name fragments of each stage's kernel in the committed rocprofv3 stats
STAGE_KERNELS = {
    # the plain form's filter: the split-operand instantiations of the MFMA kernel
    "fp32": {"filter_final": "filter_mfma_kernelIDF16_Lb1ELb0ELi8ELb1E",
             "filter_basic": "filter_mfma_kernelIDF16_Lb0ELb0ELi8ELb1E",
             "match_basic": "match_kernel<double, double, true, true, false, false>",
             "match_final": "match_kernel<float, double, true, true, false, false>"},
    "mixed": {"filter_final": "filter_mfma_kernelIDF16_Lb1ELb0ELi8ELb0E",
              "filter_basic": "filter_mfma_kernelIDF16_Lb0ELb0ELi8ELb0E",
              "match_basic": "match_walk_kernel<double, true, false", "match_final": "match_walk_kernel<float, true, false"},
}


def profiled_ms(prec: str, stage: str):
    """Average duration (ms) of a stage's kernel in the committed rocprofv3
    kernel-trace summary of the same form, or None."""
    path = STATS_FILE.format(prec=prec)
    frag = STAGE_KERNELS["fp32" if prec == "fp32" else "mixed"].get(stage)
    if prec == "bf16" or frag is None or not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            if line.startswith('"') and frag in line:
                f = line.rsplit('",', 1)[1].split(",")
                return round(float(f[2]) / 1e6, 4)
    return None


UTIL_FILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                         "bm3d_1080p_rgb_{prec}_utilisation.txt")


def measured_util(prec: str, kernel: str):
    """Pipe utilisation of `kernel` from the committed rocprofv3 SQ/GRBM
    counter summary: lines `util <stage> <valu_busy> <mfma_busy>
    <mfma_insts_per_launch> <lds_bank_conflict_per_cu_cycle> <ms>`."""
    path = UTIL_FILE.format(prec=prec)
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        for line in fh:
            f = line.split()
            if len(f) == 7 and f[0] == "util" and f[1] == kernel:
                return {"valu_busy": float(f[2]), "mfma_busy": float(f[3]), "mfma_insts_per_launch": float(f[4]),
                        "lds_bank_conflict_cycles_per_cu_cycle": float(f[5]), "ms_under_counters": float(f[6]),
                        "source": os.path.relpath(path, os.path.dirname(UTIL_FILE))}
    return None


def pipelined_steps(fs, ys, x, steps: int, warmup: int):
    """`steps` frames with len(fs) of them in flight: frame i runs on context
    i % n and its own stream (default priority), without stage events.
    Returns seconds per frame."""
    import torch
    n = len(fs)
    streams = [torch.cuda.Stream() for _ in range(n)]

    def run(k_steps):
        for i in range(k_steps):
            with torch.cuda.stream(streams[i % n]):
                fs[i % n].process(ys[i % n], x)

    run(max(warmup, 1) * n)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(steps)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def timed_steps(f, y, x, steps: int, warmup: int):
    import torch
    for _ in range(warmup):
        f.process(y, x)
    torch.cuda.synchronize()
    f.set_timing(True)
    f.timing(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        f.process(y, x)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    stages = f.timing()
    f.set_timing(False)
    return dt, stages


def bench_batch(args) -> int:
    """64 grey images of uneven sizes (W in [256, 1024], H in [256, 768]) and
    per-image sigma in [5, 50], each generated like the 512x512 case, filtered
    as one batch: one launch per stage, replayed as one hipGraph, tile shapes
    and the final GroupSize (16 or 32) per size class from the autotuner."""
    import torch
    from isp_mw_amd import BM3D, BM3DBatch, synth

    rng = np.random.default_rng(2024)
    n_img = 64
    Ws = rng.integers(256, 1025, n_img)
    Hs = rng.integers(256, 769, n_img)
    sig = rng.uniform(5.0, 50.0, n_img)
    para = make_para()
    f = BM3D(para, device=0, precision=args.precision)
    pairs = [synth.grey_pair(int(h), int(w), float(s), 1000 + k) for k, (w, h, s) in enumerate(zip(Ws, Hs, sig))]
    srcs = [torch.from_numpy(n).cuda() for _, n in pairs]
    dsts = [torch.empty_like(x) for x in srcs]
    # autotune each size class on its median image size
    classes = {}
    for w, h in zip(Ws, Hs):
        classes.setdefault(f.size_class(int(w), int(h)), []).append((int(w), int(h)))
    tuning = {}
    t0 = time.perf_counter()
    for c, sizes in sorted(classes.items()):
        w, h = sorted(sizes)[len(sizes) // 2]
        tuning[c] = f.autotune(w, h, 25.0, auto_group=True)
    tune_s = time.perf_counter() - t0
    b = BM3DBatch(f, dsts, srcs, sigmas=list(sig), auto_group=True)
    for _ in range(args.warmup):
        b.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.run()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    px = float(sum(int(w) * int(h) for w, h in zip(Ws, Hs)))
    outs = [d.cpu().numpy() for d in dsts]
    gain = float(np.mean([synth.psnr(o, c) - synth.psnr(n, c) for o, (c, n) in zip(outs, pairs)]))
    # the same launches enqueued one by one, and one call per image
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(max(1, args.steps // 2)):
        b.run(graph=False)
    torch.cuda.synchronize()
    dt_eager = (time.perf_counter() - t0) / max(1, args.steps // 2)
    # one single-image batch per image on the same tuned context, run one after the other
    ones = [BM3DBatch(f, [d], [x], sigmas=[float(sg)], auto_group=True) for d, x, sg in zip(dsts, srcs, sig)]
    for o in ones[:4]:
        o.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for o in ones:
        o.run()
    torch.cuda.synchronize()
    dt_single = time.perf_counter() - t0
    for o in ones:
        o.close()
    info = b.info()
    res = {
        "metric": "Mpix/s batched BM3D, 64 grey images of uneven size, per-image sigma, one hipGraph",
        "value": round(px / dt / 1e6, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.precision == "fp32" else f"{args.precision}+fp32",
        "data": "synthetic (seeded rectangles+gradient+gratings per image, AWGN per-image sigma)",
        "config": {"workload": "64 grey float32 images, W in [256,1024], H in [256,768], sigma in [5,50]",
                   "block": 8, "block_step": [3, 3], "bm_range": 12, "group_size": [16, "16 or 32 (autotuned)"],
                   "precision": args.precision, "pixels": px},
        "batch": {**info, "eager_ms": round(dt_eager * 1e3, 3), "one_image_per_launch_ms": round(dt_single * 1e3, 3),
                  "speedup_vs_one_image_per_launch": round(dt_single / dt, 2)},
        "autotune": {str(c): {k: v for k, v in t.items()} for c, t in tuning.items()},
        "autotune_seconds": round(tune_s, 2),
        "quality": {"mean_psnr_gain_db": round(gain, 3)},
    }
    print(json.dumps(res), flush=True)
    return 0


def pilot_network(P, kind: str, steps: int):
    """(network, training record): the shipped weights of `kind` (trained by
    isp_mw_amd.train_pilot) when steps is 0, else a network trained here by
    the same recipe for `steps` steps."""
    import torch
    from isp_mw_amd import train_pilot
    if steps <= 0:
        path = P.weights_path(kind)
        return P.pretrained(kind), dict(P.PilotCNN.meta_of(path), source=os.path.relpath(path, ROOT))
    t0 = time.perf_counter()
    net, meta = train_pilot.train_kind(kind, steps, log_every=max(1, steps // 10))
    torch.cuda.synchronize()
    meta.update(source="trained in this run", ms_per_step=round(meta["train_s"] / steps * 1e3, 3),
                total_s=round(time.perf_counter() - t0, 1))
    return net, meta


def bench_pilot(args) -> int:
    """Learned pilot: train the residual CNN on 20k random 128x128 crops of
    synthetic images (sigma in [10, 40]) against GPU step-1 targets (Adam,
    batch 32), then denoise a 1920x1080 grey frame at sigma 25 with the pilot
    as the Wiener guide, against plain two-step BM3D."""
    import torch
    from isp_mw_amd import pilot as P, synth

    net, training = pilot_network(P, "grey", args.pilot_steps)
    clean, n = synth.grey_pair(H, W, SIGMA, SEED)
    x = torch.from_numpy(n).cuda()
    para = P.default_para(SIGMA)
    plain = P.PilotBM3D(para, precision=args.precision)
    piloted = P.PilotBM3D(para, net, precision=args.precision)

    def timeit(fn, k):
        for _ in range(2):
            fn()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(k):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / k

    dt_pilot = timeit(lambda: piloted(x), args.steps)
    dt_plain = timeit(lambda: plain(x), args.steps)
    dt_guide_pilot = timeit(lambda: piloted.guide(x), args.steps)
    dt_guide_step1 = timeit(lambda: plain.guide(x), args.steps)
    out_p = piloted(x).cpu().numpy()
    out_b = plain(x).cpu().numpy()
    g_p = piloted.guide(x).cpu().numpy()
    g_b = plain.guide(x).cpu().numpy()
    p_noisy, p_plain, p_pilot = synth.psnr(n, clean), synth.psnr(out_b, clean), synth.psnr(out_p, clean)
    # the same comparison on other seeded frames of the recipe: the pilot's
    # loss depends on the image (its gratings)
    others = {}
    for sd in (77, 5, 11):
        c2, n2 = synth.grey_pair(H, W, SIGMA, sd)
        x2 = torch.from_numpy(n2).cuda()
        others[str(sd)] = round(synth.psnr(piloted(x2).cpu().numpy(), c2) - synth.psnr(plain(x2).cpu().numpy(), c2), 3)
    # the network's flops: 2*9*cin*cout per pixel per layer
    from isp_mw_amd.pilot import LAYERS
    fl = 2.0 * 9 * sum(ci * co for ci, co, _, _, _ in LAYERS) * W * H
    res = {
        "metric": "Mpix/s learned-pilot + Wiener BM3D, 1080p grey sigma=25 (vs plain two-step)",
        "value": round(W * H / dt_pilot / 1e6, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 2,
        "ms_per_step": round(dt_pilot * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 network, " + ("fp32" if args.precision == "fp32" else f"{args.precision}+fp32") + " Wiener",
        "data": "synthetic (seeded rectangles+gradient+gratings, AWGN)",
        "config": {"workload": "pilot CNN (8 layers 3x3, 48 ch, ReLU, residual) as Wiener guide, 1920x1080 grey",
                   "sigma": SIGMA, "bm3d_sigma_param": round(SIGMA * P.GREY_SIGMA_SCALE, 3),
                   "precision": args.precision},
        "plain_two_step": {"value": round(W * H / dt_plain / 1e6, 3), "ms": round(dt_plain * 1e3, 3)},
        "guide_ms": {"pilot_cnn": round(dt_guide_pilot * 1e3, 3), "bm3d_step1": round(dt_guide_step1 * 1e3, 3)},
        "pilot_cnn_tflops": round(fl / dt_guide_pilot / 1e12, 2),
        "quality": {"psnr_noisy_db": round(p_noisy, 3), "psnr_plain_two_step_db": round(p_plain, 3),
                    "psnr_pilot_db": round(p_pilot, 3), "gain_db": round(p_pilot - p_plain, 3),
                    "psnr_guide_pilot_db": round(synth.psnr(g_p, clean), 3),
                    "psnr_guide_step1_db": round(synth.psnr(g_b, clean), 3),
                    "gain_db_other_seeds": others},
        "training": training,
    }
    print(json.dumps(res), flush=True)
    return 0


def bench_full(args) -> int:
    """Full form: 8 frames of 3840x2160 RGB (one seeded base image, frames
    shifted by random 0..7 px), sigma 25, BMrange 19, GroupSize 32 in both
    steps.  The RGB learned pilot (the three OPP planes of a frame at once,
    trained on step-1 targets with these parameters) gives the guides of all 8
    frames in one network pass; the Wiener step runs on all 8 frames in one
    batched graph launch with FP16 matrix-core transforms.  Compared with the
    plain FP32 GPU two-step path on every frame, and with step 1 as the guide
    in the same batched FP16 form."""
    import torch
    from isp_mw_amd import BM3D, BM3D_Final, BM3D_Para, BM3DBatch, pilot as P, synth

    FW_, FH_, NF = 3840, 2160, 8
    rng = np.random.default_rng(4)
    base = synth.clean_rgb(FH_ + 8, FW_ + 8, 99)
    frames, cleans = [], []
    for k in range(NF):
        dy, dx = rng.integers(0, 8, 2)
        c = np.ascontiguousarray(base[:, dy:dy + FH_, dx:dx + FW_])
        cleans.append(c)
        frames.append((c + rng.standard_normal(c.shape) * (SIGMA / 255.0)).astype(np.float32))
    del base
    para = BM3D_Para("np")
    for st in (para.basic, para.final):
        st.BlockStep, st.BMrange, st.BMstep, st.GroupSize = 3, 19, 1, 32
    para.basic.lambda_ = 2.7
    para.set_sigma(SIGMA)
    # the pilot: the RGB network, trained on step-1 targets of these parameters
    net, training = pilot_network(P, "rgb", args.pilot_steps)
    x_all = torch.from_numpy(np.stack(frames)).cuda()
    xs = list(x_all)
    guide_all = torch.empty_like(x_all)
    guides = list(guide_all)
    outs = [torch.empty_like(x) for x in xs]
    fin = BM3D_Final(para.final, precision="fp16")
    batch = BM3DBatch(fin, outs, xs, refs=guides)

    def guides_of_frames():
        P.rgb_guide(net, x_all, SIGMA, out=guide_all)   # one network pass over the 8 frames' OPP planes

    def full_form():
        guides_of_frames()
        batch.run()

    full_form()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        full_form()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        guides_of_frames()
    torch.cuda.synchronize()
    dt_pilot = (time.perf_counter() - t0) / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.run()
    torch.cuda.synchronize()
    dt_wiener = (time.perf_counter() - t0) / args.steps
    psnr_full = float(np.mean([synth.psnr(o.cpu().numpy(), c) for o, c in zip(outs, cleans)]))
    # plain FP32 GPU two-step, frame by frame
    plain = BM3D(para, precision="fp32")
    ref_out = [torch.empty_like(x) for x in xs]
    t0 = time.perf_counter()
    for x, o in zip(xs, ref_out):
        plain.process(o, x)
    torch.cuda.synchronize()
    dt_plain = time.perf_counter() - t0
    psnr_plain = float(np.mean([synth.psnr(o.cpu().numpy(), c) for o, c in zip(ref_out, cleans)]))
    # step 1 as the guide in the same batched FP16 form (the method without the network)
    from isp_mw_amd import BM3D_Basic
    est1 = [torch.empty_like(x) for x in xs]
    b1 = BM3DBatch(BM3D_Basic(para.basic, precision="fp16"), est1, xs)
    b1.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b1.run()
    torch.cuda.synchronize()
    dt_step1 = (time.perf_counter() - t0) / args.steps
    # the FP16 two-step path without the pilot, batched, for reference
    both16 = BM3D(para, precision="fp16")
    out16 = [torch.empty_like(x) for x in xs]
    b16 = BM3DBatch(both16, out16, xs)
    b16.run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b16.run()
    torch.cuda.synchronize()
    dt16 = time.perf_counter() - t0
    psnr16 = float(np.mean([synth.psnr(o.cpu().numpy(), c) for o, c in zip(out16, cleans)]))
    px = float(NF * FW_ * FH_)
    # algorithmic work of the full form: the CNN on 3 planes per frame, and the
    # Wiener step's 3 transforms x 3 planes at full N plus its matching
    cnn = 2.0 * 9 * sum(ci * co for ci, co, _, _, _ in P.layers(3)) * px
    gx, gy = grid(FW_, 3), grid(FH_, 3)
    refs = gx.size * gy.size
    cand = float(window_len(gx, FW_, 19).sum() * window_len(gy, FH_, 19).sum() - refs)
    wiener = NF * (refs * 3 * 3 * transform_flops(32) + cand * B * B * 3)
    achieved = (cnn + wiener) / dt / 1e12
    res = {
        "metric": "Mpix/s full form: 8 x 4K RGB, pilot + batched FP16 Wiener, R=19, G=32",
        "value": round(px / dt / 1e6, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 network + fp16 Wiener transforms (fp32 accumulation)",
        "data": "synthetic RGB base, 8 frames shifted by random 0..7 px, AWGN sigma 25/255",
        "config": {"workload": "8 frames 3840x2160 RGB, pilot guide + one batched Wiener launch",
                   "block_step": 3, "bm_range": 19, "group_size": [32, 32], "sigma": SIGMA},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK_F16_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_F16_TFLOPS, 4), "traffic": None,
                     "basis": "CNN flops + Wiener transforms at full N + matching, over the whole form"},
        "stages_ms": {"pilot_cnn_8_frames": round(dt_pilot * 1e3, 2), "wiener_batch": round(dt_wiener * 1e3, 2),
                      "step1_batch_8_frames": round(dt_step1 * 1e3, 2)},
        "pilot_gain": {"guide_speedup_vs_step1": round(dt_step1 / dt_pilot, 2),
                       "note": "the pilot's guides of the 8 frames against batched FP16 step 1 (the same guides' "
                               "role) with these parameters"},
        "quality": {"psnr_full_form_db": round(psnr_full, 3), "psnr_plain_fp32_two_step_db": round(psnr_plain, 3),
                    "psnr_delta_vs_plain_fp32_db": round(psnr_full - psnr_plain, 3),
                    "psnr_fp16_two_step_batch_db": round(psnr16, 3)},
        "plain_fp32_two_step": {"value": round(px / dt_plain / 1e6, 3), "ms": round(dt_plain * 1e3, 1)},
        "fp16_two_step_batch": {"value": round(px / dt16 / 1e6, 3), "ms": round(dt16 * 1e3, 1)},
        "training": training,
    }
    print(json.dumps(res), flush=True)
    return 0


def bench_stack(args) -> int:
    """Multi-frame grouping on the full-form frames: 8 frames of 3840x2160 RGB
    (one seeded base image, frames shifted by random 0..7 px, AWGN sigma 25),
    BlockStep 3, BMrange 19, GroupSize 32 in both steps, two-step BM3D whose
    groups of frame t take blocks from frames t-1..t+1 (temporal radius 1),
    in the timed precision's form.  Compared with the same form frame by frame
    (radius 0) and, for PSNR, with the plain FP32 form at radius 1."""
    import torch
    from isp_mw_amd import BM3D, BM3D_Para, synth

    FW_, FH_, NF, RADIUS = 3840, 2160, 8, 1
    rng = np.random.default_rng(4)
    base = synth.clean_rgb(FH_ + 8, FW_ + 8, 99)
    cleans = np.empty((NF, 3, FH_, FW_), np.float32)
    for k in range(NF):
        dy, dx = rng.integers(0, 8, 2)
        cleans[k] = base[:, dy:dy + FH_, dx:dx + FW_]
    del base
    noisy = cleans + (rng.standard_normal(cleans.shape) * (SIGMA / 255.0)).astype(np.float32)
    para = BM3D_Para("np")
    for st in (para.basic, para.final):
        st.BlockStep, st.BMrange, st.BMstep, st.GroupSize = 3, 19, 1, 32
    para.basic.lambda_ = 2.7
    para.set_sigma(SIGMA)
    x = torch.from_numpy(noisy).cuda()
    y = torch.empty_like(x)
    f = BM3D(para, precision=args.precision)

    def run(radius, out, filt=f):
        filt.process_stack(out, x, radius=radius)

    run(RADIUS, y)
    torch.cuda.synchronize()
    f.set_timing(True)
    f.timing(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run(RADIUS, y)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    stages = f.timing()
    f.set_timing(False)
    out = y.cpu().numpy()

    def psnr_stack(a):
        return float(np.mean([synth.psnr(a[k], cleans[k]) for k in range(NF)]))

    psnr_r1 = psnr_stack(out)
    # the same form frame by frame (radius 0)
    y0 = torch.empty_like(x)
    run(0, y0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(0, y0)
    torch.cuda.synchronize()
    dt0 = time.perf_counter() - t0
    psnr_r0 = psnr_stack(y0.cpu().numpy())
    del y0
    extra = {}
    if args.precision != "fp32":
        f32 = BM3D(para, precision="fp32")
        y32 = torch.empty_like(x)
        t0 = time.perf_counter()
        run(RADIUS, y32, f32)
        torch.cuda.synchronize()
        extra = {"plain_fp32_radius1": {"value": round(NF * FW_ * FH_ / (time.perf_counter() - t0) / 1e6, 3),
                                        "psnr_db": round(psnr_stack(y32.cpu().numpy()), 4)}}
        extra["psnr_loss_vs_fp32_db"] = round(extra["plain_fp32_radius1"]["psnr_db"] - psnr_r1, 4)
        del y32
    px = float(NF * FW_ * FH_)
    # issued matrix work of the dominant filter: 12 MFMA 32x32x16 per transform,
    # 3 planes x (2 forward + 1 inverse) for the Wiener step, 3 x 2 for step 1
    gx, gy = grid(FW_, 3), grid(FH_, 3)
    refs = NF * gx.size * gy.size
    per = {k: v[0] / v[1] for k, v in stages.items() if v[1]}
    dom = max(("filter_final", "filter_basic", "match_final", "match_basic"), key=lambda k: stages[k][0])
    ms_total = stages[dom][0] / args.steps
    if dom.startswith("filter") and args.precision != "fp32":
        ntr = 3 * (3 if dom == "filter_final" else 2) * refs
        flops = ntr * 12 * 32 * 32 * 16 * 2.0
        achieved = flops / (ms_total * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": round(achieved, 2), "peak": PEAK_F16_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_F16_TFLOPS, 4), "traffic": None, "kernel": dom,
                "ms_per_step": round(ms_total, 3), "basis": "issued MFMA work per step (8 frames), all launches"}
    else:
        cand = float(window_len(gx, FW_, 19).sum() * window_len(gy, FH_, 19).sum()) * NF * (2 * RADIUS + 1)
        flops = cand * B * B * 3 if dom.startswith("match") else refs * 3 * 3 * transform_flops(32)
        # the plain form's filter: FP32-accurate products at three FP16 MFMAs each
        peak = PEAK_FP32_TFLOPS if dom.startswith("match") else PEAK_F16_TFLOPS / 3
        achieved = flops / (ms_total * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), "traffic": None, "kernel": dom,
                "ms_per_step": round(ms_total, 3),
                "basis": "matching: 3 flop per pixel per candidate over 3 frames; filter: dense transforms at N=32"}
    res = {
        "metric": "Mpix/s 8 x 4K RGB two-step BM3D with temporal groups (radius 1), R=19, G=32",
        "value": round(px / dt / 1e6, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.precision == "fp32" else f"{args.precision}+fp32",
        "data": "synthetic RGB base, 8 frames shifted by random 0..7 px, AWGN sigma 25/255",
        "config": {"workload": "8 frames 3840x2160 RGB float32 stack, temporal radius 1", "block_step": 3,
                   "bm_range": 19, "group_size": [32, 32], "sigma": SIGMA, "precision": args.precision},
        "roofline": roof,
        "quality": {"psnr_noisy_db": round(psnr_stack(noisy), 4), "psnr_radius1_db": round(psnr_r1, 4),
                    "psnr_radius0_db": round(psnr_r0, 4), "temporal_gain_db": round(psnr_r1 - psnr_r0, 4),
                    **{k: v for k, v in extra.items() if k == "psnr_loss_vs_fp32_db"}},
        "stages_ms_per_launch": {k: round(v, 4) for k, v in per.items()},
        "radius0_same_form": {"value": round(px / dt0 / 1e6, 3), "ms": round(dt0 * 1e3, 1)},
        **{k: v for k, v in extra.items() if k != "psnr_loss_vs_fp32_db"},
    }
    print(json.dumps(res), flush=True)
    return 0


def nlm_window(pos: np.ndarray, n: int, R: int, s: int) -> np.ndarray:
    # clipped search window with BMstep s (include/Block.h:716-751): k*s offsets
    # with |k*s| <= R and 0 <= pos + k*s <= n - B
    last = n - B
    return np.minimum(R // s, (last - pos) // s) + np.minimum(R // s, pos // s) + 1


def bench_nlm(args) -> int:
    """NL-means with NLMeans_Para's defaults (correction, sigma 8: strength 40,
    thMSE 400, BlockStep 5, GroupSize 16, BMrange 24, BMstep 3) of a 1920x1080
    RGB frame of 16-bit samples (8-bit values * 257, the reference's
    ImageReader frame), AWGN sigma 8/255, frames resident in HBM."""
    import torch
    from isp_mw_amd import NLMeans, NLMeans_Para, synth

    sigma = 8.0
    para = NLMeans_Para(correction=True, sigma=sigma)
    clean, noisy = synth.rgb_pair(H, W, sigma, SEED)
    frame = np.clip(np.rint(noisy * 255), 0, 255).astype(np.int32) * 257
    dev = torch.device("cuda:0")
    x = torch.from_numpy(frame).to(dev)

    def timed(prec, steps, warmup):
        f = NLMeans(para, precision=prec)
        y = torch.empty_like(x)
        for _ in range(warmup):
            f.process(y, x, floor=0, ceil=65535)
        torch.cuda.synchronize()
        f.set_timing(True)
        f.timing(reset=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            f.process(y, x, floor=0, ceil=65535)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / steps
        st = f.timing()
        f.set_timing(False)
        return dt, st, y.cpu().numpy()

    dt, stages, out = timed(args.precision, args.steps, args.warmup)
    value = W * H / dt / 1e6
    c16 = clean * 255 * 257
    psnr = lambda a: 10 * np.log10(65535.0 ** 2 / np.mean((np.asarray(a, np.float64) - c16) ** 2))  # noqa: E731
    other = None
    if args.precision != "fp32":
        dt32, st32, out32 = timed("fp32", max(2, args.steps // 2), 1)
        other = {"form": "exact (FP64 distances)", "value": round(W * H / dt32 / 1e6, 3),
                 "ms_per_step": round(dt32 * 1e3, 3), "psnr_db": round(psnr(out32), 4),
                 "stages_ms": {k: round(v[0] / v[1], 4) for k, v in st32.items() if v[1]}}
    gx, gy = grid(W, para.BlockStep), grid(H, para.BlockStep)
    refs = gx.size * gy.size
    R = para.BMrange // para.BMstep * para.BMstep
    cand = float(nlm_window(gx, W, R, para.BMstep).sum() * nlm_window(gy, H, R, para.BMstep).sum() - refs)
    match_flops = cand * B * B * 3            # difference, square, add per pixel per candidate (luma)
    # average: each pixel's 3 samples read once, 3 fixed-point sums written once,
    # the group lists (count, G positions, G keys) read once
    avg_bytes = W * H * 3 * 4 + W * H * 3 * 8 + refs * (4 + para.GroupSize * 12)
    dom = max(("nlm_match", "nlm_average"), key=lambda k: stages[k][0])
    ms = stages[dom][0] / stages[dom][1]
    if dom == "nlm_match":
        achieved = match_flops / (ms * 1e-3) / 1e12
        peak = PEAK_FP64_TFLOPS if args.precision == "fp32" else PEAK_FP32_TFLOPS
        roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), "traffic": None, "kernel": dom, "ms_per_launch": round(ms, 4),
                "engine": "fp64 VALU" if args.precision == "fp32" else "fp32 VALU",
                "algorithmic_flops_per_launch": match_flops,
                "flops_basis": "3 flop per pixel per candidate, clipped BMstep-3 windows, luma"}
    else:
        achieved = avg_bytes / (ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": None, "kernel": dom,
                "ms_per_launch": round(ms, 4), "algorithmic_bytes_per_launch": avg_bytes}
    res = {
        "metric": "Mpix/s NL-means 1080p RGB 16-bit, NLMeans_Para defaults",
        "value": round(value, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp32" else "f32 matching + f64 averaging",
        "data": "synthetic (seeded rectangles+gradient+gratings, mixed RGB, AWGN sigma 8/255, 8-bit * 257)",
        "config": {"workload": "NL-means, 1920x1080 RGB int32 frame (Floor 0, Ceil 65535)",
                   "correction": True, "sigma": sigma, "strength": para.strength, "thMSE": para.thMSE,
                   "block": 8, "block_step": para.BlockStep, "group_size": para.GroupSize,
                   "bm_range": para.BMrange, "bm_step": para.BMstep, "precision": args.precision},
        "roofline": roof,
        "quality": {"psnr_noisy_db": round(psnr(frame), 4), "psnr_out_db": round(psnr(out), 4)},
        "stages": {k: {"ms_per_launch": round(v[0] / v[1], 4) if v[1] else None, "launches": v[1]}
                   for k, v in stages.items()},
        "exact_form": other,
        "work": {"refs": refs, "match_candidates": cand, "match_flops": match_flops,
                 "average_algorithmic_bytes": avg_bytes},
    }
    if not args.no_cpu:
        import oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        oracle.set_threads(threads)
        op = oracle.nlm_params(True, sigma)
        hc = 64
        t0 = time.perf_counter()
        oracle.nlm_rgb_int(op, frame[:, :hc], floor=0, ceil=65535)
        rate = W * hc / max(time.perf_counter() - t0, 1e-6)
        hc = int(min(H, max(64, rate * args.cpu_seconds / W)))
        times = []
        for _ in range(2):
            t0 = time.perf_counter()
            oracle.nlm_rgb_int(op, frame[:, :hc], floor=0, ceil=65535)
            times.append(time.perf_counter() - t0)
        t = min(times)
        res["cpu_baseline"] = {"value": round(W * hc / t / 1e6, 4), "unit": "Mpix/s", "cores": threads,
                               "kind": "port",
                               "sample": f"oracle (FP64 C restatement, {threads} threads over grid rows) NL-means "
                                         f"on the top {W}x{hc} crop of the same frame, best of 2 runs of {t:.2f} s"}
        res["gpu_over_cpu"] = round(value / res["cpu_baseline"]["value"], 1)
    print(json.dumps(res), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--precision", default="fp16", choices=("fp32", "fp16", "bf16"),
                    help="transform/shrinkage precision of the timed form (fp16: the metric's mixed form)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--tune", action="store_true", help="autotune the tile shapes of the frame's size class first "
                                                        "(default: the built-in shapes, which measure as fast)")
    ap.add_argument("--tiles", default="", help="mx,my,fx,fy: given tile shapes of the timed form (no autotuning; "
                                                "for profiler runs, whose averages the tuner's trials would mix)")
    ap.add_argument("--cpu-seconds", type=float, default=4.0, help="target seconds per CPU baseline run")
    ap.add_argument("--frames-per-step", type=int, default=24,
                    help="frames of the headline workload per step (a step times a batch of frames, so that "
                         "the default 20 steps keep the GPU busy for about a second)")
    ap.add_argument("--streams", type=int, default=2,
                    help="frames in flight in the timed region of the headline workload, one context and one "
                         "stream each (1: one frame after another)")
    ap.add_argument("--workload", default="rgb1080",
                    choices=("rgb1080", "batch64", "pilot1080", "full4k", "nlm1080", "stack4k"),
                    help="rgb1080: the headline 1080p RGB frame; batch64: 64 uneven grey images as one batch; "
                         "pilot1080: train the learned pilot, then pilot + Wiener on 1080p grey; "
                         "full4k: 8 x 4K RGB frames, pilot guides + one batched FP16 Wiener launch; "
                         "nlm1080: NL-means of a 1080p 16-bit RGB frame; "
                         "stack4k: 8 x 4K RGB frames with temporal groups (radius 1)")
    ap.add_argument("--pilot-steps", type=int, default=0,
                    help="pilot workloads: 0 uses the shipped weights (isp_mw_amd/weights); N > 0 trains a network "
                         "for N steps by the same recipe first")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 256 MiB device-to-device copy (3 launches): the known-traffic "
                         "kernel that scales rocprofv3 FETCH_SIZE / WRITE_SIZE readings")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        a = torch.empty(64 * 1024 * 1024, dtype=torch.float32, device="cuda:0").fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_each_way": a.numel() * 4, "launches": 3}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(images shard across GPUs with no collectives, see the design notes)", file=sys.stderr)
        return 2
    if args.workload == "batch64":
        return bench_batch(args)
    if args.workload == "pilot1080":
        return bench_pilot(args)
    if args.workload == "full4k":
        return bench_full(args)
    if args.workload == "nlm1080":
        return bench_nlm(args)
    if args.workload == "stack4k":
        return bench_stack(args)

    import torch
    from isp_mw_amd import BM3D, synth

    para = make_para()
    clean, noisy = synth.rgb_pair(H, W, SIGMA, SEED)
    dev = torch.device("cuda:0")
    x = torch.from_numpy(noisy).to(dev)
    y = torch.empty_like(x)
    f = BM3D(para, device=0, precision=args.precision)
    # tile shapes of this size class from the autotuner, or given (tiles never change results)
    tuning = None
    if args.tiles:
        mx, my, fx, fy = (int(v) for v in args.tiles.split(","))
        tuning = {"match_tile": (mx, my), "filter_tile": (fx, fy), "final_group_size": 0, "given": True}
        f.set_tuning(f.size_class(W, H), tuning)
    elif args.tune:
        tuning = f.autotune(W, H, SIGMA, rgb=True)

    # value: K frames, args.streams of them in flight (one context and stream
    # each), no stage events; then the same K frames one after another, and
    # once more with per-stage HIP events for the stage times and the roofline
    nstr = max(1, args.streams)
    fs = [f] + [BM3D(para, device=0, precision=args.precision) for _ in range(nstr - 1)]
    if tuning is not None:
        for g in fs[1:]:
            g.set_tuning(g.size_class(W, H), {k: v for k, v in tuning.items() if k != "given"})
    ys = [y] + [torch.empty_like(x) for _ in range(nstr - 1)]
    F = max(1, args.frames_per_step)
    nfr = args.steps * F
    dt = pipelined_steps(fs, ys, x, nfr, args.warmup)          # seconds per frame
    value = W * H / dt / 1e6
    contexts_equal = all(bool(torch.equal(ys[0], o)) for o in ys[1:])
    dt1 = pipelined_steps(fs[:1], ys[:1], x, nfr, 1) if nstr > 1 else dt
    dti, stages = timed_steps(f, y, x, nfr, 1)
    out = y.cpu().numpy()
    psnr_out = synth.psnr(out, clean)
    psnr_in = synth.psnr(noisy, clean)
    quality_extra = {}
    fp32_form = None
    if args.precision != "fp32":
        # the metric's bar: PSNR loss against the plain FP32 GPU path, timed beside it
        f32 = BM3D(para, device=0, precision="fp32")
        if args.tune and not args.tiles:
            f32.autotune(W, H, SIGMA, rgb=True)
        y32 = torch.empty_like(x)
        dt32, st32 = timed_steps(f32, y32, x, max(2, nfr // 4), 1)
        ref32 = y32.cpu().numpy()
        quality_extra = {"psnr_fp32_db": round(synth.psnr(ref32, clean), 4),
                         "psnr_loss_vs_fp32_db": round(synth.psnr(ref32, clean) - psnr_out, 4),
                         "max_abs_vs_fp32": float(np.abs(out - ref32).max())}
        fp32_form = {"value": round(W * H / dt32 / 1e6, 3), "ms_per_step": round(dt32 * 1e3, 3),
                     "stages_ms": {k: round(v[0] / v[1], 4) for k, v in st32.items() if v[1]}}
        # the FP32-throughout form (FP32 distances, no FP64 re-decision) beside it
        ff = BM3D(para, device=0, precision="fp32fast")
        yf = torch.empty_like(x)
        dtf, stf = timed_steps(ff, yf, x, max(2, nfr // 4), 1)
        of = yf.cpu().numpy()
        fp32_form["fp32fast"] = {"value": round(W * H / dtf / 1e6, 3), "ms_per_step": round(dtf * 1e3, 3),
                                 "stages_ms": {k: round(v[0] / v[1], 4) for k, v in stf.items() if v[1]},
                                 "psnr_db": round(synth.psnr(of, clean), 4),
                                 "max_abs_vs_fp32": float(np.abs(of - ref32).max())}
        quality_extra["psnr_loss_vs_fp32fast_db"] = round(synth.psnr(of, clean) - psnr_out, 4)
        # the same arithmetic with FP32-accurate products on the matrix cores
        fm = BM3D(para, device=0, precision="fp32mfma")
        ym = torch.empty_like(x)
        dtm, stm = timed_steps(fm, ym, x, max(2, nfr // 4), 1)
        om = ym.cpu().numpy()
        fp32_form["fp32mfma"] = {"value": round(W * H / dtm / 1e6, 3), "ms_per_step": round(dtm * 1e3, 3),
                                 "stages_ms": {k: round(v[0] / v[1], 4) for k, v in stm.items() if v[1]},
                                 "psnr_db": round(synth.psnr(om, clean), 4),
                                 "max_abs_vs_fp32": float(np.abs(om - ref32).max()),
                                 "max_abs_vs_fp32fast": float(np.abs(om - of).max())}
    # end-to-end rate with host buffers (PCIe both ways); never the headline value
    hin = noisy
    hout = np.empty_like(hin)
    f.process(hout, hin)
    host_dev = [float(np.abs(hout - out).max())]
    t0 = time.perf_counter()
    for _ in range(2):
        f.process(hout, hin)
    host_rate = W * H / ((time.perf_counter() - t0) / 2) / 1e6
    host_dev.append(float(np.abs(hout - out).max()))

    # group sizes of the last run (actual-N work) and full-N work
    gshape = f.grid_shape(W, H, False)
    n1 = f.last_groups(False, gshape).astype(np.float64)
    n2 = f.last_groups(True, f.grid_shape(W, H, True)).astype(np.float64)
    refs = n1.size
    gx, gy = grid(W, 3), grid(H, 3)
    cand = float(window_len(gx, W, 12).sum() * window_len(gy, H, 12).sum() - refs)
    match_flops = cand * B * B * 3              # per step, luma only
    filt2_actual = 3 * 3 * transform_flops(n2).sum()   # 3 channels x 3 transforms
    filt1_actual = 3 * 2 * transform_flops(n1).sum()
    filt2_full = refs * 3 * 3 * transform_flops(32)
    per_stage = {k: {"ms_per_launch": round(v[0] / v[1], 4) if v[1] else None, "launches": v[1]}
                 for k, v in stages.items()}
    # dominant kernel = the stage with the largest share of device time
    dom = max(("filter_final", "filter_basic", "match_basic", "match_final"), key=lambda k: stages[k][0])
    ms = stages[dom][0] / stages[dom][1]
    if dom.startswith("filter"):
        flops = filt2_actual if dom == "filter_final" else filt1_actual
        full = filt2_full if dom == "filter_final" else refs * 3 * 2 * transform_flops(16)
        achieved = flops / (ms * 1e-3) / 1e12
        # the plain form's FP32-accurate products take three FP16 MFMAs each
        # (split operands): its peak is a third of the dense FP16 one
        split = args.precision == "fp32"
        peak = PEAK_F16_TFLOPS / 3 if split else PEAK_F16_TFLOPS
        roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": round(peak, 1), "unit": "TFLOP/s",
                "frac": round(achieved / peak, 4), "traffic": None, "kernel": dom,
                "ms_per_launch": round(ms, 4),
                "engine": ("fp16 MFMA 32x32x16, three per FP32-accurate product (hi/lo split operands)" if split
                           else f"{args.precision} MFMA 32x32x16"),
                "algorithmic_flops_per_launch": flops,
                "flops_basis": "dense separable 2*B^2*N*(2B+N) per transform at each group's actual N",
                # the same count with every group at its full GroupSize (the work-count basis that
                # states the path's work independently of the image)
                "achieved_full_n": round(full / (ms * 1e-3) / 1e12, 3),
                "frac_full_n": round(full / (ms * 1e-3) / 1e12 / peak, 4),
                "algorithmic_flops_per_launch_full_n": full}
        util = measured_util(args.precision, dom)
        if util is not None and util.get("mfma_insts_per_launch"):
            # issued matrix work: the counted 32x32x16 MFMAs of one launch, against the dense FP16 peak
            roof["issued_mfma_flops_per_launch"] = util["mfma_insts_per_launch"] * 32 * 32 * 16 * 2.0
            roof["issued_frac"] = round(roof["issued_mfma_flops_per_launch"] / (ms * 1e-3) / 1e12 / PEAK_F16_TFLOPS,
                                        4)
    else:
        achieved = match_flops / (ms * 1e-3) / 1e12
        roof = {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / PEAK_FP64_TFLOPS, 4), "traffic": None, "kernel": dom,
                "ms_per_launch": round(ms, 4),
                "engine": "fp64 VALU", "algorithmic_flops_per_launch": match_flops,
                "flops_basis": "3 flop (sub+FMA) per pixel per candidate, clipped windows"}

    util = measured_util(args.precision, dom)
    if util is not None:
        # counter-measured pipe utilisation of the same kernel and form
        roof["pipe_utilisation"] = util
    rp = profiled_ms(args.precision, dom)
    if rp is not None:
        # the committed rocprofv3 kernel-trace average of the same kernel: it times
        # a short profiled run (one frame at a time), the events time this run
        roof["rocprof_ms_per_launch"] = rp
        roof["rocprof_over_event_ms"] = round(rp / ms, 3)
    tr = measured_traffic(args.precision, dom)
    if tr is not None:
        roof["traffic"] = tr["bytes"]
        roof["traffic_detail"] = tr
        roof["traffic_gbs_at_measured_time"] = round(tr["bytes"] / (ms * 1e-3) / 1e9, 1)

    res = {
        "metric": "Mpix/s two-step BM3D 1080p RGB sigma=25 at <=0.05 dB PSNR vs plain FP32 GPU",
        "value": round(value, 3),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(dt * F * 1e3, 3),
        "frames_per_step": F,
        "ms_per_frame": round(dt * 1e3, 4),
        "frames_in_flight": nstr,
        "one_frame_at_a_time": {"value": round(W * H / dt1 / 1e6, 3), "ms_per_frame": round(dt1 * 1e3, 4)},
        "instrumented": {"value": round(W * H / dti / 1e6, 3), "ms_per_frame": round(dti * 1e3, 4),
                         "note": "per-stage HIP events on the launch stream; the stage times and the roofline"},
        "contexts_equal": contexts_equal,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.precision == "fp32" else f"{args.precision}+fp32",
        "data": "synthetic (seeded rectangles+gradient+gratings, mixed RGB, AWGN sigma 25/255)",
        "config": {"workload": "two-step BM3D, 1920x1080 RGB float32, sigma=25, OPP colour",
                   "frames_per_step": F, "block": 8, "block_step": 3, "bm_range": 12, "bm_step": 1,
                   "group_size": 32, "group_size_basic": 16, "lambda": 2.7, "precision": args.precision},
        "roofline": roof,
        "quality": {"psnr_noisy_db": round(psnr_in, 3), "psnr_out_db": round(psnr_out, 3),
                    "mean_group_size": [round(float(n1.mean()), 3), round(float(n2.mean()), 3)],
                    **quality_extra},
        "stages": per_stage,
        "fp32_form": fp32_form,
        "tuning": tuning,
        "host_buffers_mpix_s": round(host_rate, 3),
        "host_buffers_max_abs_vs_device": host_dev,
        "work": {"refs_per_step": refs, "match_candidates_per_step": cand, "match_flops_per_step": match_flops,
                 "filter_final_flops_actual_n": filt2_actual, "filter_final_flops_full_n": filt2_full,
                 "filter_basic_flops_actual_n": filt1_actual},
    }
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
        res["cpu_baseline"] = cpu_baseline(para, noisy, threads, args.cpu_seconds)
        res["gpu_over_cpu"] = round(value / res["cpu_baseline"]["value"], 1)
        # the reference's own behaviour: one thread (source/BM3D.cpp runs single-threaded)
        res["cpu_baseline_single_thread"] = cpu_baseline(para, noisy, 1, args.cpu_seconds)
        res["gpu_over_cpu_single_thread"] = round(value / res["cpu_baseline_single_thread"]["value"], 1)
    print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
