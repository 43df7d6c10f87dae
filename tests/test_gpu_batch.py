"""Batches (isp_bm3d_batch_*) and the autotuner on the GPU.

A batch runs every stage of many images in one launch each, replayed as a
hipGraph.  Each image's output must equal a single call on that image with the
same parameters bit for bit: matching per reference block does not depend on
the launch, and aggregation is integer (order-independent).  Tile shapes do
not change results either.
"""
import numpy as np
import pytest

import oracle
from isp_mw_amd import BM3D, BM3D_Basic, BM3DBatch, synth
from tests.gpu_common import assert_close, para

pytestmark = pytest.mark.gpu

SIZES = [(64, 80), (96, 72), (40, 56), (128, 100), (33, 47)]
SIGMAS = [10.0, 25.0, 40.0, 0.0, 18.0]   # 0: the context's (25)


def say(msg):
    print(msg, flush=True)


def single(p_sigma, n, prec, cls=BM3D, **kw):
    import torch
    p = para(sigma=p_sigma, **kw)
    f = cls(p if cls is BM3D else p.basic, precision=prec)
    x = torch.from_numpy(n).cuda()
    return f(x).cpu().numpy()


@pytest.mark.parametrize("prec", ["fp32", "fp16", "fp32fast", "fp32mfma"])
def test_grey_batch_equals_single_calls(gpu, prec):
    import torch
    imgs = [synth.grey_pair(h, w, s or 25.0, 50 + k)[1] for k, ((h, w), s) in enumerate(zip(SIZES, SIGMAS))]
    f = BM3D(para(), precision=prec)
    srcs = [torch.from_numpy(a).to(gpu) for a in imgs]
    dsts = [torch.empty_like(t) for t in srcs]
    b = BM3DBatch(f, dsts, srcs, sigmas=SIGMAS)
    say(f"batch {prec}: {b.info()}")
    b.run()
    torch.cuda.synchronize()
    for k, a in enumerate(imgs):
        want = single(SIGMAS[k] or 25.0, a, prec)
        got = dsts[k].cpu().numpy()
        say(f"  image {k} {a.shape} sigma {SIGMAS[k]}: max |batch - single| = {np.abs(got - want).max():.3g}")
        np.testing.assert_array_equal(got, want)


def test_rgb_batch_equals_single_calls(gpu):
    import torch
    shapes = [(48, 64), (72, 56), (40, 40)]
    sig = [25.0, 12.0, 35.0]
    imgs = [synth.rgb_pair(h, w, s, 70 + k)[1] for k, ((h, w), s) in enumerate(zip(shapes, sig))]
    f = BM3D(para(), precision="fp16")
    srcs = [torch.from_numpy(a).to(gpu) for a in imgs]
    dsts = [torch.empty_like(t) for t in srcs]
    b = BM3DBatch(f, dsts, srcs, sigmas=sig)
    b.run()
    torch.cuda.synchronize()
    for k, a in enumerate(imgs):
        np.testing.assert_array_equal(dsts[k].cpu().numpy(), single(sig[k], a, "fp16"))


def test_rgb_batch_with_guides_and_basic_mode(gpu):
    import torch
    shapes = [(48, 64), (40, 48)]
    pairs = [synth.rgb_pair(h, w, 25.0, 90 + k) for k, (h, w) in enumerate(shapes)]
    f = BM3D_Basic(para().basic)
    srcs = [torch.from_numpy(n).to(gpu) for _, n in pairs]
    refs = [torch.from_numpy(c.astype(np.float32)).to(gpu) for c, _ in pairs]
    dsts = [torch.empty_like(t) for t in srcs]
    b = BM3DBatch(f, dsts, srcs, refs=[refs[0], None])
    b.run()
    torch.cuda.synchronize()
    g = BM3D_Basic(para().basic)
    want0 = g.process(np.empty_like(pairs[0][1]), pairs[0][1], pairs[0][0].astype(np.float32))
    want1 = g(pairs[1][1])
    np.testing.assert_array_equal(dsts[0].cpu().numpy(), want0)
    np.testing.assert_array_equal(dsts[1].cpu().numpy(), want1)


def test_batch_against_oracle(gpu):
    import torch
    pairs = [synth.grey_pair(56, 64, 25.0, 3), synth.grey_pair(48, 40, 25.0, 4)]
    p = para()
    f = BM3D(p)
    srcs = [torch.from_numpy(n).to(gpu) for _, n in pairs]
    dsts = [torch.empty_like(t) for t in srcs]
    BM3DBatch(f, dsts, srcs).run()
    torch.cuda.synchronize()
    for k, (c, n) in enumerate(pairs):
        assert_close(dsts[k].cpu().numpy(), oracle.bm3d_plane(p.basic, p.final, n), c, label=f"batch image {k}")


def test_graph_replays_with_eager_work_between(gpu):
    import torch
    imgs = [synth.grey_pair(h, w, 25.0, 20 + k)[1] for k, (h, w) in enumerate(SIZES[:3])]
    f = BM3D(para(), precision="fp16")
    srcs = [torch.from_numpy(a).to(gpu) for a in imgs]
    dsts = [torch.empty_like(t) for t in srcs]
    b = BM3DBatch(f, dsts, srcs)
    b.run(graph=False)
    torch.cuda.synchronize()
    ref = [d.cpu().numpy() for d in dsts]
    other = BM3D(para(sigma=40.0), precision="fp16")
    for rep in range(3):
        for d in dsts:
            d.fill_(-1.0)
        other(srcs[0])                 # eager work on the same device between replays
        _ = (srcs[1] * 2).sum().item()
        b.run(graph=True)
        torch.cuda.synchronize()
        for k in range(3):
            np.testing.assert_array_equal(dsts[k].cpu().numpy(), ref[k], err_msg=f"replay {rep} image {k}")


def test_update_and_degenerate_image(gpu):
    import torch
    imgs = [synth.grey_pair(h, w, 25.0, 30 + k)[1] for k, (h, w) in enumerate(SIZES[:3])]
    f = BM3D(para(sigma=0.0))          # the context's sigma 0: images without their own copy through
    srcs = [torch.from_numpy(a).to(gpu) for a in imgs]
    dsts = [torch.empty_like(t) for t in srcs]
    b = BM3DBatch(f, dsts, srcs, sigmas=[25.0, 0.0, 25.0])
    b.run()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dsts[1].cpu().numpy(), imgs[1])
    np.testing.assert_array_equal(dsts[0].cpu().numpy(), single(25.0, imgs[0], "fp32"))
    new = [torch.from_numpy(a[::-1].copy()).to(gpu) for a in imgs]
    b.update(dsts, new, sigmas=[25.0, 0.0, 25.0])
    b.run()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(dsts[2].cpu().numpy(), single(25.0, imgs[2][::-1].copy(), "fp32"))


def test_autotune_keeps_results_and_auto_group(gpu):
    import torch
    f = BM3D(para(), precision="fp16")
    _, n = synth.grey_pair(200, 160, 25.0, 7)
    x = torch.from_numpy(n).to(gpu)
    before = f(x).cpu().numpy()
    t = f.autotune(160, 200, 25.0, auto_group=True)
    say(f"autotune 160x200: {t}")
    cls = f.size_class(160, 200)
    assert f.get_tuning(cls)["match_tile"] == t["match_tile"]
    after = f(x).cpu().numpy()
    np.testing.assert_array_equal(after, before)   # tiles never change results
    # a tuned final GroupSize of 16 applies to batches that ask for it
    f.set_tuning(cls, {"match_tile": t["match_tile"], "filter_tile": t["filter_tile"], "final_group_size": 16})
    d = torch.empty_like(x)
    BM3DBatch(f, [d], [x], auto_group=True).run()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d.cpu().numpy(), single(25.0, n, "fp16", G2=16))
    d2 = torch.empty_like(x)
    BM3DBatch(f, [d2], [x]).run()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(d2.cpu().numpy(), before)


@pytest.mark.parametrize("rgb", [False, True])
def test_final_step_batch_with_guides(gpu, rgb):
    import torch
    from isp_mw_amd import BM3D_Final
    shapes = [(48, 64), (56, 40), (40, 72)]
    pairs = [(synth.rgb_pair if rgb else synth.grey_pair)(h, w, 25.0, 110 + k) for k, (h, w) in enumerate(shapes)]
    p = para()
    f = BM3D_Final(p.final, precision="fp16")
    srcs = [torch.from_numpy(n).to(gpu) for _, n in pairs]
    # guides: the clean image for the first two, none (src) for the last
    refs = [torch.from_numpy(c.astype(np.float32)).to(gpu) for c, _ in pairs[:2]] + [None]
    dsts = [torch.empty_like(t) for t in srcs]
    BM3DBatch(f, dsts, srcs, refs=refs).run()
    torch.cuda.synchronize()
    g = BM3D_Final(p.final, precision="fp16")
    for k, (c, n) in enumerate(pairs):
        ref = c.astype(np.float32) if k < 2 else n
        want = g.process(np.empty_like(n), n, ref)
        np.testing.assert_array_equal(dsts[k].cpu().numpy(), want, err_msg=f"image {k}")


def test_rgb_autotune_keeps_results(gpu):
    import torch
    f = BM3D(para(), precision="fp16")
    _, n = synth.rgb_pair(120, 200, 25.0, 9)
    x = torch.from_numpy(n).to(gpu)
    before = f(x).cpu().numpy()
    t = f.autotune(200, 120, 25.0, rgb=True)
    say(f"rgb autotune 200x120: {t}")
    np.testing.assert_array_equal(f(x).cpu().numpy(), before)


@pytest.mark.parametrize("prec", ["fp32", "fp16"])
def test_batch64_scale(gpu, prec):
    """The many-image batch at its full scale: 64 grey images, W uniform in [256, 1024],
    H in [256, 768], per-image sigma in [5, 50], each generated like the
    512x512 case, in one batch (one launch per stage, replayed as a hipGraph)
    with autotuned tiles and final GroupSize per size class.  Every image
    equals a single call on it with the same tiles and GroupSize, bit for
    bit; two images (the smallest and a large one) against the oracle."""
    import torch
    rng = np.random.default_rng(2024)
    n_img = 64
    Ws = rng.integers(256, 1025, n_img)
    Hs = rng.integers(256, 769, n_img)
    sig = rng.uniform(5.0, 50.0, n_img)
    pairs = [synth.grey_pair(int(h), int(w), float(s), 1000 + k) for k, (w, h, s) in enumerate(zip(Ws, Hs, sig))]
    p = para()
    f = BM3D(p, precision=prec)
    classes = {}
    for w, h in zip(Ws, Hs):
        classes.setdefault(f.size_class(int(w), int(h)), []).append((int(w), int(h)))
    for cls, dims in classes.items():
        w, h = sorted(dims)[len(dims) // 2]
        t = f.autotune(w, h, 25.0, auto_group=True)
        say(f"size class {cls} ({len(dims)} images): tiles {t['match_tile']} / {t['filter_tile']}, "
            f"final GroupSize {t['final_group_size'] or p.final.GroupSize}")
    srcs = [torch.from_numpy(nz).cuda() for _, nz in pairs]
    dsts = [torch.empty_like(x) for x in srcs]
    b = BM3DBatch(f, dsts, srcs, sigmas=list(sig), auto_group=True)
    say(f"batch {prec}: {b.info()}")
    b.run()
    torch.cuda.synchronize()
    worst = 0.0
    for k, (x, d) in enumerate(zip(srcs, dsts)):
        one = BM3DBatch(f, [torch.empty_like(x)], [x], sigmas=[float(sig[k])], auto_group=True)
        one.run()
        torch.cuda.synchronize()
        diff = float((one._keep[0][0] - d).abs().max())
        worst = max(worst, diff)
        assert diff == 0.0, (k, diff)
        one.close()
    say(f"64 images: max |batch - single image| = {worst}")
    # against the oracle: the single-image semantics with that image's sigma and GroupSize
    for k in (int(np.argmin(Ws * Hs)), 7):
        q = para(sigma=float(sig[k]))
        tg = f.get_tuning(f.size_class(int(Ws[k]), int(Hs[k])))
        if tg and tg["final_group_size"]:
            q.final.GroupSize = tg["final_group_size"]
        clean, nz = pairs[k]
        ref = oracle.bm3d_plane(q.basic, q.final, nz)
        got = dsts[k].cpu().numpy()
        err = float(np.abs(got - ref).max())
        d = synth.psnr(got, clean) - synth.psnr(ref, clean)
        say(f"image {k} {Ws[k]}x{Hs[k]} sigma {sig[k]:.1f}: max |batch - oracle| {err:.3g}, dPSNR {d:+.4f} dB")
        if prec == "fp32":
            assert err <= 2e-3 and abs(d) <= 0.01
        else:
            assert d >= -0.05
