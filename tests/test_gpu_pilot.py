"""The learned pilot's HIP kernels (include/isp_pilot.h) against the FP64
oracle (oracle/pilot_oracle.py), and what the trained pilot does as the guide
of the Wiener step.

Tolerances: the convolutions accumulate exact BF16 products in FP32, so their
FP32 outputs agree with FP64 to ~1e-5 relative; the BF16 outputs to one BF16
rounding.  The whole-network gradients agree with the oracle run with the
same BF16 rounding of stored tensors (``bf16=True``) to 2% in norm per layer,
the residue of FP32 accumulation flipping a few BF16 roundings.
"""
import ctypes

import numpy as np
import pytest

from oracle import pilot_oracle as po

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def bits(a):
    f = np.ascontiguousarray(po.bf16_round(a), np.float32)
    return (f.view(np.uint32) >> 16).astype(np.uint16).view(np.int16)


def unbits(t):
    u = t.cpu().numpy().view(np.uint16).astype(np.uint32) << 16
    return u.view(np.float32).astype(np.float64)


def dev(a, torch):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def pad_ch(a, c):
    out = np.zeros(a.shape[:-1] + (c,))
    out[..., :a.shape[-1]] = a
    return out


CONFIGS = [(16, 48, 2, 48), (48, 48, 48, 48), (48, 16, 48, 1)]


@pytest.fixture(scope="module")
def P(gpu):
    from isp_mw_amd import pilot
    return pilot


def packed(P, torch, w, cin, cout, cinp, coutp):
    lib = P._plib()
    wf = torch.empty(coutp * lib.isp_pilot_kpad(cinp), dtype=torch.int16, device="cuda")
    wb = torch.empty(cinp * lib.isp_pilot_kpad(coutp), dtype=torch.int16, device="cuda")
    wt = dev(w.astype(np.float32), torch)
    assert lib.isp_pilot_pack(P._ptr(wt), P._ptr(wf), P._ptr(wb), cin, cout, cinp, coutp, None) == 0
    return wf, wb


@pytest.mark.parametrize("cinp,coutp,cin,cout", CONFIGS)
def test_conv_forward_vs_oracle(P, cinp, coutp, cin, cout):
    import torch
    rng = np.random.default_rng(cinp + coutp)
    N, H, W = 2, 37, 45   # partial tiles on both axes
    x = po.bf16_round(rng.uniform(-1, 1, (N, H, W, cin)))
    w = rng.standard_normal((cout, cin, 9)) * np.sqrt(2.0 / (cin * 9))
    b = rng.standard_normal(cout) * 0.1
    wf, _ = packed(P, torch, w, cin, cout, cinp, coutp)
    xt = dev(bits(pad_ch(x, cinp)), torch)
    bt = dev(pad_ch(b[None], coutp)[0].astype(np.float32), torch)
    out = torch.empty(N * H * W * coutp, dtype=torch.int16, device="cuda")
    out32 = torch.empty(N * H * W * coutp, dtype=torch.float32, device="cuda")
    lib = P._plib()
    for relu in (0, 1):
        assert lib.isp_pilot_conv3x3(P._ptr(xt), None, P._ptr(wf), P._ptr(bt), P._ptr(out), P._ptr(out32), N, H, W,
                                     cinp, coutp, relu, None) == 0
        torch.cuda.synchronize()
        ref = po.conv3x3(x, po.bf16_round(w), b)
        if relu:
            ref = np.maximum(ref, 0)
        got32 = out32.cpu().numpy().reshape(N, H, W, coutp).astype(np.float64)
        err = np.abs(got32[..., :cout] - ref).max()
        say(f"conv {cinp}->{coutp} relu {relu}: max |fp32 - oracle| = {err:.3g}")
        np.testing.assert_allclose(got32[..., :cout], ref, rtol=1e-5, atol=1e-5)
        got16 = unbits(out).reshape(N, H, W, coutp)
        np.testing.assert_allclose(got16[..., :cout], po.bf16_round(got32[..., :cout]), rtol=0, atol=0)
        assert np.all(got16[..., cout:] == 0)


def test_conv_masked_input_gradient(P):
    import torch
    rng = np.random.default_rng(5)
    N, H, W = 1, 19, 70
    cin, cout = 48, 48
    g = po.bf16_round(rng.standard_normal((N, H, W, cout)))
    mask = po.bf16_round(np.maximum(rng.standard_normal((N, H, W, cout)), 0))
    w = rng.standard_normal((cout, cin, 9)) * 0.1
    _, wb = packed(P, torch, w, cin, cout, 48, 48)
    out = torch.empty(N * H * W * cin, dtype=torch.int16, device="cuda")
    out32 = torch.empty(N * H * W * cin, dtype=torch.float32, device="cuda")
    lib = P._plib()
    assert lib.isp_pilot_conv3x3(P._ptr(dev(bits(g), torch)), P._ptr(dev(bits(mask), torch)), P._ptr(wb), None,
                                 P._ptr(out), P._ptr(out32), N, H, W, 48, 48, 0, None) == 0
    torch.cuda.synchronize()
    ref = po.conv3x3_dinput(g * (mask > 0), po.bf16_round(w))
    got = out32.cpu().numpy().reshape(N, H, W, cin)
    np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("cinp,coutp,cin,cout", CONFIGS)
def test_wgrad_vs_oracle(P, cinp, coutp, cin, cout):
    import torch
    rng = np.random.default_rng(cinp * 3 + coutp)
    N, H, W = 3, 21, 40
    x = po.bf16_round(rng.uniform(-1, 1, (N, H, W, cin)))
    dout = po.bf16_round(rng.standard_normal((N, H, W, cout)) * 0.01)
    mask = po.bf16_round(np.maximum(rng.standard_normal((N, H, W, cout)), 0)) if cout > 1 else None
    lib = P._plib()
    part = torch.empty(lib.isp_pilot_wgrad_partial_floats(N, H, W, cin, cout), dtype=torch.float32, device="cuda")
    dw = torch.zeros(cout * cin * 9, dtype=torch.float32, device="cuda")
    db = torch.zeros(cout, dtype=torch.float32, device="cuda")
    mt = dev(bits(pad_ch(mask, coutp)), torch) if mask is not None else None
    assert lib.isp_pilot_conv3x3_wgrad(P._ptr(dev(bits(pad_ch(x, cinp)), torch)),
                                       P._ptr(dev(bits(pad_ch(dout, coutp)), torch)), P._ptr(mt), P._ptr(part),
                                       P._ptr(dw), P._ptr(db), N, H, W, cinp, coutp, cin, cout, None) == 0
    torch.cuda.synchronize()
    dz = dout * (mask > 0) if mask is not None else dout
    rw, rb = po.conv3x3_wgrad(x, dz)
    scale = np.abs(rw).max()
    np.testing.assert_allclose(dw.cpu().numpy().reshape(cout, cin, 9), rw, rtol=1e-4, atol=1e-5 * scale)
    np.testing.assert_allclose(db.cpu().numpy(), rb, rtol=1e-4, atol=1e-5 * max(1e-12, np.abs(rb).max()))


def test_network_loss_and_gradients_vs_oracle(P):
    import torch
    rng = np.random.default_rng(11)
    net = P.PilotCNN(seed=4)
    N, H, W = 2, 20, 36
    noisy = rng.uniform(0, 1, (N, H, W)).astype(np.float32)
    target = np.clip(noisy + rng.standard_normal((N, H, W)).astype(np.float32) * 0.05, 0, 1).astype(np.float32)
    sigma = np.array([15.0, 30.0], np.float32)
    loss = net.loss_and_grads(dev(noisy, torch), dev(sigma, torch), dev(target, torch))
    torch.cuda.synchronize()
    L, grads = po.loss_and_grads(net.layer_params(), noisy.astype(np.float64), sigma, target.astype(np.float64),
                                 bf16=True)
    say(f"loss gpu {float(loss.item()):.6e} oracle {L:.6e}")
    assert abs(float(loss.item()) - L) <= 1e-4 * L
    for l, ((gw, gb), (rw, rb)) in enumerate(zip(net.layer_grads(), grads)):
        ew = np.linalg.norm(gw - rw) / max(1e-30, np.linalg.norm(rw))
        eb = np.linalg.norm(gb - rb) / max(1e-30, np.linalg.norm(rb))
        say(f"  layer {l}: relative error dw {ew:.2e} db {eb:.2e}")
        assert ew < 2e-2 and eb < 2e-2, (l, ew, eb)


def test_adam_matches_formula(P):
    import torch
    rng = np.random.default_rng(2)
    n = 1000
    p = rng.standard_normal(n).astype(np.float32)
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    pt, mt, vt = dev(p, torch), dev(m, torch), dev(v, torch)
    for step in (1, 2, 3):
        g = rng.standard_normal(n).astype(np.float32)
        assert P._plib().isp_pilot_adam(P._ptr(pt), P._ptr(dev(g, torch)), P._ptr(mt), P._ptr(vt), n, 1e-2, 0.9,
                                        0.999, 1e-8, step, None) == 0
        m = 0.9 * m + 0.1 * g
        v = 0.999 * v + 0.001 * g * g
        p = p - 1e-2 * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
    torch.cuda.synchronize()
    np.testing.assert_allclose(pt.cpu().numpy(), p, rtol=1e-5, atol=1e-6)


def test_training_reduces_loss(P):
    noisy, sigma, target = P.make_training_set(256, crop=64, seed=1)
    net = P.PilotCNN(seed=0)
    losses = P.train(net, noisy, sigma, target, steps=80, batch=16, lr=2e-3, seed=0)
    first, last = np.mean(losses[:10]), np.mean(losses[-10:])
    say(f"training loss {first:.3e} -> {last:.3e}")
    assert last < 0.5 * first


def test_pilot_as_wiener_guide(P):
    import torch
    from isp_mw_amd import synth
    noisy, sigma, target = P.make_training_set(2048, crop=64, seed=2)
    net = P.PilotCNN(seed=1)
    losses = P.train(net, noisy, sigma, target, steps=600, batch=32, lr=2e-3, seed=1)
    say(f"training loss {np.mean(losses[:20]):.3e} -> {np.mean(losses[-20:]):.3e}")
    clean, n = synth.grey_pair(256, 320, 25.0, 77)
    x = torch.from_numpy(n).cuda()
    para = P.default_para(25.0)
    plain = P.PilotBM3D(para)(x).cpu().numpy()           # plain step 1 as the guide
    basic = P.PilotBM3D(para).guide(x).cpu().numpy()
    piloted = P.PilotBM3D(para, net)
    guide = piloted.guide(x).cpu().numpy()
    out = piloted(x).cpu().numpy()
    p_noisy, p_plain, p_pilot = synth.psnr(n, clean), synth.psnr(plain, clean), synth.psnr(out, clean)
    say(f"PSNR noisy {p_noisy:.2f}, step-1 guide {synth.psnr(basic, clean):.2f}, "
        f"pilot guide {synth.psnr(guide, clean):.2f} (vs step 1 {synth.psnr(guide, basic):.2f}); "
        f"two-step {p_plain:.2f}, pilot + Wiener {p_pilot:.2f}: gain {p_pilot - p_plain:+.3f} dB")
    assert p_pilot > p_noisy + 5.0
    assert p_pilot > p_plain - 1.0


def test_opp_conversions_vs_oracle(gpu):
    import oracle
    import torch
    from isp_mw_amd import _lib as L
    rng = np.random.default_rng(8)
    H, W = 37, 61
    rgb = rng.uniform(-0.1, 1.1, (3, H, W)).astype(np.float32)
    x = torch.from_numpy(rgb).cuda()
    yuv = torch.empty_like(x)
    back = torch.empty_like(x)
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    lib = L.lib()
    assert lib.isp_bm3d_rgb_to_opp(p(yuv[0]), p(yuv[1]), p(yuv[2]), p(x[0]), p(x[1]), p(x[2]), W, H, W, None) == 0
    torch.cuda.synchronize()
    want = oracle.rgb2opp(rgb.astype(np.float64))
    np.testing.assert_allclose(yuv.cpu().numpy(), want, rtol=0, atol=2e-7)
    for clip in (0, 1):
        assert lib.isp_bm3d_opp_to_rgb(p(back[0]), p(back[1]), p(back[2]), p(yuv[0]), p(yuv[1]), p(yuv[2]), W, H, W,
                                       clip, None) == 0
        torch.cuda.synchronize()
        ref = oracle.opp2rgb(yuv.cpu().numpy().astype(np.float64), clip=bool(clip))
        np.testing.assert_allclose(back.cpu().numpy(), ref, rtol=0, atol=2e-7)
        if not clip:
            np.testing.assert_allclose(back.cpu().numpy(), rgb, rtol=0, atol=1e-6)


def test_pilot_guides_an_rgb_frame(P):
    import torch
    from isp_mw_amd import BM3D_Para, synth
    noisy, sigma, target = P.make_training_set(1024, crop=64, seed=3)
    net = P.PilotCNN(seed=2)
    P.train(net, noisy, sigma, target, steps=300, batch=32, lr=2e-3, seed=2)
    clean, n = synth.rgb_pair(96, 128, 25.0, 5)
    para = BM3D_Para("np")
    for st in (para.basic, para.final):
        st.BlockStep, st.BMrange = 3, 12
    para.set_sigma(25.0)
    x = torch.from_numpy(n).cuda()
    plain = P.PilotBM3D(para, rgb=True)(x).cpu().numpy()
    out = P.PilotBM3D(para, net, rgb=True)(x).cpu().numpy()
    pn, pp, po_ = synth.psnr(n, clean), synth.psnr(plain, clean), synth.psnr(out, clean)
    say(f"RGB: noisy {pn:.2f} dB, two-step {pp:.2f} dB, pilot + Wiener {po_:.2f} dB ({po_ - pp:+.3f})")
    assert po_ > pn + 4.0
