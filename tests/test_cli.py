"""The `isp_mw --bm3d` / `--nlm` front end (isp_mw_amd/cli/isp_mw.cpp): the
flag and file semantics of ISP_MW's BM3D_IO / NLMeans_IO / FilterIO
(include/BM3D.h:363-526, include/NLMeans.h:62-171, include/Filter.h:220-340),
its PNG / PNM / PFM codecs against PIL (an independent implementation,
test-only), and on the GPU its output against the library's integer Frame
paths and the oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

from isp_mw_amd import build as B

CLI = B.CLI


def run(*args, check=True):
    r = subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=300)
    if check and r.returncode != 0:
        raise AssertionError(r.stdout + r.stderr)
    return r


def params(*args):
    return json.loads(run("--bm3d", "in.png", "--print-params", *args).stdout)


def test_defaults_follow_the_reference():
    p = params()
    # BM3D_Para() is profile "fast", sigma {10,10,10}; thMSE_Default 400+80s / 200+10s
    assert p["profile"] == "fast" and p["basic"]["sigma"] == [10, 10, 10] and p["final"]["sigma"] == [10, 10, 10]
    assert p["basic"]["BlockStep"] == 8 and p["final"]["BlockStep"] == 7 and p["basic"]["BMrange"] == 9
    assert p["basic"]["thMSE"] == 400 + 80 * 10 and p["final"]["thMSE"] == 200 + 10 * 10
    assert p["output"] == "in.BM3D.png"


def test_sigma_repeats_and_thmse_recomputed():
    p = params("-P", "np", "-S", "20", "-S", "10")
    assert p["basic"]["sigma"] == [20, 10, 10] and p["final"]["sigma"] == [20, 10, 10]
    assert p["basic"]["thMSE"] == 400 + 80 * 20 and p["final"]["thMSE"] == 200 + 10 * 20
    assert p["basic"]["GroupSize"] == 16 and p["final"]["GroupSize"] == 32 and p["final"]["BlockStep"] == 3
    p = params("-P", "vn", "-S", "5")
    assert p["basic"]["thMSE"] == 1000 + 150 * 5 and p["final"]["thMSE"] == 400 + 40 * 5
    assert p["basic"]["lambda"] == 2.8 and p["final"]["BlockSize"] == 11


def test_explicit_values_win():
    p = params("-TH1", "3000", "-TH2", "123", "-LD", "3.1", "-BS1", "8", "-BSP1", "5", "-GS1", "8", "-MR1", "7",
               "-MS1", "2", "-BSP2", "4", "-GS2", "24", "-MR2", "11", "-MS2", "1")
    b, f = p["basic"], p["final"]
    assert (b["thMSE"], f["thMSE"], b["lambda"]) == (3000, 123, 3.1)
    assert (b["BlockStep"], b["GroupSize"], b["BMrange"], b["BMstep"]) == (5, 8, 7, 2)
    assert (f["BlockStep"], f["GroupSize"], f["BMrange"], f["BMstep"]) == (4, 24, 11, 1)


def test_precision_names():
    for prec in ("fp32", "fp32fast", "fp32mfma", "fp16", "bf16"):
        assert params("--precision", prec)["precision"] == prec
    assert run("--bm3d", "in.png", "--print-params", "--precision", "fp8", check=False).returncode != 0


def test_output_naming_and_errors():
    p = json.loads(run("--bm3d", "dir/a.b.png", "--print-params", "-T", "_x", "-F", "ppm").stdout)
    assert p["output"] == "dir/a.b_x.ppm"
    assert "Not enough arguments" in run("--bm3d").stdout
    assert run("--gaussian", "x.png", check=False).returncode == 1
    r = run("--bm3d", "x.png", "-S", check=False)
    assert r.returncode == 2 and "No argument specified for option -S" in r.stderr
    assert run("--bm3d", "x.png", "--bogus", "1", check=False).returncode == 2
    assert run("--bm3d", "missing.png", check=False).returncode == 2


def nlm_params(*args):
    return json.loads(run("--nlm", "in.png", "--print-params", *args).stdout)


def test_nlm_defaults_follow_the_reference():
    # NLMeans_Para (include/NLMeans.h:11-22), tag ".NLMeans"
    p = nlm_params()
    assert p["filter"] == "nlm" and p["correction"] is True and p["sigma"] == 8 and p["strength"] == 40
    assert (p["BlockSize"], p["BlockStep"], p["GroupSize"], p["BMrange"], p["BMstep"]) == (8, 5, 16, 24, 3)
    assert p["thMSE"] == 400 and p["output"] == "in.NLMeans.png"


def test_nlm_flags_and_derived_values():
    # strength and thMSE follow the final sigma and correction unless given (:148-149)
    p = nlm_params("-S", "12", "-C", "FALSE")
    assert (p["correction"], p["strength"], p["thMSE"]) == (False, 18, 300)
    p = nlm_params("-C", "false", "-S", "12", "-H", "2.5", "-TH", "77", "-BS", "8", "-BSP", "3", "-GS", "8",
                   "-MR", "9", "-MS", "1", "-T", "_x", "-F", "ppm")
    assert (p["strength"], p["thMSE"], p["BlockStep"], p["GroupSize"], p["BMrange"], p["BMstep"]) == \
        (2.5, 77, 3, 8, 9, 1)
    assert p["output"] == "in_x.ppm"
    for name in ("--NLMeans", "--nonlocalmeans", "--NLM"):
        assert json.loads(run(name, "a.png", "--print-params").stdout)["filter"] == "nlm"
    r = run("--nlm", "x.png", "-C", "maybe", check=False)
    assert r.returncode == 2 and 'must be "true" or "false"' in r.stderr


PIL = pytest.importorskip("PIL.Image")


@pytest.mark.parametrize("mode,dtype", [("RGB", np.uint8), ("L", np.uint8), ("RGBA", np.uint8), ("LA", np.uint8),
                                        ("P", np.uint8), ("I;16", np.uint16)])
def test_png_decoder_against_pil(tmp_path, mode, dtype):
    rng = np.random.default_rng(len(mode))
    h, w = 23, 37
    if mode == "P":
        img = PIL.fromarray(rng.integers(0, 255, (h, w, 3), dtype=np.uint8), "RGB").quantize(17)
    elif mode == "I;16":
        img = PIL.fromarray(rng.integers(0, 65535, (h, w), dtype=np.uint16))
    else:
        ch = {"RGB": 3, "L": 1, "RGBA": 4, "LA": 2}[mode]
        # noise on gradients makes PIL's encoder pick Sub / Up / Average / Paeth rows
        grad = (np.arange(w) * 3) % 192
        if ch > 1:
            a = rng.integers(0, 64, (h, w, ch)) + grad[None, :, None]
        else:
            a = rng.integers(0, 64, (h, w)) + grad[None, :]
        img = PIL.fromarray(a.astype(np.uint8), mode)
    src = tmp_path / "in.png"
    img.save(src, optimize=False)
    out = run("--bm3d", src, "--copy").stdout.strip()
    got = np.asarray(PIL.open(out).convert("RGB"), np.int64)
    rgb = img.convert("RGB") if mode != "I;16" else None
    if mode == "I;16":
        v = np.asarray(img, np.int64)
        want = np.repeat(((v + 128) // 257)[..., None], 3, axis=2)   # 16-bit samples, RoundDiv to 8 bits
    else:
        want = np.asarray(rgb, np.int64)
    np.testing.assert_array_equal(got, want)


def test_png_encoder_read_by_pil(tmp_path):
    rng = np.random.default_rng(7)
    a = rng.integers(0, 256, (31, 45, 3), dtype=np.uint8)
    src = tmp_path / "in.ppm"
    src.write_bytes(b"P6\n45 31\n255\n" + a.tobytes())
    out = run("--bm3d", src, "--copy", "-F", ".png").stdout.strip()
    np.testing.assert_array_equal(np.asarray(PIL.open(out)), a)


def test_pnm_and_pfm_round_trips(tmp_path):
    rng = np.random.default_rng(8)
    g = rng.integers(0, 65536, (12, 17), dtype=np.uint16)
    src = tmp_path / "g16.pgm"
    src.write_bytes(b"P5\n# a comment\n17 12\n65535\n" + g.astype(">u2").tobytes())
    out = run("--bm3d", src, "--copy", "-F", ".ppm").stdout.strip()
    data = open(out, "rb").read()
    assert data.startswith(b"P6\n17 12\n255\n")
    got = np.frombuffer(data[len(b"P6\n17 12\n255\n"):], np.uint8).reshape(12, 17, 3)
    np.testing.assert_array_equal(got, np.repeat(((g.astype(np.int64) + 128) // 257)[..., None], 3, axis=2))
    f = rng.uniform(0, 1, (9, 11, 3)).astype(np.float32)
    pfm = tmp_path / "f.pfm"
    pfm.write_bytes(b"PF\n11 9\n-1.0\n" + f[::-1].astype("<f4").tobytes())   # rows bottom to top
    out = run("--bm3d", pfm, "--copy", "-F", ".pfm").stdout.strip()
    data = open(out, "rb").read()
    hdr = b"PF\n11 9\n-1.0\n"
    assert data.startswith(hdr)
    np.testing.assert_array_equal(np.frombuffer(data[len(hdr):], "<f4").reshape(9, 11, 3)[::-1], f)


def test_bad_headers_rejected(tmp_path):
    # a short IHDR chunk, and PNM / PFM dimensions that are negative, zero or
    # too large: the reader stops with an error instead of reading past the data
    import struct
    import zlib

    def chunk(t, data):
        return struct.pack(">I", len(data)) + t + data + struct.pack(">I", zlib.crc32(t + data) & 0xFFFFFFFF)
    png = b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">II", 4, 4) + b"\x08\x02\x00") + chunk(b"IEND", b"")
    bad = [("short.png", png),
           ("huge.png", b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", 1 << 31, 4, 8, 2, 0, 0, 0))),
           ("neg.ppm", b"P6\n-4 4\n255\n" + bytes(48)),
           ("zero.pgm", b"P5\n0 4\n255\n"),
           ("wrap.ppm", b"P6\n4294967297 1\n255\n" + bytes(3)),
           ("big.pfm", b"PF\n100000 100000\n-1.0\n" + bytes(12))]
    for name, data in bad:
        f = tmp_path / name
        f.write_bytes(data)
        r = run("--bm3d", f, "--copy", check=False)
        assert r.returncode != 0, name
        assert "Traceback" not in r.stderr


# ---------------------------------------------------------------- GPU
def noisy_png(tmp_path, h=48, w=64, seed=3):
    from isp_mw_amd import synth
    clean, n = synth.rgb_pair(h, w, 25.0, seed)
    a = np.clip(np.rint(np.transpose(n, (1, 2, 0)) * 255), 0, 255).astype(np.uint8)
    p = tmp_path / "noisy.png"
    PIL.fromarray(a, "RGB").save(p)
    return p, a


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp32", "fp32mfma", "fp16"])
def test_cli_equals_library_integer_frame_path(gpu, tmp_path, prec):
    from isp_mw_amd import BM3D, BM3D_Para
    p, a = noisy_png(tmp_path)
    out = run("--bm3d", p, "-P", "np", "-S", "25", "--precision", prec, "-T", f".{prec}").stdout.strip()
    got = np.asarray(PIL.open(out), np.int64)
    para = BM3D_Para("np").set_sigma(25.0)
    frame = np.ascontiguousarray(np.transpose(a, (2, 0, 1)).astype(np.int32) * 257)
    res = BM3D(para, precision=prec)(frame, floor=0, ceil=65535)
    want = np.transpose((res.astype(np.int64) + 128) // 257, (1, 2, 0))
    np.testing.assert_array_equal(got, want)


@pytest.mark.gpu
@pytest.mark.parametrize("flags", [("-P", "vn", "-S", "20"),
                                   ("-P", "np", "-S", "20", "-BS1", "6", "-GS1", "40", "-BS2", "10", "-GS2", "48")],
                         ids=["vn", "BS-GS"])
def test_cli_block_and_group_sizes(gpu, tmp_path, flags):
    # BlockSize other than 8 and GroupSize above 32 through the front end
    # (include/BM3D.h:416-452), equal to the library's integer Frame path
    from isp_mw_amd import BM3D
    p, a = noisy_png(tmp_path, seed=6)
    out = run("--bm3d", p, *flags).stdout.strip()
    got = np.asarray(PIL.open(out), np.int64)
    prm = params(*flags)
    from isp_mw_amd import BM3D_Para
    para = BM3D_Para(prm["profile"])
    for st, d in ((para.basic, prm["basic"]), (para.final, prm["final"])):
        st.BlockSize, st.BlockStep, st.GroupSize = d["BlockSize"], d["BlockStep"], d["GroupSize"]
        st.BMrange, st.BMstep, st.thMSE = d["BMrange"], d["BMstep"], d["thMSE"]
        st.sigma = list(d["sigma"])
    para.basic.lambda_ = prm["basic"]["lambda"]
    frame = np.ascontiguousarray(np.transpose(a, (2, 0, 1)).astype(np.int32) * 257)
    res = BM3D(para)(frame, floor=0, ceil=65535)
    want = np.transpose((res.astype(np.int64) + 128) // 257, (1, 2, 0))
    np.testing.assert_array_equal(got, want)


@pytest.mark.gpu
def test_cli_against_oracle_and_guide(gpu, tmp_path):
    import oracle
    from isp_mw_amd import BM3D_Para
    p, a = noisy_png(tmp_path, seed=4)
    out = run("--bm3d", p, "-P", "np", "-S", "25", "-T", ".o").stdout.strip()
    got = np.asarray(PIL.open(out), np.int64)
    para = BM3D_Para("np").set_sigma(25.0)
    frame = np.transpose(a, (2, 0, 1)).astype(np.int32) * 257
    ref16 = oracle.bm3d_rgb_int(para.basic, para.final, frame, mode=3, floor=0, ceil=65535)
    want = np.transpose((ref16.astype(np.int64) + 128) // 257, (1, 2, 0))
    d = np.abs(got - want)
    print(f"cli vs oracle: max {d.max()} levels, {np.mean(d > 0):.2e} of samples differ", flush=True)
    assert d.max() <= 1 and np.mean(d > 0) < 1e-3
    # a guide frame (--ref) other than the input: it steers step 1's matching
    # (its luma), so the output follows the oracle run with the same guide
    from isp_mw_amd import synth
    clean = synth.rgb_pair(48, 64, 25.0, 4)[0]
    ga = np.clip(np.rint(np.transpose(clean, (1, 2, 0)) * 255), 0, 255).astype(np.uint8)
    g = tmp_path / "guide.png"
    PIL.fromarray(ga, "RGB").save(g)
    out2 = run("--bm3d", p, "-P", "np", "-S", "25", "--ref", g, "-T", ".g").stdout.strip()
    got2 = np.asarray(PIL.open(out2), np.int64)
    guide16 = np.transpose(ga, (2, 0, 1)).astype(np.int32) * 257
    ref16g = oracle.bm3d_rgb_int(para.basic, para.final, frame, ref=guide16, mode=3, floor=0, ceil=65535)
    want2 = np.transpose((ref16g.astype(np.int64) + 128) // 257, (1, 2, 0))
    d2 = np.abs(got2 - want2)
    changed = float(np.mean(got2 != got))
    print(f"cli --ref vs oracle: max {d2.max()} levels, {np.mean(d2 > 0):.2e} differ; "
          f"{changed:.3f} of samples changed by the guide", flush=True)
    assert d2.max() <= 1 and np.mean(d2 > 0) < 1e-3
    assert changed > 0.05  # an ignored guide would leave the output as it was


@pytest.mark.gpu
def test_cli_nlm_equals_library_and_oracle(gpu, tmp_path):
    import oracle
    from isp_mw_amd import NLMeans, NLMeans_Para
    p, a = noisy_png(tmp_path, seed=5)
    out = run("--nlm", p, "-S", "20").stdout.strip()
    assert out.endswith("noisy.NLMeans.png")
    got = np.asarray(PIL.open(out), np.int64)
    para = NLMeans_Para(sigma=20.0)
    frame = np.ascontiguousarray(np.transpose(a, (2, 0, 1)).astype(np.int32) * 257)
    res = NLMeans(para)(frame, floor=0, ceil=65535)
    want = np.transpose((res.astype(np.int64) + 128) // 257, (1, 2, 0))
    np.testing.assert_array_equal(got, want)
    op = oracle.nlm_params(True, 20.0)
    ref16 = oracle.nlm_rgb_int(op, frame, floor=0, ceil=65535)
    d = np.abs(got - np.transpose((ref16.astype(np.int64) + 128) // 257, (1, 2, 0)))
    assert d.max() <= 1 and np.mean(d > 0) < 1e-3
    # a float image is not a Frame
    pfm = tmp_path / "f.pfm"
    pfm.write_bytes(b"PF\n16 16\n-1.0\n" + np.zeros((16, 16, 3), "<f4").tobytes())
    assert run("--nlm", pfm, check=False).returncode == 2
