"""Parity at the benchmark's full size: 1920x1080 RGB, sigma 25, two-step BM3D
(BlockStep 3, BMrange 12, GroupSize 16/32), HIP path vs the FP64 oracle."""
import os

import numpy as np
import pytest

import oracle
from isp_mw_amd import BM3D, synth
from tests.gpu_common import MIN_GROUPS, assert_close, group_identity, para

pytestmark = pytest.mark.gpu


def threads():
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)


def test_1080p_rgb_two_step(gpu):
    import torch
    print("1080p RGB two-step: GPU", flush=True)
    clean, noisy = synth.rgb_pair(1080, 1920, 25.0, 1)
    p = para()
    f = BM3D(p)
    x = torch.from_numpy(noisy).to(gpu)
    out = f(x).cpu().numpy()
    g1 = f.last_groups(False, (359, 639), with_pos=True)
    g2 = f.last_groups(True, (359, 639), with_pos=True)
    print("1080p RGB two-step: oracle", flush=True)
    oracle.set_threads(threads())
    fast = oracle.host_supports_fast()
    oracle.set_threads(threads(), fast=fast)
    ref = oracle.bm3d_rgb(p.basic, p.final, noisy, fast=fast)
    assert_close(out, ref, clean, label="1080p rgb")
    yuv = oracle.rgb2opp(noisy)
    oc, op, _ = oracle.match_all(p.basic, yuv[0])
    frac1 = group_identity(*g1, oc, op)
    print(f"step-1 groups identical: {frac1:.6f}", flush=True)
    assert frac1 >= MIN_GROUPS
    basic = oracle.kernel(p.basic, list(yuv), list(yuv), wiener=False, fast=fast)
    oc2, op2, _ = oracle.match_all(p.final, basic[0])
    frac2 = group_identity(*g2, oc2, op2)
    print(f"step-2 groups (GPU basic vs oracle basic as guide) identical: {frac2:.6f}", flush=True)
    assert frac2 >= MIN_GROUPS
    oracle.set_threads(1)
    oracle.set_threads(1, fast=fast)
