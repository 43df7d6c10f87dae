"""bench.py's host-side pieces, on the CPU: the reference grid and the clipped
window counts behind its algorithmic-work figures (brute force on small sizes,
and the 1080p counts), the transform flop count (a MAC count of the separable
loop nest), the readers of the committed rocprofv3 summaries, and the one-GPU
contract of --gpus."""
import os
import subprocess
import sys

import numpy as np
import pytest

import bench
from tests import npref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def brute_candidates(W, H, s, R):
    # every reference block's window positions, clipped to the image, self excluded
    n = 0
    for j in npref.grid(H, s):
        for i in npref.grid(W, s):
            ys = [y for y in range(j - R, j + R + 1) if 0 <= y <= H - 8]
            xs = [x for x in range(i - R, i + R + 1) if 0 <= x <= W - 8]
            n += len(ys) * len(xs) - 1
    return n


@pytest.mark.parametrize("W,H,s,R", [(40, 33, 3, 12), (64, 20, 4, 9), (17, 50, 3, 5), (8, 8, 3, 12)])
def test_candidate_count_matches_brute_force(W, H, s, R):
    gx, gy = bench.grid(W, s), bench.grid(H, s)
    assert list(gx) == npref.grid(W, s) and list(gy) == npref.grid(H, s)
    cand = bench.window_len(gx, W, R).sum() * bench.window_len(gy, H, R).sum() - gx.size * gy.size
    assert cand == brute_candidates(W, H, s, R)


def test_1080p_counts():
    gx, gy = bench.grid(1920, 3), bench.grid(1080, 3)
    assert (gx.size, gy.size, gx[-1], gy[-1]) == (639, 359, 1912, 1072)
    assert gx.size * gy.size == 229401
    cand = bench.window_len(gx, 1920, 12).sum() * bench.window_len(gy, 1080, 12).sum() - gx.size * gy.size
    assert cand == 141454248


@pytest.mark.parametrize("n", [1, 2, 7, 16, 32])
def test_transform_flops_is_the_separable_mac_count(n):
    macs = 0
    macs += n * 8 * 8 * 8      # rows: N*8 rows, 8 outputs of 8 terms
    macs += n * 8 * 8 * 8      # columns
    macs += 64 * n * n         # group axis: 64 pixels, N outputs of N terms
    assert bench.transform_flops(n) == 2 * macs
    assert np.allclose(bench.transform_flops(np.array([n, n])), 2 * macs)


@pytest.mark.parametrize("prec", ["fp16", "fp32"])
@pytest.mark.parametrize("stage", ["filter_final", "filter_basic", "match_basic", "match_final"])
def test_committed_profiles_parse(prec, stage):
    t = bench.measured_traffic(prec, stage)
    assert t is not None and t["fetch_bytes"] > 0 and t["write_bytes"] > 0
    assert t["bytes"] == t["fetch_bytes"] + t["write_bytes"]
    u = bench.measured_util(prec, stage)
    # valu_busy sums SQ_ACTIVE_INST_VALU over a SIMD's waves, whose VALU
    # instructions overlap in the pipeline: it passes 1 a little at the issue limit
    assert u is not None and 0 < u["valu_busy"] <= 1.25 and 0 <= u["mfma_busy"] <= 1
    # both forms filter on the matrix cores (fp32: split-FP16 products); matching is vector work
    assert (u["mfma_insts_per_launch"] > 0) == stage.startswith("filter")
    assert u["ms_under_counters"] > 0
    ms = bench.profiled_ms(prec, stage)
    assert ms is not None and 0.05 < ms < 50


def test_unknown_stage_reads_nothing():
    assert bench.measured_traffic("fp16", "no_such_stage") is None
    assert bench.measured_util("fp16", "no_such_stage") is None
    assert bench.measured_util("int8", "filter_final") is None
    assert bench.profiled_ms("fp16", "no_such_stage") is None


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, "bench.py", "--gpus", "2"], cwd=ROOT, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 2, r.stderr
    assert "one GPU" in r.stderr
