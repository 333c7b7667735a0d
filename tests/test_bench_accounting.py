"""The benchmark's roofline accounting at 2^24 (no device): the algorithmic work per kernel stage,
checked against the counts worked out by hand from the shapes, and the whole path's floor."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    sys.path.insert(0, ROOT)
    import bench as b
    return b


def test_stage_work_at_xl(lbn, bench):
    params = lbn.Params.sized(lbn.SUITES[3])  # 2^24 F162, 1024 columns, 4096 ring elements a column
    n, r, L = 1 << 24, 1024, 4096
    assert (params.witness_len(), params.columns(), params.len_ring()) == (n, r, L)
    stage = {"commit_layout": 0.2, "commit_gemm": 3.5, "commit_finish": 0.01, "fold_challenges": 0.03,
             "fold_gemm": 1.0, "fold_intt": 0.05, "eval_eq": 0.05, "eval_reduce": 0.6}
    out, floor, frac = bench.stage_rooflines(params, stage, 6.0)
    # W (24 B an element) plus two byte layouts: 4 groups of 256 columns x 8 fragments x
    # 81 x 4096 / 64 steps x 2 KB for the commit's, 16 groups of 256 ring elements x 8 x
    # 81 x 1024 / 64 x 2 KB for the fold's
    assert out["commit_layout"]["work"] == 24 * n + 4 * 8 * 5184 * 2048 + 16 * 8 * 1296 * 2048
    # the challenges' transform (8 tiles x 16 steps x 12 KB) written and read, the fold's A'
    # (8 x 1296 steps x 12 KB) written
    assert out["fold_challenges"]["work"] == 8 * (2 * 16 + 1296) * 12288
    # split limbs 3889 and 9721: 8 tiles of 81^2 MACs a ring element; quadratic 2917: 16 tiles;
    # four digit products, 2 ops each
    per_ring = 8 * 81 * 81 * 2 + 16 * 81 * 81
    assert out["commit_gemm"]["work"] == 2 * 4 * per_ring * (n // 4)
    assert out["fold_gemm"]["work"] == 8 * 81 * 81 * 4 * 2 * L * r
    assert out["eval_reduce"]["work"] == 24 * n
    assert out["commit_gemm"]["bound"] == "mfma" and out["eval_reduce"]["bound"] == "hbm"
    want = sum(v["work"] / (8000e9 if v["bound"] == "hbm" else 5000e12) for v in out.values()) * 1e3
    assert abs(floor - want) < 1e-3 and abs(frac - floor / 6.0) < 1e-3
    assert 1.9 < floor < 2.0
    # the commit GEMM executes 32 tiles x 4 groups x 5184 steps x 8 waves x 44 MFMAs of 32768 ops
    assert bench.executed_ops(params) == 32 * 4 * 5184 * 8 * 44 * 32768
