"""bench.py's byte accounting (host logic, no GPU): the algorithmic bytes per stage and the fused
sort + mark/emit stage of frames that export no stage."""
import sys
from pathlib import Path

import pytest

sys.path.insert(0, str(Path(__file__).resolve().parents[1]))
bench = pytest.importorskip("bench")


def test_stage_bytes_follow_the_reference_types():
    P, C, Np, N, R, w, h = 100, 30, 5, 1000, 700, 64, 32
    algo = bench.stage_bytes(P, C, Np, N, R, w, h)
    assert algo["transform"] == 20 * P + 4 * Np
    assert algo["count"] == 72 * C
    assert algo["sort"] == (16 + 12 + 8) * N          # sort, winding gather, winding scan
    assert algo["mark_emit"] == (20 + 16 + 24) * N + 16 * R
    assert algo["composite"] == 16 * R + 4 * w * h
    assert algo["scans+sync"] == 8 * (C + 1)            # the scan of the per-curve counts
    assert algo["chain_prep"] == (72 + 4 + 5 * 24 + 10 + 2 * 10) * C
    io = bench.stage_io_bytes(P, C, Np, N, R, w, h)
    assert io["scans+sync"] == 8 * (C + 1) + 8 * (Np + 1)
    assert set(algo) == set(io)                         # every stage rated both ways


def test_fused_mark_emit_is_one_sort_stage():
    N, R = 1000, 700
    algo = bench.stage_bytes(1, 1, 1, N, R, 8, 8)
    io = bench.stage_io_bytes(1, 1, 1, N, R, 8, 8)
    stages = {"sort": 1.25, "mark_emit": 0.005}
    bench.fold_mark_emit(stages, algo, io, N, R)
    assert stages["sort"] == pytest.approx(1.255)
    assert "mark_emit" not in algo and "mark_emit" not in io
    assert algo["sort"] == 96 * N + 16 * R                # every reference stage the kernel replaces
    assert io["sort"] == 12 * N + 16 * R


def test_every_table_row_has_algorithmic_and_own_bytes():
    """The table bench.py prints: every row that a production frame times is rated on its
    algorithmic bytes and on its own I/O (no null rates), and the stage times add up."""
    sizes = (1_000, 250, 40, 20_000, 15_000, 256, 256)
    ms = [0.01 * (k + 1) for k in range(len(bench.STAGE_NAMES))]
    stages, algo, io = bench.fold_stages(ms, sizes)
    assert set(stages) == set(algo) == set(io)
    assert "gen_fragment" not in stages and "bucket_scatter" not in stages
    dropped = bench.STAGE_NAMES.index("gen_fragment")
    assert sum(stages.values()) == pytest.approx(sum(ms) - ms[dropped])
    table = bench.rate_stages(stages, algo, io, None)
    for name, row in table.items():
        assert row["algo_GB/s"] is not None and row["io_GB/s"] is not None, name
