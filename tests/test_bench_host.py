"""Host logic of bench.py (CPU only): the algorithmic bytes behind `roofline.achieved` and the
committed rocprofv3 PMC summaries behind `roofline.traffic`."""
import os
import re
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_gather_bytes_follow_the_reference_functions():
    """Indexing's stable scatter 60 B/particle and CollideParticles 16 B/cell on every step;
    Sampling adds 12 B/particle and 280 B/cell (the 17-double CELL read and written, and the
    count) on sampling steps. The fused kernel's own count keeps 13 sums (208 B/cell) and no
    velocity re-read."""
    n, m = 1000, 10
    assert bench.stage_bytes("gather", n, m, False) == 60 * n + 16 * m
    assert bench.stage_bytes("gather", n, m, True) == 72 * n + 296 * m
    assert bench.kernel_bytes("gather", n, m, False) == 60 * n + 16 * m
    assert bench.kernel_bytes("gather", n, m, True) == 60 * n + 224 * m
    assert bench.stage_bytes("move", n, m, True) == 28 * n + 8 * m
    assert bench.stage_bytes("scatter", n, m, False) == 8 * n + 8 * m
    for stage in ("move", "scatter", "scan"):
        assert bench.kernel_bytes(stage, n, m, True) == bench.stage_bytes(stage, n, m, True)


@pytest.mark.parametrize("workload", ["c2", "c3", "c4"])
def test_profiles_hold_traffic_for_every_step_kind(workload):
    """Every workload bench.py runs by name has PMC traffic lines for its streaming kernels on
    steps of both kinds, and the weighting interpolates between them."""
    stages = ["gather", "move"] + (["scatter"] if workload != "c2" else [])
    for stage in stages:
        plain, _ = bench.profiled_traffic(workload, stage, 0.0)
        samp, src = bench.profiled_traffic(workload, stage, 1.0)
        mixed, _ = bench.profiled_traffic(workload, stage, 0.6)
        assert plain and samp and src.startswith("profiles/"), (workload, stage)
        assert mixed == pytest.approx(0.4 * plain + 0.6 * samp)
        # measured traffic sits near the algorithmic bytes (at most about 1.5x of them)
        if stage == "gather" and workload != "c2":
            m = {"c3": 4.0e6, "c4": 18390635}[workload]
            n = {"c3": 8.0193523e7, "c4": 4.153192e8}[workload]
            assert 1.0 < plain / bench.kernel_bytes("gather", n, m, False) < 1.5


def test_each_traffic_line_is_stated_once():
    """One line per (workload, stage, kind) across the summaries, so the figure bench.py reads
    is the one the summary's table shows."""
    seen = {}
    d = os.path.join(ROOT, "profiles")
    for name in sorted(os.listdir(d)):
        if not name.endswith(".md"):
            continue
        for line in open(os.path.join(d, name)):
            f = line.split()
            if f and f[0] == "traffic_bytes_per_launch":
                key = tuple(f[1:3] + f[4:5])
                assert key not in seen, (key, name, seen.get(key))
                seen[key] = name
                assert re.fullmatch(r"\d+", f[3])
    assert len(seen) >= 14


def test_step_bytes_are_the_reference_step_functions():
    """A step of the reference's functions: 136 B/particle + 28 B/cell plain (move 56, key and
    histogram 20, stable scatter 60; collide 16 and scan 8 per cell, histogram 4), 148 and 308
    on sampling steps; 28 B per new particle and 48 B per collision."""
    n, m = 1000, 10
    assert bench.step_bytes(n, m, 0, 0, False) == 136 * n + 28 * m
    assert bench.step_bytes(n, m, 0, 0, True) == 148 * n + 308 * m
    assert bench.step_bytes(n, m, 5, 3, False) - bench.step_bytes(n, m, 0, 0, False) == 28 * 5 + 48 * 3
