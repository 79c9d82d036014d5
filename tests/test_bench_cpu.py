"""bench.py host logic (no GPU): the one-GPU contract, the profiled-traffic lookup against the
committed rocprofv3 summaries, and the CPU-share rule of the baseline leg."""
from __future__ import annotations

import importlib.util
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_profiled_traffic_reads_the_committed_summaries():
    b = _bench()
    # the summaries the C4 line and the 1025x513 single-case line look up
    inner = b.profiled_traffic("mix_march_inner", "c4_batch64_rows37")
    last = b.profiled_traffic("mix_march_last", "c4_batch64_rows37")
    single = b.profiled_traffic("mix_march", "naca0012_1025x513")
    assert inner is not None and last is not None and single is not None
    # per launch between the design bytes and twice the fixed algorithmic figure
    cells = 3_805_184
    assert 88 * cells * 0.95 <= inner <= 152 * cells * 2
    assert 156 * cells * 0.95 <= last <= 2 * 156 * cells
    assert b.profiled_traffic("mix_march_inner", "no_such_workload") is None


def test_cpu_share_follows_omp_num_threads(monkeypatch):
    b = _bench()
    monkeypatch.setenv("OMP_NUM_THREADS", "7")
    assert b._cpu_share() == 7
    monkeypatch.setenv("OMP_NUM_THREADS", "not-a-number")
    assert b._cpu_share() >= 1
    monkeypatch.delenv("OMP_NUM_THREADS")
    assert b._cpu_share() >= 1


def test_design_bytes_are_below_the_fixed_figure():
    b = _bench()
    assert b.BYTES == {"fp64": 184.0, "mixed": 152.0}
    assert b.DESIGN_BYTES["mixed_inner"] < b.BYTES["mixed"] < b.DESIGN_BYTES["mixed_last"]
