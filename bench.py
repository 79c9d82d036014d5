#!/usr/bin/env python3
"""Throughput benchmark of the JST + SA residual-and-update loop on one MI355X.

A "step" is one explicit RK step (5 fused stage kernels + the step-end reduction kernel)
over the workload's whole mesh.  The metric is cell-stage updates/s = NI*NJ*5*steps / s,
with the state resident in HBM before the timed region starts.  One JSON line is printed.

    python bench.py [--steps K] [--warmup W] [--workload naca0012_257x129] [--no-cpu]

roofline: the dominant kernel is the fused stage kernel; its algorithmic bytes per launch
are 184 B/cell-stage (FP64: read U^(k-1) 40, read U^n 40, write U^(k) 40, geometry 56,
frozen V/dt 8) x NI*NJ, divided by its mean launch duration from HIP events recorded on the
solver's stream around every stage launch.  ``traffic`` comes from the rocprofv3 FETCH_SIZE /
WRITE_SIZE summary committed under profiles/ (calibrated by a copy kernel), if present.
cpu_baseline: the C restatement of jst-rs (the oracle, OpenMP row loops) timed on this host
on a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
BYTES_FP64 = 184.0


def load_workload(name: str) -> dict:
    from purinecfd_amd import workloads
    table = {"naca0012_257x129": workloads.c2_naca0012,
             "cylinder_40x17": workloads.c1_cylinder,
             "naca0012_1025x513": workloads.c3_naca_large}
    if name not in table:
        raise SystemExit(f"unknown workload {name}; choose from {sorted(table)}")
    return table[name]()


def profiled_traffic(kernel: str, workload: str):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, or None."""
    p = ROOT / "profiles" / f"{kernel}_{workload}_traffic.md"
    if not p.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch\s*[:=]\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


def cpu_baseline(w: dict, budget_s: float = 12.0) -> dict:
    from oracle.pyoracle import OracleSolver, lib
    s = OracleSolver(json.loads(w["config"]), w["mesh"])
    cells = s.ni * s.nj
    t0 = time.perf_counter()
    s.run(2)
    per = (time.perf_counter() - t0) / 2
    n = max(2, int(budget_s / max(per, 1e-6)))
    t0 = time.perf_counter()
    s.run(n)
    dt = time.perf_counter() - t0
    return {"value": cells * 5 * n / dt, "unit": "cell-stage updates/s",
            "cores": int(lib().jo_threads()), "kind": "port",
            "sample": f"{w['name']}: {n} steps of the C restatement of jst-rs (FP64, unfused "
                      f"kernels, OpenMP row loops) after 2 warm-up steps; {dt:.1f} s"}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--workload", default="naca0012_257x129")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; the path shards by case (or by "
              "circumferential slab with a 3-layer halo) but no multi-GPU run is part of it",
              file=sys.stderr)
        return 2
    from purinecfd_amd import jst

    w = load_workload(args.workload)
    s = jst.Solver(w["config"], w["mesh"], jst.FP64)
    cells = s.n_cells()
    if args.warmup > 0:
        s.time_steps(args.warmup)
    t0 = time.perf_counter()
    ev_ms, _ = s.time_steps(args.steps)
    wall = time.perf_counter() - t0
    value = cells * 5 * args.steps / wall
    # dominant kernel: mean stage-kernel duration from per-launch HIP events (separate pass)
    nprof = max(10, min(100, args.steps // 4))
    _, stage_ms = s.time_steps(nprof, per_stage=True)
    achieved = BYTES_FP64 * cells / (stage_ms * 1e-3) / 1e9
    traffic = profiled_traffic("stage_fp64", w["name"])
    out = {
        "metric": "cell-stage updates/s FP32-batched vs FP64-plain (ΔCL,ΔCD≤1e-4); p99 solve latency",
        "value": value,
        "unit": "cell-stage updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: generated O-mesh, freestream initial state",
        "config": {"workload": w["name"], "path": "fp64-plain", "cells": cells,
                   "event_ms_per_step": ev_ms / args.steps},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "stage_kernel<8,32,256> (FP64)",
                     "stage_kernel_ms": stage_ms,
                     "bytes_per_cell_stage": BYTES_FP64},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
