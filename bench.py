#!/usr/bin/env python3
"""Throughput benchmark of the JST + SA residual-and-update loop on one MI355X.

Default workload (the metric's "FP32-batched vs FP64-plain"): the 64-case uneven C4 batch
(NACA 4-digit O-meshes 129x65 .. 513x257 drawn from seed 20261203).  A "step" is one RK step
(5 stage launches + 1 step-end launch) of every case.  Tuning: the row-segment length
and grid of the stage kernel, picked by the autotuner.  Reported:

* value: FP32-mixed batched cell-stage updates/s = sum(NI*NJ)*5*steps / wall seconds, all
  64 cases resident in HBM and advanced by the same launches (HIP-graph replays), tile shape
  picked by the autotuner;
* fp64_plain: the FP64 plain path on the same 64 cases run one case at a time;
* accuracy: per case, mixed vs FP64 plain after the same number of steps from the same start:
  max |dCL|, |dCD| and field L2 (the FP64 plain path is pinned to the CPU oracle at 1e-10 per
  residual evaluation by the test suite);
* converged_accuracy: the same 64 geometries and flows with the settings that converge them
  (local time stepping at CFL 2 with the viscous spectral radius, the negative-nu~ SA
  branch), run to a 6-decade drop in one
  mixed batch and on the FP64 plain path one case at a time: outcome per case, |dCL|, |dCD|
  and field L2 where both converge, and the time to solution of the whole distribution;
* c2_convergence: the NACA 0012 257x129 case under its own stop rule (8-decade drop within
  20 000 steps) on both paths, and on the FP64 path with the step cap raised: the step at
  which each decade is reached, wall time, CL/CD;
* serving: p50/p99 time to completion, and to convergence over the converged requests, of a
  Poisson request stream (mean 20 solves/s) run by the serving loop (jst_serve);
* roofline: the dominant kernel, the FP32 row-march stage of RK stages 1-4 (4 of the 5
  launches per step): 152 algorithmic B/cell-stage x cells / its mean launch duration (HIP
  events around every launch on the batch's stream); stage 5 and the step end beside, and
  the bytes this design's launches move (88 B/cell inner, 156 stage 5);
* cpu_baseline: the oracle (C restatement of jst-rs, OpenMP) timed on this host's cores on a
  bounded sample (one 257x129 case of the batch).

    python bench.py [--steps K] [--warmup W] [--workload c4|naca0012_257x129|...] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import random
import re
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per cell-stage, the fixed per-unit figures of the roofline:
#   fp64 plain 184 = U^(k-1) 40 + U^n 40 + U^(k) 40 + FP64 geometry 56 + frozen V/dt 8
#   fp32-mixed 152 = U^(k-1) FP32 20 + U^n FP64 40 + U^(k) FP64 40 + FP32 shadow 20
#                    + FP32 geometry 28 + V/dt 4
# (a leaner storage scheme shows up as a higher rate, not as a smaller figure).  What this
# build's launches actually have to move, per cell, is reported beside it:
#   mixed RK stages 1-4: dU^(k-1) 20 + dU^n 20 + dU^(k) 20 + FP32 geometry 24 (tau and n
#                        normals, V, 1/d^2; 1/V is formed from V) + V/dt 4 = 88
#   mixed stage 5: dU^(4) 20 + U^n 40 + U^{n+1} 40 + dU^{n+1} 20 + geometry 24 + V/dt 4
#                  + next dt and V/dt 8 = 156
BYTES = {"fp64": 184.0, "mixed": 152.0}
DESIGN_BYTES = {"mixed_inner": 88.0, "mixed_last": 156.0}
METRIC = "cell-stage updates/s FP32-batched vs FP64-plain (ΔCL,ΔCD≤1e-4); p99 solve latency"


def profiled_traffic(kernel: str, workload: str):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, or None."""
    p = ROOT / "profiles" / f"{kernel}_{workload}_traffic.md"
    if not p.exists():
        return None
    m = re.search(r"traffic_bytes_per_launch\s*[:=]\s*([0-9.eE+]+)", p.read_text())
    return float(m.group(1)) if m else None


def _cpu_model() -> str:
    try:
        for ln in Path("/proc/cpuinfo").read_text().splitlines():
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_share() -> int:
    """Threads this process may use: OMP_NUM_THREADS when set (a GPU host grants each GPU a
    share of its cores and says so there), else every core of the machine.  The calling
    thread's affinity mask is no guide: the GPU runtime may have narrowed it."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    return int(env) if env.isdigit() and int(env) > 0 else (os.cpu_count() or 1)


def cpu_baseline(cfg: str, mesh: str, name: str, budget_s: float = 4.0) -> dict:
    """The oracle (C restatement of jst-rs: same unfused kernel sequence, OpenMP row loops)
    on this host's cores, timed like jst-rs bin/jst.rs:94-118 (the step loop only), at two
    thread counts: every core of this process's share, and the reference CLI's choice
    ceil(cells / 24576) (bin/jst.rs:52-56); OMP_PROC_BIND=close; median of 3 runs each."""
    os.environ.setdefault("OMP_PROC_BIND", "close")  # read when the OpenMP runtime starts
    from oracle.pyoracle import OracleSolver, lib
    L = lib()
    share = _cpu_share()
    cells = None
    runs = {}
    for label in ("all_cores", "reference_heuristic"):
        s = OracleSolver(json.loads(cfg), mesh)
        cells = s.ni * s.nj
        threads = share if label == "all_cores" else min(share, max(1, -(-cells // 24576)))
        L.jo_set_threads(threads)
        s.run(2)  # warm-up
        t0 = time.perf_counter()
        s.run(2)
        per = (time.perf_counter() - t0) / 2
        n = max(3, int(budget_s / 3 / max(per, 1e-6)))
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            s.run(n)
            times.append((time.perf_counter() - t0) / n)
        ms = sorted(times)[1] * 1e3
        runs[label] = {"threads": threads, "steps_per_run": n, "ms_per_step": ms,
                       "ns_per_cell_step": ms * 1e6 / cells,
                       "cell_stage_per_s": cells * 5 / (ms * 1e-3)}
    L.jo_set_threads(share)
    best = max(runs.values(), key=lambda r: r["cell_stage_per_s"])
    return {"value": best["cell_stage_per_s"], "unit": "cell-stage updates/s",
            "cores": best["threads"], "kind": "port", "cpu_model": _cpu_model(),
            "ms_per_step": best["ms_per_step"], "ns_per_cell_step": best["ns_per_cell_step"],
            "runs": runs, "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "host_cpus": os.cpu_count(), "main_thread_affinity": len(os.sched_getaffinity(0)),
            "sample": f"{name} ({cells} cells), one case of the batch: median of 3 runs of "
                      f"the C restatement of jst-rs (FP64, unfused kernels, OpenMP row loops) at "
                      f"{runs['all_cores']['threads']} threads (this process's cores) and "
                      f"{runs['reference_heuristic']['threads']} (jst-rs bin/jst.rs:52-56 "
                      f"heuristic), after 4 warm-up steps; value = the faster",
            "reference_published": "jst-rs bin/jst.rs:43-47, i7-14650HX, 32768 cells: 23.0 ms/step "
                                   "at 1 thread, 15.7 at 2 (best)"}


def fp64_plain_pass(jst, cases, acc_steps: int, time_steps: int):
    """FP64 plain path one case at a time: CL/CD after acc_steps, then timed steps."""
    out, cells_steps, secs = [], 0.0, 0.0
    for w in cases:
        s = jst.Solver(w["config"], w["mesh"], jst.FP64)
        s.run(acc_steps)
        f = s.forces()
        fields = {n: s.get(n) for n in ("rho", "u", "v", "p")}
        t0 = time.perf_counter()
        s.time_steps(time_steps)
        secs += time.perf_counter() - t0
        cells_steps += s.n_cells() * 5 * time_steps
        out.append((f, fields))
        s.close()
    return out, cells_steps / secs


def mixed_accuracy_pass(jst, cases, acc_steps: int):
    import numpy as np  # noqa: F401
    b = jst.Batch(hbm_bytes=8 << 30, max_cases=len(cases))
    ids = [b.add(w["config"], w["mesh"], acc_steps) for w in cases]
    b.run(steps_per_round=acc_steps)
    res = []
    for w, cid in zip(cases, ids):
        info = b.info(cid)
        assert info["step"] == acc_steps, info
        res.append((b.forces(cid), {n: b.get(cid, n, w["ni"], w["nj"]) for n in ("rho", "u", "v", "p")}))
    b.close()
    return res


def converged_accuracy_leg(jst, cases, max_steps: int = 20000):
    """The C4 distribution with the settings that converge it (workloads.converging_config:
    local time stepping at CFL 2 with the viscous spectral radius, the negative-nu~ SA branch,
    6-decade drop, at most max_steps): all 64 cases in one mixed
    batch, and the FP64 plain path one case at a time, each to its own convergence; per case
    the outcome on both paths and, where both converge, |dCL|, |dCD| and the field L2."""
    import numpy as np
    from purinecfd_amd import workloads
    cfgs = [workloads.converging_config(w, max_steps) for w in cases]
    b = jst.Batch(hbm_bytes=16 << 30, max_cases=len(cases))
    ids = [b.add(c, w["mesh"], max_steps) for c, w in zip(cfgs, cases)]
    t0 = time.perf_counter()
    b.run(steps_per_round=500)
    t_mixed = time.perf_counter() - t0
    names = {jst.CASE_CONVERGED: "converged", jst.CASE_STOPPED: "step cap",
             jst.CASE_NONPHYSICAL: "non-physical"}
    rows, both, t_plain, n_mixed, n_plain = [], [], 0.0, 0, 0
    for w, c, cid in zip(cases, cfgs, ids):
        im = b.info(cid)
        fm = b.forces(cid) if im["status"] != jst.CASE_NONPHYSICAL else None
        fields = {n: b.get(cid, n, w["ni"], w["nj"]) for n in ("rho", "u", "v", "p")}
        s = jst.Solver(c, w["mesh"], jst.FP64)
        t0 = time.perf_counter()
        try:
            rep = s.run(max_steps)
            st, steps = (jst.CASE_CONVERGED if rep.converged else jst.CASE_STOPPED), rep.steps
        except jst.NonPhysical:
            st, steps = jst.CASE_NONPHYSICAL, s.step
        t_plain += time.perf_counter() - t0
        row = [w["name"], names[im["status"]], im["step"], names[st], steps]
        n_mixed += im["status"] == jst.CASE_CONVERGED
        n_plain += st == jst.CASE_CONVERGED
        if im["status"] == jst.CASE_CONVERGED and st == jst.CASE_CONVERGED:
            f = s.forces()
            l2 = max(float(np.linalg.norm(fields[n] - s.get(n)) / np.linalg.norm(s.get(n)))
                     for n in fields)
            both.append((abs(fm["cl"] - f["cl"]), abs(fm["cd"] - f["cd"]), l2))
            row += both[-1]
        rows.append(row)
        s.close()
    b.close()
    return {"settings": "local time stepping with the viscous spectral radius, CFL 2, "
                        f"negative-nu~ SA branch, 6-decade drop, at most {max_steps} steps",
            "converged_mixed": n_mixed, "converged_fp64": n_plain, "converged_both": len(both),
            "max_abs_dcl": max(r[0] for r in both) if both else None,
            "max_abs_dcd": max(r[1] for r in both) if both else None,
            "max_field_l2_rel": max(r[2] for r in both) if both else None,
            "time_to_solution_s": {"mixed_batch_all_64": t_mixed, "fp64_one_at_a_time": t_plain},
            "reference": "FP64 plain GPU path (oracle-pinned), each case to its own convergence",
            "per_case_columns": ["case", "mixed", "mixed_steps", "fp64", "fp64_steps", "abs_dcl",
                                 "abs_dcd", "max_field_l2_rel"],
            "per_case": rows}


def c2_convergence_leg(jst, long_cap: int = 60000):
    """The NACA 0012 257x129 case (M 0.75, AoA 2 deg, Re 6e6; local time steps at CFL 2 with the
    negative-nu~ SA branch) under its own stop rule, an 8-decade drop within 20 000 steps, on
    both paths, and on the FP64 path again with the step cap raised to long_cap: the step at
    which each decade is first reached, the wall time of the run and CL/CD."""
    import json
    import numpy as np
    from purinecfd_amd import workloads
    w = workloads.c2_naca0012()
    runs = {}
    for key, prec, cap in (("fp64", jst.FP64, 0), ("mixed", jst.MIXED, 0),
                           ("fp64_cap_raised", jst.FP64, long_cap)):
        cfg = w["config"]
        if cap:
            c = json.loads(cfg)
            c["solver"]["iteration"] = cap
            cfg = json.dumps(c)
        s = jst.Solver(cfg, w["mesh"], prec)
        hist = []
        t0 = time.perf_counter()
        rep = s.run(on_step=lambda k, r: hist.append(r))
        wall = time.perf_counter() - t0
        h = np.array(hist)
        f = s.forces()
        runs[key] = {"steps": rep.steps, "converged": rep.converged,
                     "decades": float(np.log10(h[0] / h.min())),
                     "first_step_reaching": {str(d): int(np.argmax(h < h[0] * 10.0 ** -d)) + 1
                                             for d in range(1, 9) if (h < h[0] * 10.0 ** -d).any()},
                     "wall_s": wall, "cl": f["cl"], "cd": f["cd"]}
        s.close()
    return {"case": w["name"], "stop_rule": "8-decade drop of the density residual, at most 20 000 steps",
            "step_cap_raised_to": long_cap,
            "abs_dcl_mixed_vs_fp64": abs(runs["mixed"]["cl"] - runs["fp64"]["cl"]),
            "abs_dcd_mixed_vs_fp64": abs(runs["mixed"]["cd"] - runs["fp64"]["cd"]),
            **runs}


def serving_leg(jst, seconds: float, rate: float, seed: int):
    from purinecfd_amd import workloads
    reqs = workloads.c5_requests(seconds=seconds, rate=rate, seed=seed)
    res, summ = jst.serve([(r["arrival_s"], r["config"], r["mesh"], r["max_steps"]) for r in reqs],
                          hbm_bytes=16 << 30, max_cases=96, steps_per_round=50, tune=False)
    conv = sum(1 for r in res if r["status"] == jst.CASE_CONVERGED)
    summ.update({"requests": len(reqs), "converged": conv, "rate_per_s": rate,
                 "arrival_window_s": seconds,
                 "cell_stage_rate": summ["cell_stages"] / max(summ["wall_s"], 1e-9)})
    return summ


def bench_c4(args, jst):
    import numpy as np
    from purinecfd_amd import workloads
    cases = workloads.c4_batch()
    total = sum(w["ni"] * w["nj"] for w in cases)
    per_case = []
    if args.throughput_only:
        plain_rate, dcl, dcd, l2 = float("nan"), [float("nan")], [float("nan")], {}
    else:
        # accuracy: same start, same number of steps, mixed (batched) vs FP64 plain
        plain, plain_rate = fp64_plain_pass(jst, cases, args.acc_steps, args.plain_steps)
        mixed = mixed_accuracy_pass(jst, cases, args.acc_steps)
        dcl = [abs(m[0]["cl"] - p[0]["cl"]) for m, p in zip(mixed, plain)]
        dcd = [abs(m[0]["cd"] - p[0]["cd"]) for m, p in zip(mixed, plain)]
        rel = [{n: float(np.linalg.norm(m[1][n] - p[1][n]) / np.linalg.norm(p[1][n]))
                for n in ("rho", "u", "v", "p")} for m, p in zip(mixed, plain)]
        l2 = {n: max(r[n] for r in rel) for n in ("rho", "u", "v", "p")}
        per_case = [[w["name"], a, b, max(r.values())] for w, a, b, r in zip(cases, dcl, dcd, rel)]
    # throughput: every case resident, autotuned tiles, HIP-graph replayed steps
    def resident_batch():
        bb = jst.Batch(hbm_bytes=8 << 30, max_cases=len(cases))
        return bb, [bb.add(w["config"], w["mesh"], 10**9) for w in cases]

    b, ids = resident_batch()
    if args.tiling:
        variant, grid = (int(x) for x in args.tiling.split(","))
    else:
        # the tuner advances the cases by a few hundred steps; the timed run starts again
        # from the freestream so that every case stays well inside its stable window
        variant, grid = b.tune(6)
        b.close()
        b, ids = resident_batch()
    b.set_tiling(variant, grid)
    # the segment length of the chosen variant (variant 10: each case its own)
    lens = {b.segment_rows(cid) for cid in ids}
    rows = lens.pop() if len(lens) == 1 else "per-case"
    if args.warmup > 0:
        b.time_steps(args.warmup)
    # untimed: stops off, launch lists and the step graphs built and launched once; the timed
    # region is then exactly `steps` steps of all 64 cases and one wait for the stream
    b.time_begin(args.steps)
    t0 = time.perf_counter()
    b.step(args.steps, sync=True)
    wall = time.perf_counter() - t0
    b.time_end()
    value = total * 5 * args.steps / wall
    ev_steps = min(args.steps, 200)
    ev_ms, _ = b.time_steps(ev_steps)  # HIP events around as many steps (at most 200)
    kern = b.time_kernels(max(5, min(50, args.steps // 4)))
    b.close()
    achieved = BYTES["mixed"] * total / (kern["inner_ms"] * 1e-3) / 1e9
    last_gbs = BYTES["mixed"] * total / (kern["last_ms"] * 1e-3) / 1e9
    out = {
        "metric": METRIC, "value": value, "unit": "cell-stage updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: seeded NACA 4-digit O-meshes, freestream start",
        "config": {"workload": "C4 batch: 64 uneven NACA 4-digit O-mesh cases 129x65..513x257 "
                               "(seed 20261203), FP32-mixed batched, FP64 accumulation",
                   "cases": len(cases), "cells": total, "segment_rows": rows,
                   "grid": grid,
                   "event_ms_per_step": ev_ms / ev_steps},
        "fp64_plain": {"value": plain_rate, "unit": "cell-stage updates/s",
                       "how": f"same 64 cases, FP64 plain path, one case at a time, "
                              f"{args.plain_steps} timed steps each"},
        "speedup_vs_fp64_plain": value / plain_rate,
        "accuracy": {"steps": args.acc_steps, "max_abs_dcl": max(dcl), "max_abs_dcd": max(dcd),
                     "max_field_l2_rel": l2, "cases": len(cases),
                     "reference": "FP64 plain GPU path (oracle-pinned) at the same step",
                     "per_case_columns": ["case", "abs_dcl", "abs_dcd", "max_field_l2_rel"],
                     "per_case": per_case},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "whole_step_frac": value * BYTES["mixed"] / (HBM_PEAK_GBS * 1e9),
                     "traffic": profiled_traffic("mix_march_inner", f"c4_batch64_rows{rows}"),
                     "kernel": "mix_march_kernel<MIX_INNER> (RK stages 1-4, all 64 cases per launch)",
                     "kernel_ms": kern["inner_ms"], "bytes_per_cell_stage": BYTES["mixed"],
                     "design_bytes_per_cell": DESIGN_BYTES["mixed_inner"],
                     "stage5": {"kernel_ms": kern["last_ms"], "achieved_gbs": last_gbs,
                                "design_bytes_per_cell": DESIGN_BYTES["mixed_last"],
                                "traffic": profiled_traffic("mix_march_last", f"c4_batch64_rows{rows}")},
                     "step_end_ms": kern["end_ms"]},
    }
    if args.throughput_only:
        return out
    if not args.no_converged:
        out["converged_accuracy"] = converged_accuracy_leg(jst, cases)
        out["c2_convergence"] = c2_convergence_leg(jst)
    if not args.no_serve:
        out["serving"] = serving_leg(jst, args.serve_seconds, 20.0, 20261204)
    if not args.no_cpu:
        w = next(w for w in cases if w["ni"] == 128)
        out["cpu_baseline"] = cpu_baseline(w["config"], w["mesh"], w["name"])
    return out


def bench_single(args, jst):
    from purinecfd_amd import workloads
    table = {"naca0012_257x129": workloads.c2_naca0012, "cylinder_40x17": workloads.c1_cylinder,
             "naca0012_1025x513": workloads.c3_naca_large}
    w = table[args.workload]()
    prec = jst.MIXED if args.mixed else jst.FP64
    s = jst.Solver(w["config"], w["mesh"], prec)
    cells = s.n_cells()
    if args.warmup > 0:
        s.time_steps(args.warmup)
    s.time_begin(args.steps)  # untimed set-up, as for the batch
    t0 = time.perf_counter()
    s.time_run(args.steps)
    wall = time.perf_counter() - t0
    s.time_end()
    ev_ms, _ = s.time_steps(args.steps)
    _, stage_ms = s.time_steps(max(10, min(100, args.steps // 4)), per_stage=True)
    nbytes = BYTES["mixed"] if args.mixed else BYTES["fp64"]
    achieved = nbytes * cells / (stage_ms * 1e-3) / 1e9
    out = {"metric": METRIC, "value": cells * 5 * args.steps / wall, "unit": "cell-stage updates/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32" if args.mixed else "f64",
           "data": "synthetic: generated O-mesh, freestream start",
           "config": {"workload": w["name"], "path": "fp32-mixed" if args.mixed else "fp64-plain",
                      "cells": cells, "event_ms_per_step": ev_ms / args.steps},
           "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": achieved / HBM_PEAK_GBS,
                        "traffic": profiled_traffic("mix_march" if args.mixed else "stage_fp64", w["name"]),
                        "stage_kernel_ms": stage_ms, "bytes_per_cell_stage": nbytes}}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w["config"], w["mesh"], w["name"])
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="c4")
    ap.add_argument("--mixed", action="store_true", help="single-workload mode: FP32-mixed path")
    ap.add_argument("--acc-steps", type=int, default=100)
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--serve-seconds", type=float, default=10.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-converged", action="store_true",
                    help="C4: skip the runs to convergence (every C4 case, and the 257x129 case) on both paths")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--throughput-only", action="store_true",
                    help="C4: only the batched throughput pass (for profiler runs)")
    ap.add_argument("--tiling", default="",
                    help="C4: 'variant,grid' of the stage kernel instead of autotuning")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; the path shards by case (or by "
              "circumferential slab with a 3-layer halo) but no multi-GPU run is part of it",
              file=sys.stderr)
        return 2
    from purinecfd_amd import jst
    out = bench_c4(args, jst) if args.workload == "c4" else bench_single(args, jst)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
