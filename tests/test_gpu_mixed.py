"""FP32-mixed path, batching and serving on the GPU.

Tolerances: against the FP64 oracle, |dCL|, |dCD| <= 1e-4 per case, field L2 relative error
reported and bounded; within the mixed path itself, results are bitwise identical alone vs in
a mixed batch and for every launch grid (each face and cell is evaluated from the case's own
data only, reductions have a fixed per-case tree), and agree to FP32 rounding across row
segment lengths.
"""
from __future__ import annotations

import json

import numpy as np
import pytest

from conftest import config_dict, config_text, ellipse_mesh, mesh_text

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def jst():
    from purinecfd_amd import jst as mod
    mod.lib()
    return mod


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_mixed_residual_matches_oracle_fp32_level(jst, oracle_lib):
    """One residual evaluation (R = Fc - Fv - Fd per component, and S) of the mixed path
    after 3 steps of the golden 17x40 case, against the oracle: field-normed <= 2e-4."""
    text = mesh_text("cyl_17x40.txt")
    s = jst.Solver(config_text(), text, jst.MIXED)
    o = oracle_lib.OracleSolver(config_dict(), text)
    s.run(3)
    o.run(3)
    dt = s.residual_terms()
    odt = o.timestep()
    o.residual_terms()
    assert abs(dt - odt) <= 1e-5 * odt
    worst = 0.0
    for k in range(5):
        want = o.get(f"Fc{k}") - o.get(f"Fv{k}") - o.get(f"Fd{k}")
        got = s.get(f"R{k}")
        err = np.max(np.abs(got - want)) / np.max(np.abs(want))
        worst = max(worst, err)
        assert err <= 2e-4, (k, err)
    want = o.get("S")
    err = np.max(np.abs(s.get("S") - want)) / np.max(np.abs(want))
    assert err <= 2e-4, ("S", err)
    print(f"mixed residual worst field-normed error vs oracle: {max(worst, err):.2e}")


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_residual_terms_export_leaves_the_run_unchanged(jst, precision):
    """Solver::residual_terms is read-only (solver.rs:86-94): a run with an export in the
    middle ends bitwise where the same run without it does."""
    prec = jst.FP64 if precision == "fp64" else jst.MIXED
    text = mesh_text("cyl_17x40.txt")
    a = jst.Solver(config_text(), text, prec)
    b = jst.Solver(config_text(), text, prec)
    a.run(10)
    a.residual_terms()
    a.run(10)
    b.run(20)
    for n in ("rho", "u", "v", "p", "nut"):
        assert np.array_equal(a.get(n), b.get(n)), n
    assert a.residual() == b.residual()


def test_mixed_cylinder_c1_cl_cd_within_1e4(jst, oracle_lib):
    """Cylinder case (40 x 17 points, M 0.2, Re 1e3, CFL 2), 2000 iterations: mixed vs FP64
    oracle |dCL|,|dCD| <= 1e-4."""
    from purinecfd_amd import workloads
    w = workloads.c1_cylinder()
    s = jst.Solver(w["config"], w["mesh"], jst.MIXED)
    o = oracle_lib.OracleSolver(json.loads(w["config"]), w["mesh"])
    rep = s.run(2000)
    orep = o.run(2000)
    assert rep.steps == orep["steps"] == 2000
    f, of = s.forces(), o.forces()
    dcl, dcd = abs(f["cl"] - of["cl"]), abs(f["cd"] - of["cd"])
    l2 = {n: rel_l2(s.get(n), o.get(n)) for n in ("rho", "p", "u", "v", "nut")}
    print(f"C1 mixed vs oracle: dCL {dcl:.2e} dCD {dcd:.2e} (CD {of['cd']:.6f}); L2 {l2}")
    assert dcl <= 1e-4 and dcd <= 1e-4
    assert l2["rho"] <= 1e-5 and l2["p"] <= 1e-5 and l2["u"] <= 1e-4


def _cases():
    cfg_a = config_dict()
    cfg_b = config_dict()
    cfg_b["simulation"].update({"AOA": 3.0, "Ma": 0.3, "CFL": 1.0})
    cfg_c = config_dict()
    cfg_c["simulation"].update({"AOA": -2.0, "Ma": 0.5})
    return [
        (json.dumps(cfg_a), mesh_text("cyl_17x40.txt")),
        (json.dumps(cfg_b), ellipse_mesh(33, 96, 1.08)),
        (json.dumps(cfg_c), ellipse_mesh(21, 64, 1.15)),
    ]


def _fields(b, cid, text):
    from purinecfd_amd.jst import Mesh
    mesh = Mesh.parse(text)
    ni, nj = mesh.ni(), mesh.nj()
    return {n: b.get(cid, n, ni, nj) for n in ("rho", "u", "v", "p", "nut")}


def test_batch_alone_equals_in_the_mix_bitwise(jst):
    cases = _cases()
    alone = []
    for cfg, text in cases:
        b = jst.Batch(1 << 30, 4)
        cid = b.add(cfg, text, 60)
        b.run(steps_per_round=25)
        alone.append((_fields(b, cid, text), b.forces(cid), b.info(cid)))
        b.close()
    b = jst.Batch(1 << 30, 8)
    ids = [b.add(cfg, text, 60) for cfg, text in cases]
    b.run(steps_per_round=25)
    for (cfg, text), cid, (fa, fo, ia) in zip(cases, ids, alone):
        fm = _fields(b, cid, text)
        for n in fa:
            assert np.array_equal(fa[n], fm[n]), n
        assert b.forces(cid) == fo
        assert b.info(cid)["step"] == ia["step"] == 60


def test_segment_length_and_grid(jst):
    """The launch grid never changes results (bitwise).  The row-segment length (the tuning
    variant) only moves where a segment's edge rows are evaluated: results agree to FP32
    rounding (relative 1e-6 after 40 steps), and bitwise for a fixed segment length."""
    cfg, text = _cases()[1]
    runs = {}
    for variant, grid in ((0, 0), (0, 3), (1, 0), (1, 16), (2, 0), (3, 0), (4, 0), (5, 0), (5, 1024), (6, 0), (6, 5),
                          (7, 0), (7, 7), (8, 0), (8, 3), (9, 0), (9, 5), (10, 0), (10, 3), (11, 0),
                          (11, 5), (137, 0), (137, 3)):
        b = jst.Batch(1 << 30, 2)
        b.set_tiling(variant, grid)
        cid = b.add(cfg, text, 40)
        b.run(steps_per_round=20)
        runs[(variant, grid)] = _fields(b, cid, text)
        b.close()
    for (variant, grid), f in runs.items():
        ref = runs[(variant, 0)]
        for n in f:
            assert np.array_equal(f[n], ref[n]), (variant, grid, n)
        base = runs[(0, 0)]
        for n in f:
            assert rel_l2(f[n], base[n]) <= 1e-6, (variant, n, rel_l2(f[n], base[n]))


def test_batch_fitted_segment_length(jst):
    """Variant 11 gives every case of the batch one segment length: the one minimising
    ceil(items / 2048) x (L + 6), so the shortest whose work items fit the 2048 resident waves
    in one round where one does.  Known answers: a lone 1025 x 513 case gets its own length
    (5 rows: 18 column groups x 103 segments = 1854 items; 4 rows would give 2304); thirty 128 x 256-cell cases get 10
    rows (5 groups x 13 segments x 30 = 1950 items; 9 rows would give 2250, a second round);
    the 64-case NACA batch gets 37 rows (2019 items; 36 rows give 2172).  The length follows
    the batch as cases come and go, and a refit moves results only at FP32 rounding."""
    from purinecfd_amd import workloads
    b = jst.Batch(1 << 30, 1)
    b.set_tiling(11, 0)
    w = workloads.c3_naca_large()
    cid = b.add(w["config"], w["mesh"], 10)
    assert b.segment_rows(cid) == 5
    b.set_tiling(10, 0)
    assert b.segment_rows(cid) == 5
    b.close()

    cfg, _ = _cases()[1]
    text = ellipse_mesh(129, 256, 1.04)  # 128 x 256 cells
    b = jst.Batch(4 << 30, 32)
    b.set_tiling(11, 0)
    ids = [b.add(cfg, text, 20) for _ in range(30)]
    assert {b.segment_rows(c) for c in ids} == {10}
    b.run(steps_per_round=10)
    one = jst.Batch(1 << 30, 1)
    one.set_tiling(11, 0)
    c1 = one.add(cfg, text, 20)
    assert one.segment_rows(c1) == 1  # alone: 5 groups x 128 one-row segments = 640 items
    one.run(steps_per_round=10)
    f1 = one.forces(c1)
    one.close()
    for c in ids:
        f = b.forces(c)
        assert b.info(c)["step"] == 20
        assert abs(f["cl"] - f1["cl"]) <= 1e-5 and abs(f["cd"] - f1["cd"]) <= 1e-5
    for c in ids[1:]:
        b.remove(c)
    assert b.segment_rows(ids[0]) == 1  # refitted to the one case left
    with pytest.raises(jst.JstError):
        b.segment_rows(ids[1])  # removed
    with pytest.raises(jst.JstError):
        b.segment_rows(999)
    b.close()

    cases = workloads.c4_batch()
    b = jst.Batch(8 << 30, 64)
    b.set_tiling(11, 512)
    ids = [b.add(w["config"], w["mesh"], 10) for w in cases]
    assert {b.segment_rows(c) for c in ids} == {37}
    b.close()


def test_graph_replays_match_eager_launches(jst):
    """A HIP-graph replayed step sequence checked replay by replay against the per-step
    residual history of a fresh batch stepped one synchronised step at a time."""
    cfg, text = _cases()[2]
    b1 = jst.Batch(1 << 30, 2)
    c1 = b1.add(cfg, text, 30)
    for _ in range(30):
        b1.step(1)
    h1 = b1.history(c1, 1, 30)
    b2 = jst.Batch(1 << 30, 2)
    c2 = b2.add(cfg, text, 30)
    for n in (16, 10, 4):  # a 16-step graph, then single-step graphs
        b2.step(n)
    h2 = b2.history(c2, 1, 30)
    assert np.array_equal(h1, h2)


def test_diverging_case_is_evicted_others_unaffected(jst):
    cases = _cases()[:2]
    bad = config_dict()
    bad["simulation"]["CFL"] = 200.0
    b = jst.Batch(1 << 30, 4)
    ids = [b.add(cfg, text, 50) for cfg, text in cases]
    bid = b.add(json.dumps(bad), ellipse_mesh(13, 40), 50)
    b.run(steps_per_round=10)
    assert b.info(bid)["status"] == jst.CASE_NONPHYSICAL
    b_ok = jst.Batch(1 << 30, 4)
    ids_ok = [b_ok.add(cfg, text, 50) for cfg, text in cases]
    b_ok.run(steps_per_round=10)
    for (cfg, text), a, c in zip(cases, ids, ids_ok):
        assert b.info(a)["status"] == jst.CASE_STOPPED
        fa, fc = _fields(b, a, text), _fields(b_ok, c, text)
        for n in fa:
            assert np.array_equal(fa[n], fc[n]), n


def test_serving_loop_results_match_alone(jst):
    """Requests arriving while others run are admitted into the running batch; each result
    equals the same case run alone (bitwise), and latency percentiles are reported."""
    cases = _cases()
    reqs = [(0.0, cases[0][0], cases[0][1], 80), (0.05, cases[1][0], cases[1][1], 60),
            (0.1, cases[2][0], cases[2][1], 70), (0.12, cases[0][0], cases[0][1], 40)]
    res, summ = jst.serve(reqs, hbm_bytes=1 << 30, max_cases=3, steps_per_round=10,
                          segment_rows=-1)  # own lengths at any load: bitwise vs alone
    assert summ["completed"] == 4 and summ["nonphysical"] == 0 and summ["rejected"] == 0
    assert summ["p99_s"] >= summ["p50_s"] > 0
    assert summ["converged"] == 0 and np.isnan(summ["p50_converged_s"])  # all step-capped
    for (t, cfg, text, steps), r in zip(reqs, res):
        b = jst.Batch(1 << 30, 1)
        b.set_tiling(10, 0)  # the serving loop's default: each case its own segment length
        cid = b.add(cfg, text, steps)
        b.run(steps_per_round=10)
        f = b.forces(cid)
        assert r["steps"] == steps and r["status"] == jst.CASE_STOPPED
        assert r["cl"] == f["cl"] and r["cd"] == f["cd"]
        b.close()


def test_admission_by_footprint(jst):
    """A batch whose arena holds one case admits the second only after the first retires."""
    cfg, text = _cases()[1]
    fp = jst.Batch.footprint(cfg, text)
    b = jst.Batch(int(fp * 1.5), 4)
    a = b.add(cfg, text, 10)
    with pytest.raises(jst.JstError) as e:
        b.add(cfg, text, 10)
    assert e.value.status == jst.JST_ERR_NOSPACE
    b.run(steps_per_round=10)
    b.remove(a)
    b.add(cfg, text, 10)


def test_serving_requests_wait_for_memory(jst):
    """An arena that holds one case at a time: requests that have arrived queue until the
    running case retires, are admitted in arrival order, and still finish with the result
    they give alone."""
    cfg, text = _cases()[1]
    fp = jst.Batch.footprint(cfg, text)
    reqs = [(0.0, cfg, text, 30), (0.0, cfg, text, 20), (0.01, cfg, text, 25)]
    res, summ = jst.serve(reqs, hbm_bytes=int(fp * 1.5), max_cases=4, steps_per_round=10,
                          segment_rows=-1)
    assert summ["completed"] == 3 and summ["max_concurrent"] == 1
    for q in (1, 2):
        assert res[q]["admit_s"] >= res[q - 1]["done_s"]
    for (t, c, m, steps), r in zip(reqs, res):
        assert r["steps"] == steps and r["status"] == jst.CASE_STOPPED
    b = jst.Batch(1 << 30, 1)
    b.set_tiling(10, 0)
    cid = b.add(cfg, text, 20)
    b.run(steps_per_round=10)
    f = b.forces(cid)
    assert res[1]["cl"] == f["cl"] and res[1]["cd"] == f["cd"]
    b.close()


def test_tuner_choice_is_valid_and_its_steps_are_real_work(jst):
    """jst_batch_tune times every (segment length, grid) candidate on the live batch: it
    returns one of them, leaves every case advanced by the same whole number of steps, and
    that state matches a fixed-tiling run of as many steps to FP32 rounding."""
    cases = _cases()
    b = jst.Batch(1 << 30, 4)
    ids = [b.add(cfg, text, 10 ** 6) for cfg, text in cases]
    nvar = 12  # row-segment lengths 16 24 32 40 48 64 8 4 2 1, each case its own, batch-fitted
    b.set_tiling(nvar - 1, 0)
    with pytest.raises(jst.JstError):
        b.set_tiling(nvar, 0)
    v, g = b.tune(1)
    assert 0 <= v < nvar and g in (0, 512, 1024)
    steps = {b.info(cid)["step"] for cid in ids}
    assert len(steps) == 1
    n = steps.pop()
    assert n >= 2 * 3 * nvar  # two rounds x three grids x every variant, >= 1 step each
    ref = jst.Batch(1 << 30, 4)
    rids = [ref.add(cfg, text, n) for cfg, text in cases]
    ref.run(steps_per_round=50)
    for (cfg, text), a, r in zip(cases, ids, rids):
        assert b.info(a)["status"] == jst.CASE_RUNNING and ref.info(r)["step"] == n
        fa, fr = _fields(b, a, text), _fields(ref, r, text)
        for name in fa:
            assert rel_l2(fa[name], fr[name]) <= 1e-5, (name, rel_l2(fa[name], fr[name]))
    b.close()
    ref.close()


def test_serving_with_grid_tuning_matches_alone(jst):
    """With tuning on, the serving loop re-tunes the launch grid once four or more cases run;
    the grid never changes results, so every request still ends bitwise where it ends alone."""
    cases = _cases()
    reqs = [(0.0, cfg, text, 30 + 5 * q) for q, (cfg, text) in enumerate(cases + cases[:2])]
    res, summ = jst.serve(reqs, hbm_bytes=1 << 30, max_cases=8, steps_per_round=10, tune=True,
                          segment_rows=-1)
    assert summ["completed"] == len(reqs) and summ["max_concurrent"] >= 4
    for (t, cfg, text, steps), r in zip(reqs, res):
        b = jst.Batch(1 << 30, 1)
        b.set_tiling(10, 0)
        cid = b.add(cfg, text, steps)
        b.run(steps_per_round=10)
        f = b.forces(cid)
        assert r["steps"] == steps and r["cl"] == f["cl"] and r["cd"] == f["cd"]
        b.close()


def test_negative_nut_under_the_reference_sa_tracks_fp64(jst):
    """C4 case 8 (129x65) with local time stepping at CFL 2 and the reference's SA model: nu~
    turns negative next to the wall from step 195, and from step 241 a few wall cells have a
    clamped S~ and r near -2e10, where the FP32 fw would overflow (the mixed path takes the
    FP64 formula there).  On the FP64 path's step-250 state one residual evaluation of the
    mixed path is within 2e-4 field-normed of the FP64 one (R and S); over the 250 steps the
    mixed run keeps |dCL|, |dCD| <= 1e-4 and the L2 norms of density, pressure and velocity
    within 2e-4 (bursts at a wall cell near the leading edge move local values by more)."""
    from purinecfd_amd import workloads
    w = workloads.c4_batch(n=9)[8]
    cfg = json.loads(w["config"])
    cfg["simulation"].update({"CFL": 2.0, "local_time_stepping": True})
    cfg["spalart_allmaras"]["negative"] = False
    s64 = jst.Solver(json.dumps(cfg), w["mesh"], jst.FP64)
    smx = jst.Solver(json.dumps(cfg), w["mesh"], jst.MIXED)
    assert s64.run(250).steps == 250
    assert smx.run(250).steps == 250
    assert s64.get("nut").min() < 0.0
    f, g = smx.forces(), s64.forces()
    errs = {"cl": abs(f["cl"] - g["cl"]), "cd": abs(f["cd"] - g["cd"])}
    for n in ("rho", "p"):
        errs[n] = rel_l2(smx.get(n), s64.get(n))
    dv = np.hypot(smx.get("u") - s64.get("u"), smx.get("v") - s64.get("v"))
    errs["velocity"] = float(np.linalg.norm(dv) / np.linalg.norm(np.hypot(s64.get("u"), s64.get("v"))))
    u250 = np.stack([s64.get(f"U{k}") for k in range(5)])
    smx.set_state(u250)
    smx.residual_terms()
    s64.residual_terms()
    for k in range(5):
        want = s64.get(f"Fc{k}") - s64.get(f"Fv{k}") - s64.get(f"Fd{k}")
        errs[f"R{k}"] = float(np.max(np.abs(smx.get(f"R{k}") - want)) / np.max(np.abs(want)))
    want = s64.get("S")
    errs["S"] = float(np.max(np.abs(smx.get("S") - want)) / np.max(np.abs(want)))
    print("mixed vs fp64, reference SA through negative nu~:", {k: f"{v:.1e}" for k, v in errs.items()})
    assert errs["cl"] <= 1e-4 and errs["cd"] <= 1e-4, errs
    for n in ("rho", "p", "velocity"):
        assert errs[n] <= 2e-4, (n, errs[n])
    for n in ("R0", "R1", "R2", "R3", "R4", "S"):
        assert errs[n] <= 2e-4, (n, errs[n])


def test_serving_under_load_switches_tiling_within_tolerance(jst):
    """Thirty requests arriving at once saturate the GPU at the cases' own (short) segment
    lengths, so the adaptive serving loop (segment_rows 0, the default) runs them with
    24-row segments; each result then differs from the same case alone at most at FP32
    rounding (|dCL|, |dCD| <= 1e-5 here), inside the |dCL|, |dCD| <= 1e-4 contract."""
    cfg_b = _cases()[1][0]
    text = ellipse_mesh(129, 256, 1.04)  # 128 x 256 cells: 640 one-row work items each
    reqs = [(0.0, cfg_b, text, 20)] * 30
    res, summ = jst.serve(reqs, hbm_bytes=4 << 30, max_cases=32, steps_per_round=10)
    assert summ["completed"] == 30 and summ["nonphysical"] == 0
    assert summ["tiling_switches"] >= 1
    b = jst.Batch(1 << 30, 1)
    b.set_tiling(10, 0)
    cid = b.add(cfg_b, text, 20)
    b.run(steps_per_round=10)
    f = b.forces(cid)
    b.close()
    for r in res:
        assert r["steps"] == 20
        assert abs(r["cl"] - f["cl"]) <= 1e-5 and abs(r["cd"] - f["cd"]) <= 1e-5
    # a light stream keeps each case's own segment length: no switch
    res, summ = jst.serve(reqs[:2], hbm_bytes=4 << 30, max_cases=32, steps_per_round=10)
    assert summ["tiling_switches"] == 0


def test_serving_rejects_a_request_larger_than_the_arena(jst):
    """A request whose footprint exceeds the whole arena can never be admitted: it is
    rejected (status -1) once nothing else is resident, and the other requests, before and
    after it, still run to the result they give alone."""
    cfg, text = _cases()[1]
    big = ellipse_mesh(257, 512, 1.02)
    fp = jst.Batch.footprint(cfg, text)
    assert jst.Batch.footprint(cfg, big) > 2 * fp
    reqs = [(0.0, cfg, text, 20), (0.0, cfg, big, 20), (0.01, cfg, text, 25)]
    res, summ = jst.serve(reqs, hbm_bytes=int(fp * 1.5), max_cases=4, steps_per_round=10,
                          segment_rows=-1)
    assert res[1]["status"] == -1 and summ["rejected"] == 1
    assert summ["completed"] == 2
    for q, steps in ((0, 20), (2, 25)):
        assert res[q]["status"] == jst.CASE_STOPPED and res[q]["steps"] == steps
