"""GPU FP64 plain path against the reference golden vectors and the CPU oracle.

Mirrors jst-rs tests/golden.rs on the GPU through the C ABI, then checks the GPU against the
oracle on a 40 x 17 cylinder case (per-residual-evaluation error <= 1e-10, converged CL/CD
to 1e-8), on a 513 x 257 NACA case and on the reference's robustness properties
(tests/properties.rs).
"""
from __future__ import annotations

import json

import numpy as np
import pytest

from conftest import (GOLDEN_FILES, config_dict, config_text, ellipse_mesh, field_cmp,
                      load_golden, mesh_text)

pytestmark = pytest.mark.gpu

MESHES = {"fangdata.txt": "fangdata.txt", "cyl_17x40.txt": "cyl_17x40.txt"}


@pytest.fixture(scope="module")
def jst():
    from purinecfd_amd import jst as mod
    mod.lib()
    return mod


@pytest.fixture(params=GOLDEN_FILES, ids=lambda p: p.stem)
def golden(request, jst):
    doc = load_golden(request.param)
    return doc, mesh_text(MESHES[doc["meta"]["mesh"]])


def test_geometry_matches_golden_and_oracle_bitwise(golden, jst, oracle_lib):
    doc, text = golden
    s = jst.Solver(config_text(), text)
    o = oracle_lib.OracleSolver(config_dict(), text)
    g = doc["geometry"]
    for key, name in (("cell_vol", "vol"), ("cell_x", "cx"), ("cell_y", "cy"),
                      ("cell_sad", "wall_dist")):
        field_cmp(s.get(name), g[key], name, 1e-13, 1e-15)
    for key in ("tau_nx", "tau_ny", "tau_mx", "tau_my", "n_nx", "n_ny", "n_mx", "n_my"):
        field_cmp(s.get(key), g[key], key, 1e-13, 1e-15)
    # +,-,*,/ only (no FMA contraction on either side): bit-identical to the oracle
    for name in ("vol", "cx", "cy", "tau_nx", "tau_ny", "tau_mx", "tau_my", "n_nx", "n_ny",
                 "inv_vol"):
        assert np.array_equal(s.get(name), o.get(name)), name
    # sqrt is correctly rounded on both sides
    np.testing.assert_array_max_ulp(s.get("wall_dist"), o.get("wall_dist"), maxulp=1)


@pytest.mark.parametrize("shape", [(128, 33, 0.04), (256, 129, 0.0)])
def test_geometry_on_generated_naca_meshes_matches_oracle(jst, oracle_lib, shape):
    """FP64 geometry precompute on stretched NACA O-meshes (wall spacing 1e-5 chord, wall
    distance window min(max(15, NJ/5), NJ/2), geometry.rs:186-216): bit-identical to the
    oracle except the correctly rounded sqrt of the wall distance (1 ulp)."""
    from purinecfd_amd import meshgen
    n_theta, rings, m = shape
    text, _ = meshgen.naca_mesh(n_theta, rings, m, 0.4, 0.12, 1e-5, 30.0)
    s = jst.Solver(config_text(), text)
    o = oracle_lib.OracleSolver(config_dict(), text)
    for name in ("vol", "cx", "cy", "tau_nx", "tau_ny", "tau_mx", "tau_my", "n_nx", "n_ny",
                 "inv_vol"):
        assert np.array_equal(s.get(name), o.get(name)), name
    np.testing.assert_array_max_ulp(s.get("wall_dist"), o.get("wall_dist"), maxulp=1)


def test_initial_state_matches_golden(golden, jst):
    doc, text = golden
    s = jst.Solver(config_text(), text)
    for key, name in (("rho", "rho"), ("p", "p"), ("T", "T"), ("u", "u"), ("v", "v"),
                      ("E", "E"), ("H", "H"), ("c", "c"), ("miubl", "nut")):
        field_cmp(s.get(name), doc["init"][key], f"init.{key}", 1e-14, 1e-300)
    for k in range(5):
        field_cmp(s.get(f"U{k}"), doc["init"][f"U{k + 1}"], f"init.U{k + 1}", 1e-14, 1e-300)


GRADS = ["ugrad_x", "ugrad_y", "vgrad_x", "vgrad_y", "Tgrad_x", "Tgrad_y",
         "miublgrad_x", "miublgrad_y"]


def test_single_stage_residual_terms(golden, jst, oracle_lib):
    """golden.rs:221-275 on the GPU, plus the same terms against the oracle."""
    doc, text = golden
    s = jst.Solver(config_text(), text)
    s.run(doc["meta"]["warmup"])
    dt = s.residual_terms()
    t = doc["terms"]
    assert abs(dt - t["dt"]) < 1e-12 * t["dt"]
    o = oracle_lib.OracleSolver(config_dict(), text)
    o.run(doc["meta"]["warmup"])
    assert abs(dt - o.timestep()) <= 1e-13 * dt
    o.residual_terms()
    for k, key in enumerate(GRADS):
        field_cmp(s.get(f"grad{k}"), t[key], key, 1e-10, 1e-280)
        field_cmp(s.get(f"grad{k}"), o.get(f"grad{k}"), "oracle " + key, 1e-10, 1e-280)
    for key in ("chi", "fv1"):
        field_cmp(s.get(key), t[key], key, 1e-10, 1e-280)
    for k in range(5):
        for term in ("Fc", "Fv", "Fd"):
            field_cmp(s.get(f"{term}{k}"), t[f"{term}{k + 1}"], f"{term}{k + 1}", 1e-10, 1e-280)
            field_cmp(s.get(f"{term}{k}"), o.get(f"{term}{k}"), f"oracle {term}{k}", 1e-10, 1e-280)
    field_cmp(s.get("S"), t["S5"], "S5", 1e-10, 1e-280)
    field_cmp(s.get("S"), o.get("S"), "oracle S", 1e-10, 1e-280)


def test_final_state(golden, jst):
    """golden.rs:278-311 on the GPU."""
    doc, text = golden
    s = jst.Solver(config_text(), text)
    rep = s.run(doc["meta"]["steps"])
    f = doc["final"]
    assert rep.steps == doc["meta"]["steps"]
    assert abs(rep.residual - f["residual"]) < 1e-8 * f["residual"]
    assert abs(rep.totaltime - f["totaltime"]) < 1e-10 * f["totaltime"]
    for key, name in (("rho", "rho"), ("p", "p"), ("T", "T"), ("u", "u"), ("v", "v"),
                      ("miubl", "nut")):
        field_cmp(s.get(name), f[key], f"final.{key}", 1e-8, 1e-300)


def test_cylinder_c1_matches_oracle(jst, oracle_lib):
    """Cylinder case (40 x 17 points, M 0.2, Re 1e3, CFL 2, k4 1/64, nu~/nu 3):
    per-step residual history, every per-residual evaluation along the way and the
    converged CL/CD after 2000 iterations against the oracle."""
    from purinecfd_amd import workloads
    w = workloads.c1_cylinder()
    s = jst.Solver(w["config"], w["mesh"])
    o = oracle_lib.OracleSolver(json.loads(w["config"]), w["mesh"])
    hist = []
    for block in range(4):  # residual terms checked at four points of the run
        rep = s.run(500, on_step=lambda k, r: hist.append(r))
        orep = o.run(500)
        assert rep.steps == 500 and orep["steps"] == 500
        s.residual_terms()
        o.timestep()
        o.residual_terms()
        for term in ("Fc", "Fv", "Fd"):
            for k in range(5):
                field_cmp(s.get(f"{term}{k}"), o.get(f"{term}{k}"), f"{term}{k}@{block}", 1e-10, 1e-280)
        field_cmp(s.get("S"), o.get("S"), f"S@{block}", 1e-10, 1e-280)
        field_cmp(np.array(hist[-500:]), orep["history"], f"residual history @{block}", 1e-9)
    for name in ("rho", "p", "u", "v", "nut"):
        field_cmp(s.get(name), o.get(name), name, 1e-9, 1e-300)
    f, of = s.forces(), o.forces()
    assert abs(f["cd"] - of["cd"]) <= 1e-8 * max(1.0, abs(of["cd"]))
    assert abs(f["cl"] - of["cl"]) <= 1e-8
    assert 0.5 < f["cd"] < 2.0  # a bluff body at Re 1e3: drag coefficient of order one
    wg, wo = s.wall(), o.wall()
    for k in ("x", "y", "cp", "cf"):
        field_cmp(wg[k], wo[k], f"wall.{k}", 1e-9, 1e-300)


def test_run_is_deterministic(jst):
    """properties.rs:49-75 / solver.rs:318-326: bitwise reproducible runs."""
    text = ellipse_mesh(17, 48, 1.1)
    a = jst.Solver(config_text(), text)
    b = jst.Solver(config_text(), text)
    a.run(25)
    b.run(25)
    for name in ("rho", "u", "nut"):
        assert np.array_equal(a.get(name), b.get(name)), name
    assert a.totaltime == b.totaltime


def test_large_case_matches_oracle(jst, oracle_lib):
    """A 513 x 257-point NACA case of the uneven batch (131 072 cells, 512 stage tiles, more
    than one per CU, and a 512-tile step end): 40 steps, then one residual evaluation,
    against the oracle at 1e-10."""
    from purinecfd_amd import workloads
    w = next(c for c in workloads.c4_batch() if (c["ni"], c["nj"]) == (256, 512))
    s = jst.Solver(w["config"], w["mesh"])
    o = oracle_lib.OracleSolver(json.loads(w["config"]), w["mesh"])
    hist = []
    rep = s.run(40, on_step=lambda k, r: hist.append(r))
    orep = o.run(40)
    assert rep.steps == 40 and orep["steps"] == 40
    field_cmp(np.array(hist), orep["history"], "residual history", 1e-9)
    assert abs(rep.totaltime - o.totaltime) <= 1e-12 * o.totaltime
    s.residual_terms()
    o.timestep()
    o.residual_terms()
    for term in ("Fc", "Fv", "Fd"):
        for k in range(5):
            field_cmp(s.get(f"{term}{k}"), o.get(f"{term}{k}"), f"{term}{k}", 1e-10, 1e-280)
    field_cmp(s.get("S"), o.get("S"), "S", 1e-10, 1e-280)
    for name in ("rho", "p", "u", "v", "nut"):
        field_cmp(s.get(name), o.get(name), name, 1e-10, 1e-300)


def test_meshes_of_several_shapes_stay_physical(jst, oracle_lib):
    """properties.rs:158-172, with the oracle as the reference on each shape."""
    for rings, nj, stretch in ((8, 16, 1.0), (13, 40, 1.0), (21, 64, 1.25)):
        text = ellipse_mesh(rings, nj, stretch)
        s = jst.Solver(config_text(), text)
        o = oracle_lib.OracleSolver(config_dict(), text)
        rep = s.run(40)
        o.run(40)
        assert rep.steps == 40
        rho, p = s.get("rho"), s.get("p")
        assert (rho > 0).all() and (p > 0).all()
        field_cmp(rho, o.get("rho"), f"rho {rings}x{nj}", 1e-10)
        field_cmp(s.get("nut"), o.get("nut"), f"nut {rings}x{nj}", 1e-9)


@pytest.mark.parametrize("rings,nj,stretch", [(4, 8, 1.0), (4, 9, 1.3), (5, 61, 1.2)])
def test_smallest_meshes_match_oracle(jst, oracle_lib, rings, nj, stretch):
    """The smallest mesh Mesh::parse accepts (4 rings x 8 points, mesh.rs:159-165), an odd ring
    length, and a ring three points wider than one 58-column wave group of the mixed path:
    FP64 at 1e-10 of the oracle after 60 steps, FP32-mixed within 1e-5 of it."""
    text = ellipse_mesh(rings, nj, stretch)
    o = oracle_lib.OracleSolver(config_dict(), text)
    o.run(60)
    s64 = jst.Solver(config_text(), text, jst.FP64)
    smx = jst.Solver(config_text(), text, jst.MIXED)
    assert s64.run(60).steps == 60 and smx.run(60).steps == 60
    for name in ("rho", "u", "v", "p", "nut"):
        field_cmp(s64.get(name), o.get(name), f"fp64 {name}", 1e-10, 1e-300)
    for name in ("rho", "u", "v", "p"):
        field_cmp(smx.get(name), o.get(name), f"mixed {name}", 1e-5, 1e-300)


def test_converges_and_stops_on_the_exact_step(jst, oracle_lib):
    """properties.rs:178-201: the device stops at the same step as the reference loop."""
    text = ellipse_mesh(9, 32, 1.0)
    cfg = config_dict()
    cfg["solver"]["targetres"] = 1e-6
    s = jst.Solver(json.dumps(cfg), text)
    o = oracle_lib.OracleSolver(cfg, text)
    rep = s.run(6000)
    orep = o.run(6000)
    assert rep.converged and orep["converged"]
    assert rep.steps == orep["steps"]
    assert abs(rep.residual - orep["residual"]) <= 1e-8 * orep["residual"]


def test_blowup_is_reported_not_nan(jst, oracle_lib):
    """properties.rs:203-221: CFL 200 must raise NonPhysical at the reference's cell."""
    text = ellipse_mesh(13, 40, 1.0)
    cfg = config_dict()
    cfg["simulation"]["CFL"] = 200.0
    s = jst.Solver(json.dumps(cfg), text)
    with pytest.raises(jst.NonPhysical) as ei:
        s.run(400)
    assert "non-physical" in str(ei.value)
    o = oracle_lib.OracleSolver(cfg, text)
    with pytest.raises(oracle_lib.NonPhysical) as eo:
        o.run(400)
    assert (ei.value.i, ei.value.j) == (eo.value.i, eo.value.j)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_set_state_reports_nonphysical(jst, precision):
    """solver.rs:358-365: rho = -1 injected at (0, 0) is reported, not fatal."""
    s = jst.Solver(config_text(), mesh_text("fangdata.txt"), jst.FP64 if precision == "fp64" else jst.MIXED)
    u = np.stack([s.get(f"U{k}") for k in range(5)])
    u[0, 0] = -1.0
    with pytest.raises(jst.NonPhysical) as ei:
        s.set_state(u)
    assert (ei.value.i, ei.value.j) == (0, 0) and ei.value.rho == -1.0


def test_write_csv_format(jst, tmp_path):
    """solver.rs:229-255 columns and Rust {:.8e} formatting."""
    s = jst.Solver(config_text(), mesh_text("fangdata.txt"))
    s.run(2)
    path = tmp_path / "result.csv"
    s.write_csv(path)
    lines = path.read_text().splitlines()
    assert lines[0] == "i,j,x,y,rho,p,T,u,v,miubl"
    assert len(lines) == 1 + 9 * 12
    first = lines[1].split(",")
    assert first[:2] == ["1", "1"] and "e" in first[4] and "+" not in first[4]
    assert abs(float(first[4]) - s.get("rho")[0]) <= 1e-8 * s.get("rho")[0]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_resume_from_an_exported_state(jst, precision):
    """Checkpoint and resume through the C ABI: U exported after 15 steps (U0..U4) and
    injected into a fresh solver (jst_set_state) continues as the uninterrupted run does.
    The frozen time step of the next step is recomputed from the injected state; on the
    FP64 path that reproduces the uninterrupted run to 1e-13 after 15 more steps, on the
    FP32-mixed path to FP32 rounding (the FP32 base and dt are re-derived from FP64 U)."""
    prec = jst.FP64 if precision == "fp64" else jst.MIXED
    text = mesh_text("cyl_17x40.txt")
    a = jst.Solver(config_text(), text, prec)
    a.run(15)
    u = np.stack([a.get(f"U{k}") for k in range(5)])
    a.run(15)
    b = jst.Solver(config_text(), text, prec)
    b.set_state(u)
    b.run(15)
    tol = 1e-13 if precision == "fp64" else 1e-6
    for n in ("rho", "u", "v", "p", "nut"):
        x, y = b.get(n), a.get(n)
        assert np.max(np.abs(x - y)) <= tol * np.max(np.abs(y)), (n, np.max(np.abs(x - y)))
