"""The CPU oracle reproduces the reference golden vectors.

Mirrors jst-rs tests/golden.rs (staged: meta -> geometry -> init -> single-stage terms after
warm-up -> final state), with the same field-normed tolerances (golden.rs:11-21).  The
golden JSONs are the reference's own fixtures (tests/golden/cyl_10x12.json and
cyl_17x40.json, produced by jst-py/tools/dump_golden.py).
"""
from __future__ import annotations

import pytest

from conftest import GOLDEN_FILES, config_dict, field_cmp, load_golden, mesh_text

MESHES = {"fangdata.txt": "fangdata.txt", "cyl_17x40.txt": "cyl_17x40.txt"}


@pytest.fixture(params=GOLDEN_FILES, ids=lambda p: p.stem)
def golden(request, oracle_lib):
    doc = load_golden(request.param)
    text = mesh_text(MESHES[doc["meta"]["mesh"]])
    return doc, text, oracle_lib


def test_golden_files_present():
    assert len(GOLDEN_FILES) == 2


def test_meta_agrees(golden):
    doc, text, po = golden
    s = po.OracleSolver(config_dict(), text)
    m = doc["meta"]
    assert s.ni + 1 == m["i_total"] and s.nj == m["j_total"] and s.n_cells == m["n_cells"]
    for key, val in (("gamma", s.cfg.gamma), ("R", s.cfg.r_gas), ("CFL", s.cfg.cfl),
                     ("Ma", s.cfg.mach), ("k2", s.cfg.k2), ("k4", s.cfg.k4)):
        assert abs(m[key] - val) < 1e-15, key


def test_geometry_agrees(golden):
    doc, text, po = golden
    s = po.OracleSolver(config_dict(), text)
    g = doc["geometry"]
    field_cmp(s.get("vol"), g["cell_vol"], "vol", 1e-14)
    field_cmp(s.get("cx"), g["cell_x"], "cx", 1e-13, 1e-15)
    field_cmp(s.get("cy"), g["cell_y"], "cy", 1e-13, 1e-15)
    field_cmp(s.get("wall_dist"), g["cell_sad"], "wall_dist", 1e-14)
    for key in ("tau_nx", "tau_ny", "tau_mx", "tau_my", "n_nx", "n_ny", "n_mx", "n_my"):
        field_cmp(s.get(key), g[key], key, 1e-13, 1e-15)


def test_initial_state_agrees(golden):
    doc, text, po = golden
    s = po.OracleSolver(config_dict(), text)
    for key, name in (("rho", "rho"), ("p", "p"), ("T", "T"), ("u", "u"), ("v", "v"),
                      ("E", "E"), ("H", "H"), ("c", "c"), ("miubl", "nut")):
        field_cmp(s.get(name), doc["init"][key], f"init.{key}", 1e-14, 1e-300)
    for k in range(5):
        field_cmp(s.get(f"U{k}"), doc["init"][f"U{k + 1}"], f"init.U{k + 1}", 1e-14, 1e-300)


GRADS = ["ugrad_x", "ugrad_y", "vgrad_x", "vgrad_y", "Tgrad_x", "Tgrad_y",
         "miublgrad_x", "miublgrad_y"]


def test_single_stage_residual_terms_agree(golden):
    """golden.rs:221-275: warm-up, timestep, residual_terms, compare at 1e-10."""
    doc, text, po = golden
    s = po.OracleSolver(config_dict(), text)
    s.run(doc["meta"]["warmup"])
    dt = s.timestep()
    s.residual_terms()
    t = doc["terms"]
    assert abs(dt - t["dt"]) < 1e-12 * t["dt"]
    for k, key in enumerate(GRADS):
        field_cmp(s.get(f"grad{k}"), t[key], key, 1e-10, 1e-280)
    field_cmp(s.get("chi"), t["chi"], "chi", 1e-10, 1e-280)
    field_cmp(s.get("fv1"), t["fv1"], "fv1", 1e-10, 1e-280)
    for k in range(5):
        for term in ("Fc", "Fv", "Fd"):
            field_cmp(s.get(f"{term}{k}"), t[f"{term}{k + 1}"], f"{term}{k + 1}", 1e-10, 1e-280)
    field_cmp(s.get("S"), t["S5"], "S5", 1e-10, 1e-280)


def test_final_state_agrees(golden):
    """golden.rs:278-311: run meta.steps, compare residual, totaltime and fields."""
    doc, text, po = golden
    s = po.OracleSolver(config_dict(), text)
    rep = s.run(doc["meta"]["steps"])
    f = doc["final"]
    assert abs(rep["residual"] - f["residual"]) < 1e-8 * f["residual"]
    assert abs(rep["totaltime"] - f["totaltime"]) < 1e-10 * f["totaltime"]
    for key, name in (("rho", "rho"), ("p", "p"), ("T", "T"), ("u", "u"), ("v", "v"),
                      ("miubl", "nut")):
        field_cmp(s.get(name), f[key], f"final.{key}", 1e-8, 1e-300)


def test_planted_mistake_is_caught(golden, monkeypatch):
    """A wrong-index class of error (swapping two gradient components) fails the terms
    check by many orders of magnitude more than the 1e-10 tolerance."""
    doc, text, po = golden
    s = po.OracleSolver(config_dict(), text)
    s.run(doc["meta"]["warmup"])
    s.timestep()
    s.residual_terms()
    with pytest.raises(AssertionError):
        field_cmp(s.get("grad1"), doc["terms"]["vgrad_x"], "swapped", 1e-10, 1e-280)
