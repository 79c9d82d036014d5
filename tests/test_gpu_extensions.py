"""The config extensions (no reference counterpart; parity unpinned beyond the oracle, which
implements them with the same formulas): local time stepping, relative residual drop, on
both device paths against the oracle."""
from __future__ import annotations

import json

import numpy as np
import pytest

from conftest import config_dict, ellipse_mesh, field_cmp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def jst():
    from purinecfd_amd import jst as mod
    mod.lib()
    return mod


def _lts_case():
    cfg = config_dict()
    cfg["simulation"].update({"AOA": 2.0, "Ma": 0.4, "CFL": 1.0, "local_time_stepping": True})
    return cfg, ellipse_mesh(25, 64, 1.1)


def test_local_time_stepping_fp64_matches_oracle(jst, oracle_lib):
    cfg, text = _lts_case()
    s = jst.Solver(json.dumps(cfg), text)
    o = oracle_lib.OracleSolver(cfg, text)
    s.run(60)
    o.run(60)
    for name in ("rho", "p", "u", "v", "nut"):
        field_cmp(s.get(name), o.get(name), name, 1e-9, 1e-300)
    o.timestep()  # the device already holds the next step's local dt
    field_cmp(s.get("localdt"), o.get("localdt"), "localdt", 1e-9)


def test_local_time_stepping_mixed_within_fp32_level(jst, oracle_lib):
    cfg, text = _lts_case()
    s = jst.Solver(json.dumps(cfg), text, jst.MIXED)
    o = oracle_lib.OracleSolver(cfg, text)
    s.run(60)
    o.run(60)
    for name in ("rho", "p", "u", "v"):
        a, b = s.get(name), o.get(name)
        assert np.linalg.norm(a - b) / np.linalg.norm(b) <= 1e-5, name


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_residual_drop_stops_where_the_oracle_does(jst, oracle_lib, precision):
    """solver.residual_drop = 1.5: converged once residual < residual(step 1) * 10^-1.5."""
    cfg = config_dict()
    cfg["simulation"].update({"Ma": 0.3, "CFL": 1.0})
    cfg["solver"]["residual_drop"] = 1.5
    cfg["solver"]["targetres"] = 0.0
    text = ellipse_mesh(13, 40)
    s = jst.Solver(json.dumps(cfg), text, jst.FP64 if precision == "fp64" else jst.MIXED)
    o = oracle_lib.OracleSolver(cfg, text)
    rep = s.run(3000)
    orep = o.run(3000)
    assert orep["converged"] and rep.converged
    if precision == "fp64":
        assert rep.steps == orep["steps"]
    else:  # the drop is crossed within a step or two of the FP64 step
        assert abs(rep.steps - orep["steps"]) <= 2


def _visc_case():
    cfg = config_dict()
    cfg["simulation"].update({"AOA": 2.0, "Ma": 0.4, "CFL": 1.0, "local_time_stepping": True,
                              "viscous_time_step": True})
    cfg["physics"]["mu0"] *= 1e3  # viscous enough that the viscous radius moves the local dt
    return cfg, ellipse_mesh(25, 64, 1.1)


def test_viscous_time_step_fp64_matches_oracle(jst, oracle_lib):
    cfg, text = _visc_case()
    s = jst.Solver(json.dumps(cfg), text)
    o = oracle_lib.OracleSolver(cfg, text)
    hist = []
    s.run(60, on_step=lambda k, r: hist.append(r))
    orep = o.run(60)
    field_cmp(np.array(hist), orep["history"], "residual history", 1e-9)
    for name in ("rho", "p", "u", "v", "nut"):
        field_cmp(s.get(name), o.get(name), name, 1e-9, 1e-300)
    o.timestep()
    field_cmp(s.get("localdt"), o.get("localdt"), "localdt", 1e-9)
    conv = oracle_lib.OracleSolver({**cfg, "simulation": {**cfg["simulation"], "viscous_time_step": False}}, text)
    conv.set_state(np.stack([o.get(f"U{k}") for k in range(5)]))
    conv.timestep()
    assert np.all(o.get("localdt") < conv.get("localdt"))  # the viscous radius is active


def test_viscous_time_step_mixed_within_fp32_level(jst, oracle_lib):
    cfg, text = _visc_case()
    s = jst.Solver(json.dumps(cfg), text, jst.MIXED)
    o = oracle_lib.OracleSolver(cfg, text)
    s.run(60)
    o.run(60)
    for name in ("rho", "p", "u", "v"):
        a, b = s.get(name), o.get(name)
        assert np.linalg.norm(a - b) / np.linalg.norm(b) <= 1e-5, name
    o.timestep()
    a, b = s.get("localdt"), o.get("localdt")
    assert np.max(np.abs(a - b) / b) <= 1e-4
