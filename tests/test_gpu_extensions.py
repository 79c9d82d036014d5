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


# ---- negative-nu~ SA branch (extension spalart_allmaras.negative) ----

def _neg_case():
    """C4 case 8 (129x65, M 0.42, Re 4.2e6) with local time stepping at CFL 2: nu~ turns
    negative next to the wall after about 195 steps, and chi passes the fv1 pole (-cv1 = -7.1)
    at step 256 and reaches -14 by step 300."""
    from purinecfd_amd import workloads
    w = workloads.c4_batch(n=9)[8]
    cfg = json.loads(w["config"])
    cfg["simulation"].update({"CFL": 2.0, "local_time_stepping": True})
    cfg["spalart_allmaras"]["negative"] = True
    return cfg, w["mesh"]


@pytest.fixture(scope="module")
def neg_oracle(oracle_lib):
    """The oracle's run to step 240 (history and fields), and its state at step 300."""
    cfg, text = _neg_case()
    o = oracle_lib.OracleSolver(cfg, text)
    rep = o.run(240)
    assert o.get("nut").min() < 0.0
    at240 = {n: o.get(n) for n in ("rho", "p", "u", "v", "nut")}
    at240["forces"] = o.forces()
    o.run(60)
    assert o.get("chi").min() < -7.1
    u300 = np.stack([o.get(f"U{k}") for k in range(5)])
    return cfg, text, rep["history"], at240, u300


def test_negative_sa_fp64_matches_oracle(jst, oracle_lib, neg_oracle):
    """Per residual evaluation at 1e-10 on the step-300 state (chi down to -14), and the run
    from the free stream to step 240 at 1e-9.  Past step 260 the run goes through bursts at a
    wall cell near the leading edge (the residual rises 50 % in three steps, on both sides)
    that amplify the ulp-level libm differences by about 1e5 each, so longer runs are no
    parity measure."""
    cfg, text, hist_ref, at240, u300 = neg_oracle
    s = jst.Solver(json.dumps(cfg), text)
    hist = []
    s.run(240, on_step=lambda k, r: hist.append(r))
    field_cmp(np.array(hist), hist_ref, "residual history", 1e-9)
    for name in ("rho", "p", "u", "v", "nut"):
        field_cmp(s.get(name), at240[name], name, 1e-9, 1e-300)
    o = oracle_lib.OracleSolver(cfg, text)
    o.set_state(u300)
    s.set_state(u300)
    dt = s.residual_terms()
    dt_ref = o.timestep()
    o.residual_terms()
    assert abs(dt - dt_ref) <= 1e-12 * dt_ref
    for name in ["S"] + [f"{t}{k}" for t in ("Fc", "Fv", "Fd") for k in range(5)]:
        field_cmp(s.get(name), o.get(name), name, 1e-10)


def test_negative_sa_mixed_within_fp32_level(jst, oracle_lib, neg_oracle):
    """One residual evaluation on the step-300 state within 2e-4 field-normed (the mixed path's
    per-evaluation bar), and the run to step 240 within 1e-4 in CL, CD and the field L2 norms
    (velocity as a vector): the trailing-edge transient of this run amplifies FP32 rounding."""
    cfg, text, _, at240, u300 = neg_oracle
    s = jst.Solver(json.dumps(cfg), text, jst.MIXED)
    o = oracle_lib.OracleSolver(cfg, text)
    o.set_state(u300)
    o.timestep()
    o.residual_terms()
    s.set_state(u300)
    s.residual_terms()
    errs = {}
    for k in range(5):
        want = o.get(f"Fc{k}") - o.get(f"Fv{k}") - o.get(f"Fd{k}")
        errs[f"R{k}"] = float(np.max(np.abs(s.get(f"R{k}") - want))) / float(np.max(np.abs(want)))
    want = o.get("S")
    errs["S"] = float(np.max(np.abs(s.get("S") - want))) / float(np.max(np.abs(want)))
    s = jst.Solver(json.dumps(cfg), text, jst.MIXED)
    s.run(240)
    for name in ("rho", "p"):
        errs[name] = float(np.linalg.norm(s.get(name) - at240[name]) / np.linalg.norm(at240[name]))
    dv = np.hypot(s.get("u") - at240["u"], s.get("v") - at240["v"])
    errs["velocity"] = float(np.linalg.norm(dv) / np.linalg.norm(np.hypot(at240["u"], at240["v"])))
    f, fo = s.forces(), at240["forces"]
    errs["cl"], errs["cd"] = abs(f["cl"] - fo["cl"]), abs(f["cd"] - fo["cd"])
    print("mixed vs oracle:", {k: f"{v:.1e}" for k, v in errs.items()})
    for k in ("R0", "R1", "R2", "R3", "R4", "S"):
        assert errs[k] <= 2e-4, (k, errs[k])
    for k in ("rho", "p", "velocity", "cl", "cd"):
        assert errs[k] <= 1e-4, (k, errs[k])
    assert s.get("nut").min() < 0.0


def test_negative_sa_kernels_leave_other_cases_bitwise_unchanged(jst):
    """A batch holding a case with the switch runs the kernels compiled with the branch; a case
    without it gets bitwise the results it gets alone (kernels without the branch)."""
    from purinecfd_amd.jst import Mesh
    cfg_neg, text_neg = _neg_case()
    cfg = config_dict()
    cfg["simulation"].update({"AOA": 3.0, "Ma": 0.3, "CFL": 1.0})
    text = ellipse_mesh(33, 96, 1.08)
    mesh = Mesh.parse(text)
    ni, nj = mesh.ni(), mesh.nj()
    alone = jst.Batch(1 << 30, 2)
    a = alone.add(json.dumps(cfg), text, 250)
    alone.run(steps_per_round=50)
    mix = jst.Batch(1 << 30, 2)
    m = mix.add(json.dumps(cfg), text, 250)
    n = mix.add(json.dumps(cfg_neg), text_neg, 250)
    mix.run(steps_per_round=50)
    for name in ("rho", "u", "v", "p", "nut"):
        assert np.array_equal(alone.get(a, name, ni, nj), mix.get(m, name, ni, nj)), name
    assert mix.info(n)["step"] == 250
    nm = Mesh.parse(text_neg)
    assert mix.get(n, "nut", nm.ni(), nm.nj()).min() < 0.0
