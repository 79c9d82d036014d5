"""Cross-module invariants of jst-rs tests/properties.rs, and the mathematical pins of the
extensions, on both GPU paths through the C ABI.

* Free-stream preservation (properties.rs:126-156): the initial free stream has a zero
  mean-flow residual away from the wall on the FP64 path (1e-13 of the flux scale) and exactly
  zero on the FP32-mixed path, whose fluxes are perturbations of the free stream; a uniform
  state that is not the free stream gives 1e-13 (FP64) and FP32 level (mixed).  The wall
  mirror and the far-field ghost are not uniform, so rows within the stencil reach (three
  rows) of either boundary are left out.
* Discrete mass conservation and zero wall mass flux (properties.rs:82-123): ring by ring,
  the mass residual summed over rows 0..i-1 equals the central mass flux through ring i
  computed from the exported state, with nothing through the wall (FP64 at 1e-13); the
  mixed path's ring sums of R0 = Fc0 - Fd0 follow the oracle's on the same state at FP32
  level.
* CL/CD: AoA alpha on a mesh against AoA 0 on the mesh rotated by -alpha.
* nu~ ratio: initial field and far-field inflow ghosts; residual_drop: first step below the
  drop; local time stepping: the steady state is that of the global time step.
"""
from __future__ import annotations

import json
import math

import numpy as np
import pytest

from conftest import check_mass_conservation, config_dict, ellipse_mesh, field_cmp, rotate_mesh

pytestmark = pytest.mark.gpu

PRECISIONS = ["fp64", "mixed"]


@pytest.fixture(scope="module")
def jst():
    from purinecfd_amd import jst as mod
    mod.lib()
    return mod


def _prec(jst, name):
    return jst.FP64 if name == "fp64" else jst.MIXED


def _residual(s, precision, k):
    if precision == "fp64":
        return s.get(f"Fc{k}") - s.get(f"Fv{k}") - s.get(f"Fd{k}")
    return s.get(f"R{k}")


def _packed(cfg, rho, vx, vy, p, nut, n):
    gamma = cfg["physics"]["gamma"]
    e = p / (rho * (gamma - 1.0)) + 0.5 * (vx * vx + vy * vy)
    return np.stack([np.full(n, v) for v in (rho, rho * vx, rho * vy, rho * e, rho * nut)])


def _flux_scale(s, rho, vx, vy, p, gamma=1.4):
    """Largest face flux of the state (jst-rs properties.rs:141-147 takes the largest tau
    face flux): the energy flux (rho E + p) |V| |n| or the momentum flux (p + rho |V|^2) |n|,
    with the largest face normal."""
    n = max(float(np.max(np.hypot(s.get("tau_nx"), s.get("tau_ny")))),
            float(np.max(np.hypot(s.get("n_nx"), s.get("n_ny")))))
    v2 = vx * vx + vy * vy
    rho_e = p / (gamma - 1.0) + 0.5 * rho * v2
    return max(p + rho * v2, (rho_e + p) * math.sqrt(v2)) * n


@pytest.mark.parametrize("precision", PRECISIONS)
def test_free_stream_is_preserved(jst, precision):
    cfg = config_dict()
    cfg["simulation"]["AOA"] = 7.0
    text = ellipse_mesh(21, 40, 1.1)
    s = jst.Solver(json.dumps(cfg), text, _prec(jst, precision))
    ni, nj = s.ni(), s.nj()
    d = jst.Config.from_str(json.dumps(cfg)).derived
    inner = slice(4 * nj, (ni - 4) * nj)  # rows 4 .. NI-5
    # the initial free stream
    s.residual_terms()
    scale = _flux_scale(s, d["rho_inf"], d["u_inf"], d["v_inf"], cfg["simulation"]["P"])
    for k in range(4):
        r = _residual(s, precision, k)[inner]
        if precision == "mixed":
            assert np.all(r == 0.0), (k, float(np.abs(r).max()))  # perturbation form: exact
        else:
            assert float(np.abs(r).max()) < 1e-13 * scale, (k, float(np.abs(r).max()), scale)
    # a uniform state that is not the free stream (properties.rs:129-130)
    rho, vx, vy, p, nut = 1.176, 69.4, 17.3, 101325.0, 1.5e-4
    s.set_state(_packed(cfg, rho, vx, vy, p, nut, ni * nj))
    s.residual_terms()
    scale = _flux_scale(s, rho, vx, vy, p)
    tol = 1e-13 if precision == "fp64" else 1e-6
    for k in range(4):
        r = _residual(s, precision, k)[inner]
        assert float(np.abs(r).max()) < tol * scale, (k, float(np.abs(r).max()), scale)
    # the wall row does see the non-uniform ghosts (the check above is not vacuous)
    assert float(np.abs(_residual(s, precision, 3)[:nj]).max()) > 1e-3 * scale


def test_mass_conservation_and_zero_wall_flux_fp64(jst):
    s = jst.Solver(json.dumps(config_dict()), ellipse_mesh(17, 48, 1.1), jst.FP64)
    s.run(30)
    s.residual_terms()
    ni, nj = s.ni(), s.nj()
    check_mass_conservation(s.get("Fc0"), s.get("U1"), s.get("U2"), s.get("tau_nx"),
                            s.get("tau_ny"), ni, nj, 1e-13)


def test_mass_conservation_mixed_follows_the_oracle(jst, oracle_lib):
    cfg = config_dict()
    text = ellipse_mesh(17, 48, 1.1)
    s = jst.Solver(json.dumps(cfg), text, jst.MIXED)
    s.run(30)
    u = np.stack([s.get(f"U{k}") for k in range(5)])
    s.residual_terms()
    o = oracle_lib.OracleSolver(cfg, text)
    o.set_state(u)
    o.timestep()
    o.residual_terms()
    ni, nj = s.ni(), s.nj()
    want = (o.get("Fc0") - o.get("Fd0")).reshape(ni, nj)
    got = s.get("R0").reshape(ni, nj)
    cum_w, cum_g = np.cumsum(want.sum(axis=1)), np.cumsum(got.sum(axis=1))
    # FP32 level of the face fluxes the mixed path sums: its mass fluxes are the face
    # perturbation momentum (m - m_inf) . n, four faces a cell
    d = jst.Config.from_str(json.dumps(cfg)).derived
    dm = (np.abs(u[1] - d["rho_inf"] * d["u_inf"]) + np.abs(u[2] - d["rho_inf"] * d["v_inf"])).reshape(ni, nj)
    nmax = max(float(np.max(np.hypot(s.get("tau_nx"), s.get("tau_ny")))),
               float(np.max(np.hypot(s.get("n_nx"), s.get("n_ny")))))
    scale = 4.0 * float(dm.sum()) * nmax
    err = float(np.abs(cum_g - cum_w).max())
    assert err <= 1e-5 * scale, (err, scale)
    # the wall row alone: no mass enters or leaves through the wall on either side
    assert abs(float(got[0].sum() - want[0].sum())) <= 1e-5 * 4.0 * float(dm[0].sum()) * nmax


@pytest.mark.parametrize("precision", PRECISIONS)
def test_forces_are_invariant_under_rotation(jst, precision):
    aoa = 5.0
    text = ellipse_mesh(17, 48, 1.1)
    cfg_a, cfg_b = config_dict(), config_dict()
    cfg_a["simulation"]["AOA"] = aoa
    a = jst.Solver(json.dumps(cfg_a), text, _prec(jst, precision))
    b = jst.Solver(json.dumps(cfg_b), rotate_mesh(text, -math.radians(aoa)), _prec(jst, precision))
    a.run(20)
    b.run(20)
    fa, fb = a.forces(), b.forces()
    assert abs(fa["cl"]) > 1e-3
    tol = 1e-9 if precision == "fp64" else 1e-5
    assert abs(fa["cl"] - fb["cl"]) <= tol * abs(fa["cl"]), (fa, fb)
    assert abs(fa["cd"] - fb["cd"]) <= tol * abs(fa["cd"]), (fa, fb)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_nut_ratio_initial_field_and_inflow_ghosts(jst, oracle_lib, precision):
    cfg = config_dict()
    cfg["spalart_allmaras"]["nut_ratio"] = 3.0
    cfg["simulation"]["AOA"] = 10.0
    text = ellipse_mesh(13, 40, 1.2)
    s = jst.Solver(json.dumps(cfg), text, _prec(jst, precision))
    ph, sim = cfg["physics"], cfg["simulation"]
    t, t0, ts = sim["T"], ph["T0"], ph["Ts"]
    mu_inf = ph["mu0"] * (t / t0) ** 1.5 * (t0 + ts) / (t + ts)
    rho_inf = sim["P"] / (ph["R"] * t)
    nut_inf = 3.0 * mu_inf / rho_inf
    np.testing.assert_allclose(s.get("nut"), nut_inf, rtol=1e-14)
    if precision != "fp64":
        return
    # interior nu~ doubled: the convective nu~ residual of the outer row is the far-face flux
    # minus the interior value's, (ghost/nu_inf - 2)/2 rho_inf nu_inf v_n: -1/2 on inflow faces
    ni, nj = s.ni(), s.nj()
    u = np.stack([s.get(f"U{k}") for k in range(5)])
    u[4] *= 2.0
    s.set_state(u)
    s.residual_terms()
    fc4 = s.get("Fc4").reshape(ni, nj)[ni - 1]
    nx = s.get("tau_nx").reshape(ni + 1, nj)[ni]
    ny = s.get("tau_ny").reshape(ni + 1, nj)[ni]
    c = math.sqrt(ph["gamma"] * ph["R"] * t)
    a = math.radians(10.0)
    vn = c * sim["Ma"] * (math.cos(a) * nx + math.sin(a) * ny)
    inflow = vn < -0.2 * c * sim["Ma"] * np.hypot(nx, ny)
    assert inflow.sum() > 5
    np.testing.assert_allclose(fc4[inflow] / (rho_inf * nut_inf * vn[inflow]), -0.5, rtol=1e-8)
    o = oracle_lib.OracleSolver(cfg, text)  # outflow faces: the oracle's extrapolation
    o.set_state(u)
    o.residual_terms()
    field_cmp(fc4, o.get("Fc4").reshape(ni, nj)[ni - 1], "Fc4 outer row", 1e-10, 1e-300)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_residual_drop_stops_at_the_first_step_below_the_drop(jst, precision):
    cfg = config_dict()
    cfg["solver"].update({"targetres": 0.0, "residual_drop": 3.0})
    s = jst.Solver(json.dumps(cfg), ellipse_mesh(9, 32, 1.0), _prec(jst, precision))
    hist = []
    rep = s.run(6000, on_step=lambda k, r: hist.append(r))
    h = np.array(hist)
    below = np.nonzero(h < h[0] * 1e-3)[0]
    assert rep.converged and below.size > 0
    assert rep.steps == len(h) == int(below[0]) + 1


def test_local_time_stepping_reaches_the_global_dt_steady_state(jst, oracle_lib):
    text = ellipse_mesh(9, 32, 1.0)
    out = {}
    for lts in (False, True):
        cfg = config_dict()
        cfg["simulation"]["local_time_stepping"] = lts
        cfg["solver"]["targetres"] = 1e-12
        s = jst.Solver(json.dumps(cfg), text, jst.FP64)
        rep = s.run(40000)
        assert rep.converged, (lts, rep)
        out[lts] = (s, rep)
    assert out[True][1].steps < out[False][1].steps
    for name in ("rho", "u", "v", "p"):
        a, b = out[False][0].get(name), out[True][0].get(name)
        assert float(np.abs(a - b).max()) <= 1e-7 * float(np.abs(a).max()), name


def test_halo_depth_beyond_three_changes_nothing(jst):
    """config.rs:33-35: IM >= 3 is accepted and gives bitwise the same run; IM < 3 is a config
    error (the stencils read three ghost layers)."""
    cfg = config_dict()
    text = ellipse_mesh(13, 40, 1.2)
    for precision in PRECISIONS:
        a = jst.Solver(json.dumps(cfg), text, _prec(jst, precision))
        cfg5 = config_dict()
        cfg5["simulation"]["IM"] = 5
        b = jst.Solver(json.dumps(cfg5), text, _prec(jst, precision))
        a.run(20)
        b.run(20)
        for name in ("rho", "u", "v", "p", "nut"):
            assert np.array_equal(a.get(name), b.get(name)), (precision, name)
    cfg2 = config_dict()
    cfg2["simulation"]["IM"] = 2
    with pytest.raises(jst.ConfigError):
        jst.Solver(json.dumps(cfg2), text)


@pytest.mark.parametrize("precision", PRECISIONS)
def test_rejected_state_leaves_the_solver_unchanged(jst, precision):
    s = jst.Solver(json.dumps(config_dict()), ellipse_mesh(13, 40, 1.2), _prec(jst, precision))
    s.run(5)
    before = {n: s.get(n) for n in ("rho", "u", "p")}
    u = np.stack([s.get(f"U{k}") for k in range(5)])
    u[0, 7] = -1.0
    with pytest.raises(jst.NonPhysical):
        s.set_state(u)
    for n, v in before.items():
        assert np.array_equal(s.get(n), v), n
    assert s.run(5).steps == 5
