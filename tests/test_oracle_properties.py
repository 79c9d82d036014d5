"""Cross-module invariants of the CPU oracle, and checks fixed by mathematics for the oracle's
functions that the reference does not have.

The first group ports jst-rs tests/properties.rs (thread-count independence 49-75, discrete
mass conservation 82-107, zero wall mass flux 110-123, free-stream preservation 126-156,
convergence 178-201, dt shrinking under refinement 224-233) to the oracle.  The second group
pins the default-off extensions, which no reference output covers, by properties that a
plausible mistake breaks:

* CL/CD: the discrete scheme is invariant under a rotation of the plane, so AoA alpha on a
  mesh equals AoA 0 on the same mesh rotated by -alpha (a wrong sign in a sin(alpha) term of
  the force rotation, or a swapped lift/drag, fails it);
* local time stepping: one RK stage moves every cell by exactly alpha * dt_local(i,j) * R / V
  of that cell, and a converged steady state does not depend on the time step;
* residual_drop: the stop step is the first step whose residual is below the step-1 residual
  times 10^-drop, read from the history;
* nut_ratio: nu~ = ratio * mu_inf / rho_inf (Sutherland mu_inf, ideal-gas rho_inf in closed
  form) in the initial field and in the far-field inflow ghosts, while outflow ghosts
  extrapolate the interior value;
* IM (halo depth): any IM >= 3 gives bitwise the same run.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import (check_mass_conservation, config_dict, ellipse_mesh, ring_mass_flux,
                      rotate_mesh)


def _solver(oracle_lib, rings, nj, stretch, **over):
    cfg = config_dict()
    for sect, kv in over.items():
        cfg[sect].update(kv)
    return oracle_lib.OracleSolver(cfg, ellipse_mesh(rings, nj, stretch))


def test_results_are_independent_of_thread_count(oracle_lib):
    """properties.rs:49-75: every output is written once, reductions are serial."""
    lib = oracle_lib.lib()
    n0 = lib.jo_threads()
    try:
        runs = []
        for threads in (1, 2, 3, 7):
            lib.jo_set_threads(threads)
            s = _solver(oracle_lib, 17, 48, 1.1)
            s.run(25)
            runs.append((s.get("rho"), s.get("u"), s.get("nut"), s.totaltime))
        for got in runs[1:]:
            for a, b in zip(runs[0][:3], got[:3]):
                assert np.array_equal(a, b)
            assert runs[0][3] == got[3]
    finally:
        lib.jo_set_threads(n0)


def test_discrete_mass_conservation(oracle_lib):
    """properties.rs:82-107: interior faces cancel; the total equals wall + far-field flux."""
    s = _solver(oracle_lib, 17, 48, 1.1)
    s.run(30)
    s.timestep()
    s.residual_terms()
    ni, nj = s.ni, s.nj
    fc0 = s.get("Fc0")
    tf = s.get("tau_flux0").reshape(ni + 1, nj)
    total = float(fc0.sum())
    boundary = float((tf[ni] - tf[0]).sum())
    scale = float(np.abs(fc0).max())
    assert abs(total - boundary) < 1e-12 * max(scale, 1.0), (total, boundary)
    # ring by ring, from the state alone (the check the GPU paths also take)
    check_mass_conservation(fc0, s.get("U1"), s.get("U2"), s.get("tau_nx"), s.get("tau_ny"),
                            ni, nj, 1e-13)
    # the oracle's interior ring fluxes are the central mass flux of the state
    phi = ring_mass_flux(s.get("U1"), s.get("U2"), s.get("tau_nx"), s.get("tau_ny"), ni, nj)
    np.testing.assert_allclose(tf[1:ni].sum(axis=1), phi, rtol=0, atol=1e-13 * np.abs(tf).sum())


def test_no_mass_flux_through_the_wall(oracle_lib):
    """properties.rs:110-123: the mirror ghost cancels the face momentum."""
    s = _solver(oracle_lib, 17, 48, 1.1)
    s.run(30)
    s.timestep()
    s.residual_terms()
    tf = s.get("tau_flux0").reshape(s.ni + 1, s.nj)
    scale = float(np.abs(tf[1]).max())
    assert float(np.abs(tf[0]).max()) < 1e-13 * max(scale, 1.0)


def test_free_stream_is_preserved_end_to_end(oracle_lib):
    """properties.rs:126-156: a uniform state (ghosts included, no BC applied) has a zero
    mean-flow residual: metric closure, sum of the normals of a cell = 0."""
    s = _solver(oracle_lib, 13, 40, 1.0)
    s.set_uniform(1.176, 69.4, 17.3, 101325.0, 1.5e-4)
    s.timestep()
    for k in ("convection", "gradient", "viscous", "dissipation"):
        s.call(k)
    scale = max(float(np.abs(s.get(f"tau_flux{k}")).max()) for k in range(5))
    for k in range(4):
        r = s.get(f"Fc{k}") - s.get(f"Fv{k}") - s.get(f"Fd{k}")
        assert float(np.abs(r).max()) < 1e-13 * scale, (k, float(np.abs(r).max()), scale)


def test_timestep_shrinks_with_mesh_refinement(oracle_lib):
    """properties.rs:224-233 (and timestep.rs:120-143)."""
    coarse = _solver(oracle_lib, 9, 32, 1.0).timestep()
    fine = _solver(oracle_lib, 17, 64, 1.0).timestep()
    assert fine < coarse
    assert fine > 0.2 * coarse


def test_converges_on_a_moderate_mesh(oracle_lib):
    """properties.rs:178-201: 6 decades of residual drop, still physical."""
    s = _solver(oracle_lib, 9, 32, 1.0, solver={"targetres": 1e-9})
    rep = s.run(6000)
    assert rep["converged"]
    assert rep["history"][0] / rep["residual"] > 1e6
    assert (s.get("rho") > 0).all() and (s.get("p") > 0).all()


# ---- extensions (no reference counterpart) ----

@pytest.mark.parametrize("aoa", [5.0, -3.0])
def test_forces_are_invariant_under_rotation(oracle_lib, aoa):
    """CL/CD at AoA alpha on a mesh = CL/CD at AoA 0 on the mesh rotated by -alpha (the
    discrete scheme commutes with rotations of the plane); the body-axis forces rotate."""
    mesh = ellipse_mesh(17, 48, 1.1)
    cfg_a = config_dict()
    cfg_a["simulation"]["AOA"] = aoa
    cfg_b = config_dict()
    cfg_b["simulation"]["AOA"] = 0.0
    a = oracle_lib.OracleSolver(cfg_a, mesh)
    b = oracle_lib.OracleSolver(cfg_b, rotate_mesh(mesh, -math.radians(aoa)))
    a.run(20)
    b.run(20)
    fa, fb = a.forces(), b.forces()
    assert abs(fa["cl"]) > 1e-3  # lift is present: the check sees the sin(alpha) terms
    assert abs(fa["cl"] - fb["cl"]) <= 1e-9 * abs(fa["cl"])
    assert abs(fa["cd"] - fb["cd"]) <= 1e-9 * abs(fa["cd"])
    t = -math.radians(aoa)
    fx = math.cos(t) * fa["fx"] - math.sin(t) * fa["fy"]
    fy = math.sin(t) * fa["fx"] + math.cos(t) * fa["fy"]
    scale = math.hypot(fa["fx"], fa["fy"])
    assert abs(fx - fb["fx"]) <= 1e-9 * scale and abs(fy - fb["fy"]) <= 1e-9 * scale
    np.testing.assert_allclose(a.get("rho"), b.get("rho"), rtol=1e-11)
    np.testing.assert_allclose(a.get("p"), b.get("p"), rtol=1e-11)


@pytest.mark.parametrize("lts", [False, True])
def test_one_stage_moves_each_cell_by_its_own_time_step(oracle_lib, lts):
    """One RK stage: (U_former - U) * V / (alpha * dt) = Fc - Fv - Fd - S per cell, dt the
    frozen global minimum, or the cell's own local dt under local time stepping."""
    s = _solver(oracle_lib, 13, 40, 1.2, simulation={"local_time_stepping": lts})
    s.run(5)
    u0 = np.stack([s.get(f"U{k}") for k in range(5)])
    dt = s.begin_step()
    alpha = 0.25
    s.stage(alpha, dt)
    u1 = np.stack([s.get(f"U{k}") for k in range(5)])
    ldt = s.get("localdt")
    assert float(dt) == float(ldt.min())
    if lts:
        assert ldt.max() > 2.0 * ldt.min()  # the local steps really differ
    step = ldt if lts else np.full_like(ldt, dt)
    vol = s.get("vol")
    for k in range(5):
        r = s.get(f"Fc{k}") - s.get(f"Fv{k}") - s.get(f"Fd{k}")
        if k == 4:
            r = r - s.get("S")
        got = (u0[k] - u1[k]) * vol / (alpha * step)
        scale = float(np.abs(r).max())
        assert float(np.abs(got - r).max()) <= 1e-9 * scale, (k, float(np.abs(got - r).max()), scale)


def test_local_time_stepping_reaches_the_same_steady_state(oracle_lib):
    """The steady state solves R(U) = 0, whatever the time step: global-dt and local-dt runs
    converged to a residual of 1e-12 agree to that level."""
    mesh = ellipse_mesh(9, 32, 1.0)
    runs = {}
    for lts in (False, True):
        cfg = config_dict()
        cfg["simulation"]["local_time_stepping"] = lts
        cfg["solver"]["targetres"] = 1e-12
        s = oracle_lib.OracleSolver(cfg, mesh)
        rep = s.run(40000)
        assert rep["converged"], (lts, rep["residual"], rep["steps"])
        runs[lts] = (s, rep)
    assert runs[True][1]["steps"] < runs[False][1]["steps"]  # local steps converge faster
    for name in ("rho", "u", "v", "p"):
        a, b = runs[False][0].get(name), runs[True][0].get(name)
        scale = float(np.abs(a).max())
        assert float(np.abs(a - b).max()) <= 1e-7 * scale, (name, float(np.abs(a - b).max()))


def test_residual_drop_stops_at_the_first_step_below_the_drop(oracle_lib):
    s = _solver(oracle_lib, 9, 32, 1.0, solver={"targetres": 0.0, "residual_drop": 3.0})
    rep = s.run(6000)
    h = rep["history"]
    below = np.nonzero(h < h[0] * 1e-3)[0]
    assert rep["converged"] and below.size > 0
    assert rep["steps"] == int(below[0]) + 1
    assert len(h) == rep["steps"]


def test_nut_ratio_sets_the_initial_field_and_the_inflow_ghosts(oracle_lib):
    cfg = config_dict()
    cfg["spalart_allmaras"]["nut_ratio"] = 3.0
    cfg["simulation"]["AOA"] = 10.0
    mesh = ellipse_mesh(13, 40, 1.2)
    s = oracle_lib.OracleSolver(cfg, mesh)
    ph, sim = cfg["physics"], cfg["simulation"]
    t, t0, ts = sim["T"], ph["T0"], ph["Ts"]
    mu_inf = ph["mu0"] * (t / t0) ** 1.5 * (t0 + ts) / (t + ts)  # Sutherland
    rho_inf = sim["P"] / (ph["R"] * t)
    nut_inf = 3.0 * mu_inf / rho_inf
    np.testing.assert_allclose(s.get("nut"), nut_inf, rtol=1e-14)
    # interior nu~ doubled, mean flow left at the free stream: inflow ghosts take nu~_inf,
    # outflow ghosts extrapolate 2 nu~_inf, so the far-field face carries 1.5 resp. 2 times
    # rho_inf nu~_inf v_n
    u = np.stack([s.get(f"U{k}") for k in range(5)])
    u[4] *= 2.0
    s.set_state(u)
    s.residual_terms()
    ni, nj = s.ni, s.nj
    f4 = s.get("tau_flux4").reshape(ni + 1, nj)[ni]
    nx = s.get("tau_nx").reshape(ni + 1, nj)[ni]
    ny = s.get("tau_ny").reshape(ni + 1, nj)[ni]
    c = math.sqrt(ph["gamma"] * ph["R"] * t)
    a = math.radians(10.0)
    vn = c * sim["Ma"] * (math.cos(a) * nx + math.sin(a) * ny)
    ratio = f4 / (rho_inf * nut_inf * vn)
    clear = np.abs(vn) > 0.2 * c * sim["Ma"] * np.hypot(nx, ny)
    inflow, outflow = clear & (vn < 0), clear & (vn > 0)
    assert inflow.sum() > 5 and outflow.sum() > 5
    np.testing.assert_allclose(ratio[inflow], 1.5, rtol=1e-9)
    # Outflow: the reference's three-point formula (boundary.rs:158-185, the same in jst-py
    # boundary.py:27-35) from the distances l1 < l2 < l3 of the three outer cell centres to
    # the face midpoint.  Its first two weights are the Lagrange weights of extrapolation to
    # l = 0; its third, l1 l2 / (d23 (-d13)), has the opposite sign of the Lagrange weight
    # l1 l2 / (d13 d23), so a uniform nu~ extrapolates to nu~ (1 - 2 w3) (floored at 1e-10)
    # rather than to nu~.  The oracle keeps the reference's formula, as parity requires.
    cx = s.get("cx").reshape(ni, nj)
    cy = s.get("cy").reshape(ni, nj)
    mx = s.get("tau_mx").reshape(ni + 1, nj)[ni]
    my = s.get("tau_my").reshape(ni + 1, nj)[ni]
    l1, l2, l3 = (np.hypot(cx[ni - 1 - k] - mx, cy[ni - 1 - k] - my) for k in range(3))
    d12, d13, d23 = l1 - l2, l1 - l3, l2 - l3
    lagrange3 = l1 * l2 / (d13 * d23)
    w = l2 * l3 / (d12 * d13) + l1 * l3 / (-d12 * d23) - lagrange3
    np.testing.assert_allclose(l2 * l3 / (d12 * d13) + l1 * l3 / (-d12 * d23) + lagrange3, 1.0,
                               rtol=1e-12)  # the true Lagrange weights sum to one
    ghost = np.maximum(2.0 * nut_inf * w, 1e-10)
    np.testing.assert_allclose(ratio[outflow], (0.5 * (2.0 * nut_inf + ghost) / nut_inf)[outflow],
                               rtol=1e-9)


def test_halo_depth_beyond_three_changes_nothing(oracle_lib):
    """config.rs:33-35: IM is the number of ghost layers; the stencils read three."""
    a = _solver(oracle_lib, 13, 40, 1.2)
    b = _solver(oracle_lib, 13, 40, 1.2, simulation={"IM": 5})
    assert b.cfg.halo == 5
    a.run(20)
    b.run(20)
    for name in ("rho", "u", "v", "p", "nut"):
        assert np.array_equal(a.get(name), b.get(name)), name


# ---- viscous time step (extension simulation.viscous_time_step) ----

def test_viscous_time_step_keeps_a_viscous_case_stable(oracle_lib):
    """At very low Reynolds number the explicit viscous limit, not the acoustic one, bounds
    the step: the reference's convective-only dt diverges within a few steps, the dt with the
    viscous spectral radius stays physical."""
    over = {"physics": {"mu0": config_dict()["physics"]["mu0"] * 1e7}}
    a = _solver(oracle_lib, 17, 48, 1.1, **over)
    with pytest.raises(oracle_lib.NonPhysical):
        a.run(300)
    b = _solver(oracle_lib, 17, 48, 1.1, simulation={"viscous_time_step": True}, **over)
    rep = b.run(300)
    assert rep["steps"] == 300 and (b.get("rho") > 0).all() and (b.get("p") > 0).all()


def test_viscous_time_step_vanishes_without_viscosity(oracle_lib):
    """The viscous radius is additive: as mu -> 0 the local dt tends to the convective one,
    and it is never larger; the JST spectral radius (dissipation) is the convective one in
    both cases, so Fd is bitwise unchanged on the same state."""
    for mu_scale, rtol in ((1e-7, 1e-9), (1.0, None)):
        over = {"physics": {"mu0": config_dict()["physics"]["mu0"] * mu_scale}}
        a = _solver(oracle_lib, 13, 40, 1.2, **over)
        b = _solver(oracle_lib, 13, 40, 1.2, simulation={"viscous_time_step": True}, **over)
        a.run(10)
        u = np.stack([a.get(f"U{k}") for k in range(5)])
        b.set_state(u)
        a.timestep()
        b.timestep()
        la, lb = a.get("localdt"), b.get("localdt")
        assert (lb <= la).all()
        if rtol is not None:
            np.testing.assert_allclose(lb, la, rtol=rtol)
        else:
            assert (lb < la).all()
        a.residual_terms()
        b.residual_terms()
        for k in range(5):
            assert np.array_equal(a.get(f"Fd{k}"), b.get(f"Fd{k}"))


def test_viscous_local_time_stepping_reaches_the_global_dt_steady_state(oracle_lib):
    mesh = ellipse_mesh(9, 32, 1.0)
    runs = []
    for sim in ({}, {"local_time_stepping": True, "viscous_time_step": True}):
        cfg = config_dict()
        cfg["simulation"].update(sim)
        cfg["solver"]["targetres"] = 1e-12
        s = oracle_lib.OracleSolver(cfg, mesh)
        rep = s.run(40000)
        assert rep["converged"], (sim, rep["residual"])
        runs.append(s)
    for name in ("rho", "u", "v", "p"):
        a, b = runs[0].get(name), runs[1].get(name)
        assert float(np.abs(a - b).max()) <= 1e-7 * float(np.abs(a).max()), name
