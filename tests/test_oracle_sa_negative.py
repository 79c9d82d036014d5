"""The negative-nu~ branch of the Spalart-Allmaras model (extension ``spalart_allmaras.negative``,
default off; Allmaras, Johnson & Spalart, "Modifications and Clarifications for the
Implementation of the Spalart-Allmaras Turbulence Model", ICCFD7-1902, 2012, eqs. 19-21).

The reference evaluates fv1 = chi^3 / (chi^3 + cv1^3) for every sign of nu~ (jst-rs
viscous.rs:27-35), which has a pole at chi = -cv1: a cell whose nu~ drifts negative through
chi = -7.1 gets an unbounded eddy viscosity, and so does the wall face of any first cell with
chi = +7.1, because its mirror ghost carries -nu~ (boundary.rs:30-51).  Where nu~ < 0 (wall
ghosts included) the branch uses no eddy viscosity, the nu~ diffusion coefficient
mu + rho nu~ fn with fn = (16 + chi^3) / (16 - chi^3), and the source
cb1 (1 - ct3) S rho nu~ + cw1 rho (nu~/d)^2 (+ the unchanged cb2 term) on the vorticity S.
No reference output covers it; the oracle's implementation is pinned by closed forms on
designed states, each of which a plausible mistake (a sign, S~ for S, fn inverted or left
out, the eddy viscosity kept) breaks:

* with nu~ >= 0 in every cell the switch changes only the viscous fluxes through the wall
  faces (whose ghosts hold -nu~), bitwise, and there it removes the reference's pole;
* where nu~ < 0 the momentum and energy viscous fluxes are bitwise those of nu~ = 0;
* the source on a uniform negative state is cw1 rho nu~^2 / d^2 V, and on a shear flow the
  production term follows from the exported vorticity;
* the nu~ viscous flux on a negative nu~ field linear in x equals a face-by-face assembly of
  sigma mu (1 + chi fn(chi)) grad nu~ from the exported gradients, mu and chi.
"""
from __future__ import annotations

import numpy as np

from conftest import config_dict, ellipse_mesh

CN1 = 16.0


def _solver(oracle_lib, negative, rings=17, nj=48, stretch=1.1):
    cfg = config_dict()
    cfg["spalart_allmaras"]["negative"] = negative
    return oracle_lib.OracleSolver(cfg, ellipse_mesh(rings, nj, stretch))


def _state(s, rho, u, v, p, nut):
    """Conserved planes (5, NI*NJ) from per-cell primitives (scalars or arrays)."""
    n = s.ni * s.nj
    rho, u, v, p, nut = (np.broadcast_to(np.asarray(x, dtype=float), (n,)) for x in (rho, u, v, p, nut))
    e = p / (rho * (s.cfg.gamma - 1.0)) + 0.5 * (u * u + v * v)
    return np.stack([rho, rho * u, rho * v, rho * e, rho * nut])


def _interior(a, s):
    """Rows 1..NI-2: cells whose faces join interior cells only."""
    return a.reshape(s.ni, s.nj)[1:-1]


def test_switch_changes_only_the_wall_faces_while_nut_stays_positive(oracle_lib):
    a = _solver(oracle_lib, False)
    a.run(25)
    assert a.get("nut").min() > 0.0
    b = _solver(oracle_lib, True)
    b.set_state(np.stack([a.get(f"U{k}") for k in range(5)]))
    for s in (a, b):
        s.timestep()
        s.residual_terms()
    for name in ["S"] + [f"{t}{k}" for t in ("Fc", "Fd") for k in range(5)]:
        assert np.array_equal(a.get(name), b.get(name)), name
    for k in range(5):
        fa, fb = a.get(f"Fv{k}").reshape(a.ni, a.nj), b.get(f"Fv{k}").reshape(b.ni, b.nj)
        assert np.array_equal(fa[1:], fb[1:]), k
    for k in (1, 2, 4):  # momentum and nu~ fluxes through the wall differ
        assert not np.array_equal(a.get(f"Fv{k}")[: a.nj], b.get(f"Fv{k}")[: b.nj]), k


def test_reference_wall_ghost_has_the_fv1_pole(oracle_lib):
    """nu~ > 0 with chi = cv1 everywhere (shear flow): the reference's wall ghost (-nu~) sits on
    the fv1 pole and its wall-face momentum flux is orders of magnitude off; with the switch the
    ghost has no eddy viscosity."""
    cfg = config_dict()
    out = {}
    for neg in (False, True):
        s = _solver(oracle_lib, neg)
        cy = s.get("cy")
        s.set_state(_state(s, 1.2, 60.0 + 30.0 * cy, 5.0, 101325.0, 0.0))
        s.residual_terms()
        nut = cfg["spalart_allmaras"]["Cv1"] * s.get("mu") / 1.2  # chi = cv1
        s.set_state(_state(s, 1.2, 60.0 + 30.0 * cy, 5.0, 101325.0, nut))
        s.residual_terms()
        assert np.allclose(s.get("chi"), cfg["spalart_allmaras"]["Cv1"], rtol=1e-12)
        out[neg] = s.get("Fv1").reshape(s.ni, s.nj)
    assert np.all(np.isfinite(out[True]))
    inner = np.abs(out[True][1:]).max()
    assert np.abs(out[True][0]).max() < 100.0 * inner
    assert not np.isfinite(np.abs(out[False][0]).max()) or np.abs(out[False][0]).max() > 1e6 * inner


def test_no_eddy_viscosity_where_nut_is_negative(oracle_lib):
    """Shear flow u = 60 + 30 y: with nu~ < 0 the momentum and energy viscous fluxes are
    bitwise those of nu~ = 0 (laminar viscosity only); the reference's formula differs."""
    runs = {}
    for key, neg, nut in (("neg", True, -2e-4), ("zero", True, 0.0), ("ref", False, -2e-4)):
        s = _solver(oracle_lib, neg)
        cy = s.get("cy")
        s.set_state(_state(s, 1.2, 60.0 + 30.0 * cy, 5.0, 101325.0, nut))
        s.residual_terms()
        runs[key] = [_interior(s.get(f"Fv{k}"), s) for k in (1, 2, 3)]
    for k in range(3):
        assert np.array_equal(runs["neg"][k], runs["zero"][k]), k
    scale = np.abs(runs["zero"][0]).max()
    assert np.abs(runs["ref"][0] - runs["zero"][0]).max() > 1e-2 * scale


def test_source_of_a_uniform_negative_state(oracle_lib):
    """Uniform flow, nu~ < 0: no vorticity and no nu~ gradient inside, so the source is the
    destruction term alone with its sign reversed, cw1 rho (nu~/d)^2 V."""
    nut, rho = -2e-4, 1.2
    s = _solver(oracle_lib, True)
    s.set_state(_state(s, rho, 60.0, 5.0, 101325.0, nut))
    s.residual_terms()
    d = oracle_lib.derived(s.cfg)
    want = d.cw1 * rho * nut * nut * s.get("inv_wd2") * s.get("vol")
    np.testing.assert_allclose(_interior(s.get("S"), s), _interior(want, s), rtol=1e-9)
    r = _solver(oracle_lib, False)
    r.set_state(_state(r, rho, 60.0, 5.0, 101325.0, nut))
    r.residual_terms()
    assert np.abs(_interior(r.get("S") - want, r)).max() > 1e-3 * np.abs(want).max()


def test_source_of_a_shear_flow_with_negative_nut(oracle_lib):
    """Production on the vorticity S = fv3 |dv/dx - du/dy| (not S~) with the factor
    cb1 (1 - ct3), from the exported gradients; destruction and cb2 terms as above."""
    nut, rho = -2e-4, 1.2
    s = _solver(oracle_lib, True)
    cx, cy = s.get("cx"), s.get("cy")
    s.set_state(_state(s, rho, 60.0 + 30.0 * cy, 5.0 - 10.0 * cx, 101325.0, nut))
    s.residual_terms()
    c, d = s.cfg, oracle_lib.derived(s.cfg)
    g = [s.get(f"grad{k}") for k in range(8)]
    vort = c.fv3 * np.abs(g[2] - g[1])
    prod = c.cb1 * (1.0 - c.ct3) * vort * rho * nut
    dest = d.cw1 * rho * nut * nut * s.get("inv_wd2")
    diff = c.cb2 * c.sigma * rho * (g[6] ** 2 + g[7] ** 2)
    want = (prod + dest + diff) * s.get("vol")
    got = s.get("S")
    assert np.abs(_interior(vort, s)).min() > 10.0  # the shear is resolved
    scale = np.abs(_interior(prod * s.get("vol"), s)).max()
    assert np.abs(_interior(got - want, s)).max() <= 1e-10 * scale


def _assemble_nut_flux(s, coef):
    """Fv4 = tau(i+1) - tau(i) + n(j+1) - n(j) of the face averages of coef * grad nu~."""
    ni, nj = s.ni, s.nj
    dx = (coef * s.get("grad6")).reshape(ni, nj)
    dy = (coef * s.get("grad7")).reshape(ni, nj)
    tnx, tny = s.get("tau_nx").reshape(ni + 1, nj), s.get("tau_ny").reshape(ni + 1, nj)
    nnx, nny = s.get("n_nx").reshape(ni, nj), s.get("n_ny").reshape(ni, nj)
    # tau face i (1..ni-1) between cells i-1 and i; n face j between cells j-1 and j
    tau = (dx[:-1] + dx[1:]) * (0.5 * tnx[1:ni]) + (dy[:-1] + dy[1:]) * (0.5 * tny[1:ni])
    nf = ((np.roll(dx, 1, axis=1) + dx) * (0.5 * nnx) + (np.roll(dy, 1, axis=1) + dy) * (0.5 * nny))
    return tau[1:] - tau[:-1] + np.roll(nf, -1, axis=1)[1:-1] - nf[1:-1]


def test_nut_diffusion_coefficient(oracle_lib):
    """nu~ = -(1e-5 + 1e-4 (x + 5)) < 0 (chi from about -1 to -15): the nu~ viscous flux is
    the assembly of sigma mu (1 + chi fn(chi)) grad nu~, fn = (16 + chi^3) / (16 - chi^3);
    the reference's coefficient sigma mu (1 + chi) gives another flux."""
    s = _solver(oracle_lib, True)
    cx = s.get("cx")
    s.set_state(_state(s, 1.2, 60.0, 5.0, 101325.0, -(1e-5 + 1e-4 * (cx + 5.0))))
    s.residual_terms()
    mu, chi = s.get("mu"), s.get("chi")
    assert chi.max() < 0.0 and chi.min() < -10.0
    c3 = chi ** 3
    fn = (CN1 + c3) / (CN1 - c3)
    want = _assemble_nut_flux(s, s.cfg.sigma * mu * (1.0 + chi * fn))
    got = _interior(s.get("Fv4"), s)
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 1e-11 * scale
    reference = _assemble_nut_flux(s, s.cfg.sigma * mu * (1.0 + chi))
    assert np.abs(reference - want).max() > 1e-2 * scale
