"""The benchmark / parity workloads, as (config JSON text, mesh text) pairs.

Base config is the reference ``config.json`` (physics, SA constants); each workload
overrides the simulation inputs and uses the default-off extensions where the configuration
asks for something the reference has no input for (Reynolds number, nu~/nu ratio, local
time stepping, relative residual drop).
"""
from __future__ import annotations

import copy
import json
import math
import random
from pathlib import Path

from . import meshgen

BASE_CONFIG = {
    "physics": {"gamma": 1.4, "R": 287.06, "T0": 288.16, "Ts": 110, "mu0": 1.7894e-05,
                "P0": 101325.0, "c0": 340.28},
    "simulation": {"AOA": 0, "Ma": 0.2, "CFL": 0.8, "IM": 3, "T": 300, "P": 101325.0},
    "solver": {"iteration": 10000, "targetres": 1e-10},
    "spalart_allmaras": {"Cv1": 7.1, "Pr": 0.71, "Prt": 0.9, "sigma": 1.5, "Cb1": 0.1355,
                         "Cb2": 0.622, "Ct3": 1.2, "Ct4": 0.5, "Cw2": 0.3, "Cw3": 2.0,
                         "fv3": 1.0, "kappa": 0.41, "rmax": 10},
    "dissipation": {"k2": 0.5, "k4": 0.0078125},
}


def make_config(*, mach: float, aoa: float, cfl: float = 2.0, k4: float = 1.0 / 64.0,
                reynolds: float | None = None, nut_ratio: float = 3.0, lts: bool = False,
                viscous_dt: bool = False, sa_negative: bool = False, iteration: int = 10000,
                targetres: float = 0.0, residual_drop: float | None = None) -> str:
    c = copy.deepcopy(BASE_CONFIG)
    c["simulation"].update({"AOA": aoa, "Ma": mach, "CFL": cfl})
    if reynolds:
        c["simulation"]["Re"] = reynolds
    if lts:
        c["simulation"]["local_time_stepping"] = True
    if viscous_dt:
        c["simulation"]["viscous_time_step"] = True
    c["dissipation"]["k4"] = k4
    c["spalart_allmaras"]["nut_ratio"] = nut_ratio
    if sa_negative:
        c["spalart_allmaras"]["negative"] = True
    c["solver"] = {"iteration": iteration, "targetres": targetres}
    if residual_drop:
        c["solver"]["residual_drop"] = residual_drop
    return json.dumps(c, indent=1)


def c1_cylinder() -> dict:
    """Plain FP64 JST + SA on a circular cylinder: 40 x 17 points (r 0.5 -> 20, geometric),
    M 0.2, Re 1e3 per diameter, nu~/nu = 3, CFL 2.0, k4 = 1/64, 2000 iterations,
    global-minimum dt (the reference's time stepping)."""
    mesh, q = meshgen.cylinder_mesh(40, 17, 0.5, 20.0)
    cfg = make_config(mach=0.2, aoa=0.0, cfl=2.0, reynolds=1e3, nut_ratio=3.0, iteration=2000)
    return {"name": "cylinder_40x17", "config": cfg, "mesh": mesh, "stretch": q, "steps": 2000}


def c2_naca0012() -> dict:
    """NACA 0012 O-mesh 257 x 129 (closed ring, first spacing 1e-5 c, far field 30 c),
    M 0.75, AoA 2 deg, Re 6e6, SA nu~/nu = 3 with the negative-nu~ branch, local time stepping
    at CFL 2, up to 20 000 iterations or an 8-decade drop.  (With the reference's global time
    step the 1e-5 c wall cells set a step that moves the far field by 2.6 decades in 20 000
    iterations; with local steps the reference's SA model meets the fv1 pole where nu~ turns
    negative near the shock foot and diverges.)"""
    mesh, q = meshgen.naca_mesh(256, 129, 0.0, 0.4, 0.12, 1e-5, 30.0, closed=True)
    cfg = make_config(mach=0.75, aoa=2.0, cfl=2.0, reynolds=6e6, nut_ratio=3.0, lts=True,
                      sa_negative=True, iteration=20000, residual_drop=8.0)
    return {"name": "naca0012_257x129", "config": cfg, "mesh": mesh, "stretch": q,
            "steps": 20000}


def c3_naca_large() -> dict:
    """NACA 0012 O-mesh 1025 x 513 (first spacing 2e-6 c), M 0.8, AoA 1.5 deg, Re 1e7, CFL 1.0
    with the reference's global time step and SA model (a fixed number of steps for the kernel
    measurements)."""
    mesh, q = meshgen.naca_mesh(1024, 513, 0.0, 0.4, 0.12, 2e-6, 30.0, closed=True)
    cfg = make_config(mach=0.8, aoa=1.5, cfl=1.0, reynolds=1e7, nut_ratio=3.0, lts=False,
                      iteration=20000, residual_drop=8.0)
    return {"name": "naca0012_1025x513", "config": cfg, "mesh": mesh, "stretch": q, "steps": 20}


C4_SIZES = [(129, 65), (193, 97), (257, 129), (385, 193), (513, 257)]


def c4_batch(n: int = 64, seed: int = 20261203, first_spacing: float = 1e-3) -> list[dict]:
    """Uneven batch from a seed: NACA 4-digit t ~ U[8%,18%], m ~ U[0,4%] (p = 0.4),
    O-meshes drawn from C4_SIZES (closed rings), M ~ U[0.3,0.85], AoA ~ U[-4,8] deg,
    Re ~ logU[1e5,1e7], SA nu~/nu = 3 with the negative-nu~ branch, CFL 1.0 with the
    reference's global time step.

    With the reference's SA model half of the cases become non-physical within 3000 steps
    (the first at step 310), where a wall cell's nu~ reaches chi = +cv1 and its mirror ghost
    sits on the fv1 pole, or nu~ turns negative through chi = -cv1; with the negative-nu~
    branch none does (GPU, 3000 steps).

    First wall spacing 1e-3 chord: the reference scheme sizes its time step from the
    convective spectral radius only, so finer wall spacing at the low-Re end of this
    distribution violates the explicit viscous limit (checked with the CPU oracle on the
    Re = 1e5 cases: at 1e-5 chord they fail in the first step, at 1e-4 chord with local time
    steps at CFL 2 in the second)."""
    rng = random.Random(seed)
    out = []
    for k in range(n):
        t = rng.uniform(0.08, 0.18)
        m = rng.uniform(0.0, 0.04)
        pts, rings = C4_SIZES[rng.randrange(len(C4_SIZES))]
        mach = rng.uniform(0.3, 0.85)
        aoa = rng.uniform(-4.0, 8.0)
        re = 10.0 ** rng.uniform(5.0, 7.0)
        mesh, q = meshgen.naca_mesh(pts - 1, rings, m, 0.4, t, first_spacing, 30.0, closed=True)
        cfg = make_config(mach=mach, aoa=aoa, cfl=1.0, reynolds=re, nut_ratio=3.0, lts=False,
                          sa_negative=True, iteration=20000, residual_drop=6.0)
        out.append({"name": f"c4_{k:02d}_naca{m*100:.1f}-{t*100:.1f}_{pts}x{rings}",
                    "config": cfg, "mesh": mesh, "stretch": q, "ni": rings - 1, "nj": pts - 1,
                    "mach": mach, "aoa": aoa, "re": re, "m": m, "t": t})
    return out


def converging_config(w: dict, max_steps: int = 20000) -> str:
    """A C4 draw's geometry and flow (Mach, incidence, Reynolds number) with the solver
    settings that converge it: local time stepping at CFL 2 with the viscous spectral radius in
    the local step, the negative-nu~ SA branch, and a 6-decade residual drop within max_steps
    (the serving stream's NACA requests).  Over the 64 cases of the C4 batch this converges 45
    within 20 000 steps and none becomes non-physical (without the viscous radius: 42 and 5;
    at CFL 2.5: 28, at CFL 3: 5)."""
    return make_config(mach=w["mach"], aoa=w["aoa"], cfl=2.0, reynolds=w["re"], nut_ratio=3.0,
                       lts=True, viscous_dt=True, sa_negative=True, iteration=max_steps,
                       residual_drop=6.0)


def golden_config_text() -> str:
    return (Path(__file__).resolve().parent.parent / "tests" / "golden" / "config.json").read_text()


CYLINDER_SIZES = [(40, 17), (80, 33), (160, 65)]


def c5_requests(seconds: float = 60.0, rate: float = 20.0, seed: int = 20261204,
                cylinder_fraction: float = 0.5) -> list[dict]:
    """Serving stream: Poisson arrivals (mean `rate` solves/s) over `seconds`; each request is
    a cylinder case (40x17, 80x33 or 160x65 points, M 0.2, Re 1e3, CFL 2 with the reference's
    global time step, converged at a 6-decade residual drop, which takes about 2900, 6100 and
    10000 steps, at most 12000 steps) with probability `cylinder_fraction`, else a draw from
    the C4 distribution (geometry, Mach, incidence, Reynolds number) solved to a 6-decade drop
    with `converging_config`, at most 20000 steps.

    Those settings converge most of the distribution in 5 000 - 20 000 steps; draws at transonic
    Mach numbers with strong incidence (shock-induced separation, no steady solution) stop at
    the step cap, and are reported as such."""
    rng = random.Random(seed)
    pool = c4_batch(n=64, seed=seed)
    out, t = [], 0.0
    while True:
        t += rng.expovariate(rate)
        if t > seconds:
            break
        if rng.random() < cylinder_fraction:
            pts, rings = CYLINDER_SIZES[rng.randrange(len(CYLINDER_SIZES))]
            mesh, _ = meshgen.cylinder_mesh(pts, rings, 0.5, 20.0)
            cfg = make_config(mach=0.2, aoa=0.0, cfl=2.0, reynolds=1e3, nut_ratio=3.0,
                              iteration=12000, residual_drop=6.0)
            out.append({"arrival_s": t, "config": cfg, "mesh": mesh, "max_steps": 12000,
                        "name": f"cyl_{pts}x{rings}"})
        else:
            w = pool[rng.randrange(len(pool))]
            out.append({"arrival_s": t, "config": converging_config(w), "mesh": w["mesh"],
                        "max_steps": 20000, "name": w["name"]})
    return out
