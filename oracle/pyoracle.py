"""ctypes binding to the C oracle ``oracle/_build/libjstoracle.so`` (test infrastructure).

The oracle restates jst-rs (src/*.rs) in plain C; see ``oracle/jst_oracle.h``.  Config
parsing here follows jst-rs config.rs:100-206 (serde field names, optional ``solver``
section with defaults 10000 / 1e-10); the default-off extension keys are read from the same
JSON (``simulation.Re``, ``simulation.ref_length``, ``simulation.local_time_stepping``,
``simulation.viscous_time_step``, ``spalart_allmaras.nut_ratio``,
``spalart_allmaras.negative``, ``solver.residual_drop``).
"""
from __future__ import annotations

import ctypes as C
import json
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libjstoracle.so"


class JoConfig(C.Structure):
    _fields_ = [
        ("gamma", C.c_double), ("r_gas", C.c_double), ("t_ref", C.c_double),
        ("t_suth", C.c_double), ("mu0", C.c_double),
        ("aoa_deg", C.c_double), ("mach", C.c_double), ("cfl", C.c_double),
        ("halo", C.c_int), ("t_inf", C.c_double), ("p_inf", C.c_double),
        ("iteration", C.c_long), ("targetres", C.c_double),
        ("cv1", C.c_double), ("pr", C.c_double), ("prt", C.c_double), ("sigma", C.c_double),
        ("cb1", C.c_double), ("cb2", C.c_double), ("ct3", C.c_double), ("ct4", C.c_double),
        ("cw2", C.c_double), ("cw3", C.c_double), ("fv3", C.c_double), ("kappa", C.c_double),
        ("rmax", C.c_double),
        ("k2", C.c_double), ("k4", C.c_double),
        ("nut_ratio", C.c_double), ("reynolds", C.c_double), ("ref_length", C.c_double),
        ("local_time_stepping", C.c_int), ("residual_drop", C.c_double),
        ("viscous_time_step", C.c_int), ("sa_negative", C.c_int),
    ]


class JoDerived(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "cv", "cp", "cw1", "cv1_cubed", "c_inf", "u_inf", "v_inf", "rho_inf", "mu_inf",
        "nut_inf")] + [("supersonic", C.c_int)] + [(n, C.c_double) for n in (
        "suth_num", "inv_t_ref", "inv_pr", "inv_prt", "inv_kappa2", "cw3_6", "one_plus_cw3_6",
        "mu0_eff")]


class JoMesh(C.Structure):
    _fields_ = [("n_rings", C.c_int), ("n_theta", C.c_int), ("xy", C.POINTER(C.c_double))]


def build() -> None:
    """Compile the oracle (make -C oracle)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        L = C.CDLL(str(LIB_PATH))
        vp, dp = C.c_void_p, C.POINTER(C.c_double)
        L.jo_solver_new.restype = vp
        L.jo_solver_new.argtypes = [C.POINTER(JoConfig), C.POINTER(JoMesh)]
        L.jo_solver_free.argtypes = [vp]
        for n in ("jo_ni", "jo_nj", "jo_halo"):
            getattr(L, n).argtypes = [vp]
            getattr(L, n).restype = C.c_int
        L.jo_advance.argtypes = [vp, dp]
        L.jo_advance.restype = C.c_int
        L.jo_residual.argtypes = [vp]
        L.jo_residual.restype = C.c_double
        L.jo_run.argtypes = [vp, C.c_long, dp, dp, dp]
        L.jo_run.restype = C.c_int
        L.jo_timestep.argtypes = [vp]
        L.jo_timestep.restype = C.c_double
        for n in ("jo_residual_terms", "jo_boundary_apply", "jo_convection", "jo_gradient",
                  "jo_viscous", "jo_dissipation", "jo_source"):
            getattr(L, n).argtypes = [vp]
        L.jo_totaltime.argtypes = [vp]
        L.jo_totaltime.restype = C.c_double
        L.jo_step.argtypes = [vp]
        L.jo_step.restype = C.c_long
        L.jo_set_uniform.argtypes = [vp] + [C.c_double] * 5
        L.jo_get.argtypes = [vp, C.c_char_p, dp]
        L.jo_get.restype = C.c_long
        L.jo_set_state.argtypes = [vp, dp, dp]
        L.jo_set_state.restype = C.c_int
        L.jo_forces.argtypes = [vp, dp]
        L.jo_wall.argtypes = [vp, dp, dp, dp, dp]
        L.jo_mesh_parse.argtypes = [C.c_char_p, C.POINTER(JoMesh), C.POINTER(C.c_long),
                                    C.POINTER(C.c_long)]
        L.jo_mesh_parse.restype = C.c_int
        L.jo_mesh_free.argtypes = [C.POINTER(JoMesh)]
        L.jo_config_finish.argtypes = [C.POINTER(JoConfig), C.POINTER(JoDerived)]
        L.jo_mu.argtypes = [C.POINTER(JoConfig), C.POINTER(JoDerived), C.c_double]
        L.jo_mu.restype = C.c_double
        L.jo_threads.restype = C.c_int
        L.jo_set_threads.argtypes = [C.c_int]
        L.jo_begin_step.argtypes = [vp]
        L.jo_begin_step.restype = C.c_double
        L.jo_stage.argtypes = [vp, C.c_double, C.c_double, dp]
        L.jo_stage.restype = C.c_int
        _lib = L
    return _lib


def config_from_dict(doc: dict) -> JoConfig:
    """jst-rs config.rs:100-206 field mapping, plus the default-off extension keys."""
    ph, sim, sa, dis = doc["physics"], doc["simulation"], doc["spalart_allmaras"], doc["dissipation"]
    sol = doc.get("solver", {})
    c = JoConfig()
    c.gamma, c.r_gas, c.t_ref, c.t_suth, c.mu0 = (
        ph["gamma"], ph["R"], ph["T0"], ph["Ts"], ph["mu0"])
    c.aoa_deg, c.mach, c.cfl, c.halo = sim["AOA"], sim["Ma"], sim["CFL"], int(sim["IM"])
    c.t_inf, c.p_inf = sim["T"], sim["P"]
    c.iteration = int(sol.get("iteration", 10000))
    c.targetres = float(sol.get("targetres", 1e-10))
    for k, n in (("cv1", "Cv1"), ("pr", "Pr"), ("prt", "Prt"), ("sigma", "sigma"),
                 ("cb1", "Cb1"), ("cb2", "Cb2"), ("ct3", "Ct3"), ("ct4", "Ct4"),
                 ("cw2", "Cw2"), ("cw3", "Cw3"), ("fv3", "fv3"), ("kappa", "kappa"),
                 ("rmax", "rmax")):
        setattr(c, k, float(sa[n]))
    c.k2, c.k4 = dis["k2"], dis["k4"]
    c.nut_ratio = float(sa.get("nut_ratio", 0.1))
    c.reynolds = float(sim.get("Re", 0.0))
    c.ref_length = float(sim.get("ref_length", 1.0))
    c.local_time_stepping = 1 if sim.get("local_time_stepping", False) else 0
    c.residual_drop = float(sol.get("residual_drop", 0.0))
    c.viscous_time_step = 1 if sim.get("viscous_time_step", False) else 0
    c.sa_negative = 1 if sa.get("negative", False) else 0
    return c


def derived(cfg: JoConfig) -> JoDerived:
    d = JoDerived()
    lib().jo_config_finish(C.byref(cfg), C.byref(d))
    return d


class MeshError(Exception):
    def __init__(self, code: int, a: int, b: int):
        super().__init__(f"mesh error {code} ({a}, {b})")
        self.code, self.a, self.b = code, a, b


def parse_mesh(text: str):
    """Returns (n_rings, n_theta, xy ndarray[n_rings*n_theta, 2]) or raises MeshError."""
    m = JoMesh()
    a, b = C.c_long(), C.c_long()
    st = lib().jo_mesh_parse(text.encode(), C.byref(m), C.byref(a), C.byref(b))
    if st != 0:
        raise MeshError(st, a.value, b.value)
    n = m.n_rings * m.n_theta
    xy = np.ctypeslib.as_array(m.xy, shape=(n * 2,)).copy().reshape(n, 2)
    lib().jo_mesh_free(C.byref(m))
    return m.n_rings, m.n_theta, xy


class NonPhysical(Exception):
    def __init__(self, err):
        i, j, rho, p = err
        super().__init__(f"non-physical state at cell ({int(i)}, {int(j)}): rho = {rho:e}, p = {p:e}")
        self.i, self.j, self.rho, self.p = int(i), int(j), rho, p


class OracleSolver:
    """Mirror of jst-rs Solver (solver.rs:25-256) over the C oracle."""

    def __init__(self, cfg: dict | JoConfig, mesh_text: str):
        L = lib()
        self.cfg = cfg if isinstance(cfg, JoConfig) else config_from_dict(cfg)
        m = JoMesh()
        a, b = C.c_long(), C.c_long()
        st = L.jo_mesh_parse(mesh_text.encode(), C.byref(m), C.byref(a), C.byref(b))
        if st != 0:
            raise MeshError(st, a.value, b.value)
        self._h = L.jo_solver_new(C.byref(self.cfg), C.byref(m))
        L.jo_mesh_free(C.byref(m))
        self.ni, self.nj = L.jo_ni(self._h), L.jo_nj(self._h)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().jo_solver_free(self._h)
            self._h = None

    @property
    def n_cells(self) -> int:
        return self.ni * self.nj

    def advance(self) -> None:
        err = (C.c_double * 4)()
        if lib().jo_advance(self._h, err) != 0:
            raise NonPhysical(list(err))

    def residual(self) -> float:
        return lib().jo_residual(self._h)

    def run(self, max_steps: int | None = None):
        n = max_steps if max_steps else self.cfg.iteration
        hist = np.zeros(n)
        out = (C.c_double * 4)()
        err = (C.c_double * 4)()
        st = lib().jo_run(self._h, n, hist.ctypes.data_as(C.POINTER(C.c_double)), out, err)
        if st != 0:
            raise NonPhysical(list(err))
        steps = int(out[0])
        return {"steps": steps, "residual": out[1], "totaltime": out[2],
                "converged": bool(out[3]), "history": hist[:steps]}

    def timestep(self) -> float:
        return lib().jo_timestep(self._h)

    def begin_step(self) -> float:
        """Step-start snapshot and frozen dt (the first half of advance())."""
        return lib().jo_begin_step(self._h)

    def stage(self, alpha: float, dt: float) -> None:
        """One RK stage from the snapshot of begin_step()."""
        err = (C.c_double * 4)()
        if lib().jo_stage(self._h, alpha, dt, err) != 0:
            raise NonPhysical(list(err))

    def residual_terms(self) -> None:
        lib().jo_residual_terms(self._h)

    def call(self, kernel: str) -> None:
        getattr(lib(), "jo_" + kernel)(self._h)

    @property
    def totaltime(self) -> float:
        return lib().jo_totaltime(self._h)

    @property
    def step(self) -> int:
        return lib().jo_step(self._h)

    def set_uniform(self, rho, vx, vy, p, nut) -> None:
        lib().jo_set_uniform(self._h, rho, vx, vy, p, nut)

    def get(self, name: str) -> np.ndarray:
        n = (self.ni + 1) * self.nj * 1
        buf = np.zeros(n)
        k = lib().jo_get(self._h, name.encode(), buf.ctypes.data_as(C.POINTER(C.c_double)))
        if k < 0:
            raise KeyError(name)
        return buf[:k].copy()

    def set_state(self, u_planes: np.ndarray) -> None:
        u = np.ascontiguousarray(u_planes, dtype=np.float64)
        err = (C.c_double * 4)()
        if lib().jo_set_state(self._h, u.ctypes.data_as(C.POINTER(C.c_double)), err) != 0:
            raise NonPhysical(list(err))

    def forces(self):
        out = (C.c_double * 4)()
        lib().jo_forces(self._h, out)
        return {"fx": out[0], "fy": out[1], "cl": out[2], "cd": out[3]}

    def wall(self) -> dict:
        a = {k: np.zeros(self.nj) for k in ("x", "y", "cp", "cf")}
        ptr = [a[k].ctypes.data_as(C.POINTER(C.c_double)) for k in ("x", "y", "cp", "cf")]
        lib().jo_wall(self._h, *ptr)
        return a


def load_config(path: str | os.PathLike) -> dict:
    return json.loads(Path(path).read_text())
