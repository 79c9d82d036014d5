"""Python mirror of the jst-rs crate surface over the gfx950 C ABI (``include/jst_gpu.h``).

``Config``, ``Mesh`` and ``Solver`` keep the reference's names, argument meaning and error
behaviour (jst-rs config.rs, mesh.rs, solver.rs): mesh errors raise ``MeshError`` carrying
the MeshError variant, a diverged state raises ``NonPhysical`` (never a silent NaN), and
``Solver.run`` returns a ``RunReport``.  All numerics run in ``libjstgpu.so``; there is no
CPU fallback -- loading fails loudly if the library has not been built.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "_build" / "libjstgpu.so"

JST_OK, JST_ERR_IO, JST_ERR_BAD_HEADER, JST_ERR_EMPTY = 0, 1, 2, 3
JST_ERR_COUNT_MISMATCH, JST_ERR_BAD_POINT, JST_ERR_TOO_SMALL = 4, 5, 6
JST_ERR_NONPHYSICAL, JST_ERR_CONFIG, JST_ERR_HIP, JST_ERR_ARG = 7, 8, 9, 10
FP64, MIXED = 0, 1

MESH_ERRORS = {JST_ERR_IO: "Io", JST_ERR_BAD_HEADER: "BadHeader", JST_ERR_EMPTY: "Empty",
               JST_ERR_COUNT_MISMATCH: "CountMismatch", JST_ERR_BAD_POINT: "BadPoint",
               JST_ERR_TOO_SMALL: "TooSmall"}

# Exported symbols, in the order include/jst_gpu.h declares them.
EXPORTS = ["jst_last_error", "jst_config_check", "jst_mesh_parse", "jst_create",
           "jst_create_from_paths", "jst_destroy", "jst_dims", "jst_advance", "jst_run",
           "jst_residual", "jst_residual_terms", "jst_get_field", "jst_set_state", "jst_forces",
           "jst_wall", "jst_write_wall_csv", "jst_write_csv", "jst_progress", "jst_time_steps",
           "jst_time_begin", "jst_time_run", "jst_time_end",
           "jst_batch_create", "jst_batch_destroy", "jst_case_footprint", "jst_batch_add",
           "jst_batch_remove", "jst_batch_step", "jst_batch_info", "jst_batch_forces", "jst_batch_set_state",
           "jst_batch_wall", "jst_batch_get_field", "jst_batch_residual_terms",
           "jst_batch_residual_terms_noalloc", "jst_batch_history",
           "jst_batch_set_tiling", "jst_batch_segment_rows", "jst_batch_tune", "jst_batch_time_begin", "jst_batch_time_end",
           "jst_batch_time_steps", "jst_batch_time_kernels",
           "jst_batch_counts",
           "jst_copy_probe", "jst_serve"]
JST_ERR_NOSPACE = 11
CASE_RUNNING, CASE_CONVERGED, CASE_STOPPED, CASE_NONPHYSICAL = 0, 1, 2, 3


class JstError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(msg)
        self.status = status


class MeshError(JstError):
    """jst-rs mesh.rs:18-25 MeshError; ``kind`` is the variant name."""

    @property
    def kind(self) -> str:
        return MESH_ERRORS.get(self.status, "?")


class ConfigError(JstError):
    pass


class NonPhysical(JstError):
    """jst-rs state.rs:241-259 NonPhysical {i, j, rho, p}."""

    def __init__(self, msg: str, rec):
        super().__init__(JST_ERR_NONPHYSICAL, msg)
        self.i, self.j, self.rho, self.p = int(rec.i), int(rec.j), rec.rho, rec.p


class _NonPhys(C.Structure):
    _fields_ = [("i", C.c_long), ("j", C.c_long), ("rho", C.c_double), ("p", C.c_double)]


class _Report(C.Structure):
    _fields_ = [("steps", C.c_long), ("residual", C.c_double), ("totaltime", C.c_double),
                ("converged", C.c_int)]


class _CaseInfo(C.Structure):
    _fields_ = [("step", C.c_long), ("status", C.c_int), ("residual", C.c_double),
                ("res_first", C.c_double), ("totaltime", C.c_double), ("cells", C.c_long)]


class _Request(C.Structure):
    _fields_ = [("arrival_s", C.c_double), ("config_json", C.c_char_p),
                ("mesh_text", C.c_char_p), ("max_steps", C.c_long)]


class _RequestResult(C.Structure):
    _fields_ = [("admit_s", C.c_double), ("done_s", C.c_double), ("steps", C.c_long),
                ("status", C.c_int), ("residual", C.c_double), ("cl", C.c_double),
                ("cd", C.c_double)]


class _ServeOpts(C.Structure):
    _fields_ = [("hbm_bytes", C.c_size_t), ("max_cases", C.c_int),
                ("steps_per_round", C.c_int), ("tune", C.c_int), ("segment_rows", C.c_int)]


class _ServeSummary(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("cell_stages", C.c_double), ("p50_s", C.c_double),
                ("p99_s", C.c_double), ("completed", C.c_int), ("nonphysical", C.c_int),
                ("max_concurrent", C.c_int), ("tiling_switches", C.c_int),
                ("converged", C.c_int), ("rejected", C.c_int),
                ("p50_converged_s", C.c_double), ("p99_converged_s", C.c_double)]


STEP_CB = C.CFUNCTYPE(None, C.c_long, C.c_double, C.c_void_p)

_lib = None


def lib():
    """Load libjstgpu.so (built by ``__graft_entry__.build()``); never falls back."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built: run __graft_entry__.build()")
        L = C.CDLL(str(LIB_PATH))
        vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.c_int
        L.jst_last_error.restype = C.c_char_p
        L.jst_config_check.argtypes = [C.c_char_p, dp]
        L.jst_mesh_parse.argtypes = [C.c_char_p, C.POINTER(ip), C.POINTER(ip), dp, C.c_size_t]
        L.jst_create.argtypes = [C.c_char_p, C.c_char_p, ip, C.POINTER(vp)]
        L.jst_create_from_paths.argtypes = [C.c_char_p, C.c_char_p, ip, C.POINTER(vp)]
        L.jst_destroy.argtypes = [vp]
        L.jst_destroy.restype = None
        L.jst_dims.argtypes = [vp, C.POINTER(ip), C.POINTER(ip)]
        L.jst_advance.argtypes = [vp, ip, dp, C.POINTER(_NonPhys)]
        L.jst_run.argtypes = [vp, C.c_long, STEP_CB, vp, C.POINTER(_Report), C.POINTER(_NonPhys)]
        L.jst_residual.argtypes = [vp, dp]
        L.jst_residual_terms.argtypes = [vp, dp]
        L.jst_get_field.argtypes = [vp, C.c_char_p, dp, C.c_size_t]
        L.jst_set_state.argtypes = [vp, dp, C.POINTER(_NonPhys)]
        L.jst_forces.argtypes = [vp, dp, dp, dp, dp]
        L.jst_write_csv.argtypes = [vp, C.c_char_p]
        L.jst_wall.argtypes = [vp, dp, dp, dp, dp]
        L.jst_write_wall_csv.argtypes = [vp, C.c_char_p]
        L.jst_progress.argtypes = [vp, C.POINTER(C.c_long), dp]
        L.jst_time_steps.argtypes = [vp, ip, dp, dp]
        L.jst_time_begin.argtypes = [vp, ip]
        L.jst_time_run.argtypes = [vp, ip]
        L.jst_time_end.argtypes = [vp]
        sz = C.c_size_t
        L.jst_batch_create.argtypes = [ip, sz, ip, C.POINTER(vp)]
        L.jst_batch_destroy.argtypes = [vp]
        L.jst_batch_destroy.restype = None
        L.jst_case_footprint.argtypes = [C.c_char_p, C.c_char_p, ip, C.POINTER(sz)]
        L.jst_batch_add.argtypes = [vp, C.c_char_p, C.c_char_p, C.c_long, C.POINTER(ip)]
        L.jst_batch_remove.argtypes = [vp, ip]
        L.jst_batch_step.argtypes = [vp, ip, ip]
        L.jst_batch_info.argtypes = [vp, ip, C.POINTER(_CaseInfo)]
        L.jst_batch_forces.argtypes = [vp, ip, dp, dp]
        L.jst_batch_set_state.argtypes = [vp, ip, dp, C.POINTER(_NonPhys)]
        L.jst_batch_wall.argtypes = [vp, ip, dp, dp, dp, dp]
        L.jst_batch_get_field.argtypes = [vp, ip, C.c_char_p, dp, sz]
        L.jst_batch_residual_terms.argtypes = [vp, ip, dp]
        L.jst_batch_residual_terms_noalloc.argtypes = [vp, ip, dp]
        L.jst_batch_history.argtypes = [vp, ip, C.c_long, C.c_long, dp]
        L.jst_batch_set_tiling.argtypes = [vp, ip, ip]
        L.jst_batch_tune.argtypes = [vp, ip, C.POINTER(ip), C.POINTER(ip)]
        L.jst_batch_segment_rows.argtypes = [vp, ip, C.POINTER(ip)]
        L.jst_batch_time_begin.argtypes = [vp, ip]
        L.jst_batch_time_end.argtypes = [vp]
        L.jst_batch_time_steps.argtypes = [vp, ip, dp, dp]
        L.jst_batch_time_kernels.argtypes = [vp, ip, dp, dp, dp]
        L.jst_batch_counts.argtypes = [vp, C.POINTER(ip), C.POINTER(C.c_long)]
        L.jst_copy_probe.argtypes = [sz, dp]
        L.jst_serve.argtypes = [C.POINTER(_Request), ip, C.POINTER(_ServeOpts),
                                C.POINTER(_RequestResult), C.POINTER(_ServeSummary)]
        for name in EXPORTS[1:]:
            if name not in ("jst_destroy", "jst_batch_destroy"):
                getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def _err() -> str:
    return lib().jst_last_error().decode()


def _check(st: int) -> None:
    if st == JST_OK:
        return
    msg = _err()
    if st in MESH_ERRORS:
        raise MeshError(st, msg)
    if st == JST_ERR_CONFIG:
        raise ConfigError(st, msg)
    raise JstError(st, msg)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


DERIVED_KEYS = ["cv", "cp", "cw1", "cv1_cubed", "c_inf", "u_inf", "v_inf", "rho_inf", "mu_inf",
                "nut_inf", "supersonic", "suth_num", "inv_t_ref", "inv_pr", "inv_prt",
                "inv_kappa2", "cw3_6", "one_plus_cw3_6", "mu0_eff", "halo"]


@dataclass
class Config:
    """jst-rs config.rs Config: the JSON text plus the derived constants (Config::finish)."""
    text: str
    derived: dict

    @classmethod
    def from_str(cls, text: str) -> "Config":
        out = np.zeros(len(DERIVED_KEYS))
        _check(lib().jst_config_check(text.encode(), _dp(out)))
        return cls(text, dict(zip(DERIVED_KEYS, out.tolist())))

    @classmethod
    def from_path(cls, path: str | os.PathLike) -> "Config":
        try:
            text = Path(path).read_text()
        except OSError as e:
            raise JstError(JST_ERR_IO, f"I/O error: {e}") from e
        return cls.from_str(text)


@dataclass
class Mesh:
    """jst-rs mesh.rs Mesh: nodes ring-major, theta wraps (node(ring, theta))."""
    n_rings: int
    n_theta: int
    xy: np.ndarray
    text: str

    @classmethod
    def parse(cls, text: str) -> "Mesh":
        a, b = C.c_int(), C.c_int()
        _check(lib().jst_mesh_parse(text.encode(), C.byref(a), C.byref(b), None, 0))
        xy = np.zeros(2 * a.value * b.value)
        _check(lib().jst_mesh_parse(text.encode(), C.byref(a), C.byref(b), _dp(xy), xy.size))
        return cls(a.value, b.value, xy.reshape(-1, 2), text)

    @classmethod
    def from_path(cls, path: str | os.PathLike) -> "Mesh":
        try:
            text = Path(path).read_text()
        except OSError as e:
            raise MeshError(JST_ERR_IO, f"I/O error: {e}") from e
        return cls.parse(text)

    def ni(self) -> int:
        return self.n_rings - 1

    def nj(self) -> int:
        return self.n_theta

    def n_cells(self) -> int:
        return self.ni() * self.nj()

    def node(self, ring: int, theta: int):
        return tuple(self.xy[ring * self.n_theta + theta % self.n_theta])


@dataclass
class RunReport:
    steps: int
    residual: float
    totaltime: float
    converged: bool


class Solver:
    """jst-rs solver.rs Solver on the GPU: state resident in HBM, one HIP stream."""

    def __init__(self, cfg: Config | str, mesh: Mesh | str, precision: int = FP64):
        ctext = cfg.text if isinstance(cfg, Config) else cfg
        mtext = mesh.text if isinstance(mesh, Mesh) else mesh
        h = C.c_void_p()
        _check(lib().jst_create(ctext.encode(), mtext.encode(), precision, C.byref(h)))
        self._h = h
        a, b = C.c_int(), C.c_int()
        _check(lib().jst_dims(h, C.byref(a), C.byref(b)))
        self._ni, self._nj = a.value, b.value

    @classmethod
    def from_paths(cls, config: str, mesh: str, precision: int = FP64) -> "Solver":
        return cls(Config.from_path(config), Mesh.from_path(mesh), precision)

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().jst_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def ni(self) -> int:
        return self._ni

    def nj(self) -> int:
        return self._nj

    def n_cells(self) -> int:
        return self._ni * self._nj

    def _nonphys(self, st: int, rec: _NonPhys) -> None:
        if st == JST_ERR_NONPHYSICAL:
            raise NonPhysical(_err(), rec)
        _check(st)

    def advance(self, nsteps: int = 1) -> np.ndarray:
        """``nsteps`` x Solver::advance; returns each step's residual."""
        res = np.zeros(max(nsteps, 1))
        rec = _NonPhys()
        self._nonphys(lib().jst_advance(self._h, nsteps, _dp(res), C.byref(rec)), rec)
        return res[:nsteps]

    def run(self, max_steps: int | None = None, on_step=None) -> RunReport:
        rep, rec = _Report(), _NonPhys()
        if on_step is None:
            cb = STEP_CB()
        else:
            cb = STEP_CB(lambda k, r, _u: on_step(int(k), float(r)))
        st = lib().jst_run(self._h, int(max_steps or 0), cb, None, C.byref(rep), C.byref(rec))
        self._nonphys(st, rec)
        return RunReport(int(rep.steps), rep.residual, rep.totaltime, bool(rep.converged))

    def residual(self) -> float:
        out = C.c_double()
        _check(lib().jst_residual(self._h, C.byref(out)))
        return out.value

    def residual_terms(self) -> float:
        """timestep::compute + Solver::residual_terms; returns dt."""
        dt = C.c_double()
        _check(lib().jst_residual_terms(self._h, C.byref(dt)))
        return dt.value

    def get(self, name: str) -> np.ndarray:
        n = (self._ni + 1) * self._nj
        buf = np.zeros(n)
        _check(lib().jst_get_field(self._h, name.encode(), _dp(buf), n))
        rows = self._ni + 1 if name.startswith("tau_") else self._ni
        return buf[: rows * self._nj].copy()

    def set_state(self, u_planes: np.ndarray) -> None:
        u = np.ascontiguousarray(u_planes, dtype=np.float64)
        rec = _NonPhys()
        self._nonphys(lib().jst_set_state(self._h, _dp(u), C.byref(rec)), rec)

    def forces(self) -> dict:
        v = [C.c_double() for _ in range(4)]
        _check(lib().jst_forces(self._h, *[C.byref(x) for x in v]))
        return dict(zip(("fx", "fy", "cl", "cd"), (x.value for x in v)))

    def wall(self) -> dict:
        """Wall distribution per wall face j: midpoint x, y, Cp and Cf (jst_wall)."""
        a = {k: np.zeros(self._nj) for k in ("x", "y", "cp", "cf")}
        _check(lib().jst_wall(self._h, *(_dp(a[k]) for k in ("x", "y", "cp", "cf"))))
        return a

    def write_wall_csv(self, path: str | os.PathLike) -> None:
        _check(lib().jst_write_wall_csv(self._h, str(path).encode()))

    def write_csv(self, path: str | os.PathLike) -> None:
        _check(lib().jst_write_csv(self._h, str(path).encode()))

    @property
    def step(self) -> int:
        s, t = C.c_long(), C.c_double()
        _check(lib().jst_progress(self._h, C.byref(s), C.byref(t)))
        return s.value

    @property
    def totaltime(self) -> float:
        s, t = C.c_long(), C.c_double()
        _check(lib().jst_progress(self._h, C.byref(s), C.byref(t)))
        return t.value

    def time_begin(self, nsteps: int) -> None:
        """Untimed set-up of nsteps timed steps (stops off, graphs built and warmed)."""
        _check(lib().jst_time_begin(self._h, int(nsteps)))

    def time_run(self, nsteps: int) -> None:
        """Exactly nsteps steps, waited for (between time_begin and time_end)."""
        _check(lib().jst_time_run(self._h, int(nsteps)))

    def time_end(self) -> None:
        _check(lib().jst_time_end(self._h))

    def time_steps(self, nsteps: int, per_stage: bool = False):
        total, stage = C.c_double(), C.c_double()
        _check(lib().jst_time_steps(self._h, nsteps, C.byref(total),
                                    C.byref(stage) if per_stage else None))
        return total.value, (stage.value if per_stage else None)


class Batch:
    """Many cases resident in HBM, advanced together by one launch per RK stage (FP32-mixed).

    No reference counterpart: the caller of the same loop for serving."""

    def __init__(self, hbm_bytes: int = 8 << 30, max_cases: int = 256):
        h = C.c_void_p()
        _check(lib().jst_batch_create(MIXED, int(hbm_bytes), int(max_cases), C.byref(h)))
        self._h = h
        self._cells = {}

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().jst_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    @staticmethod
    def footprint(config: str, mesh: str) -> int:
        out = C.c_size_t()
        _check(lib().jst_case_footprint(config.encode(), mesh.encode(), MIXED, C.byref(out)))
        return out.value

    def add(self, config: str, mesh: str, max_steps: int = 0) -> int:
        cid = C.c_int()
        _check(lib().jst_batch_add(self._h, config.encode(), mesh.encode(), int(max_steps),
                                   C.byref(cid)))
        return cid.value

    def remove(self, cid: int) -> None:
        _check(lib().jst_batch_remove(self._h, cid))

    def step(self, nsteps: int, sync: bool = True) -> None:
        _check(lib().jst_batch_step(self._h, int(nsteps), 1 if sync else 0))

    def info(self, cid: int) -> dict:
        ci = _CaseInfo()
        _check(lib().jst_batch_info(self._h, cid, C.byref(ci)))
        return {"step": ci.step, "status": ci.status, "residual": ci.residual,
                "res_first": ci.res_first, "totaltime": ci.totaltime, "cells": ci.cells}

    def run(self, max_rounds: int = 10**9, steps_per_round: int = 100) -> None:
        """Step until no case is running."""
        for _ in range(max_rounds):
            self.step(steps_per_round)
            n = C.c_int()
            _check(lib().jst_batch_counts(self._h, C.byref(n), None))
            if n.value == 0:
                return

    def forces(self, cid: int) -> dict:
        cl, cd = C.c_double(), C.c_double()
        _check(lib().jst_batch_forces(self._h, cid, C.byref(cl), C.byref(cd)))
        return {"cl": cl.value, "cd": cd.value}

    def set_state(self, cid: int, u_planes: np.ndarray) -> None:
        """Inject U (5 x NI x NJ) as the case's current state (checkpoint / resume)."""
        u = np.ascontiguousarray(u_planes, dtype=np.float64)
        rec = _NonPhys()
        st = lib().jst_batch_set_state(self._h, cid, _dp(u), C.byref(rec))
        if st == JST_ERR_NONPHYSICAL:
            raise NonPhysical(_err(), rec)
        _check(st)

    def wall(self, cid: int, nj: int) -> dict:
        """jst_wall for case cid (nj = its wall face count)."""
        a = {k: np.zeros(nj) for k in ("x", "y", "cp", "cf")}
        _check(lib().jst_batch_wall(self._h, cid, *(_dp(a[k]) for k in ("x", "y", "cp", "cf"))))
        return a

    def get(self, cid: int, name: str, ni: int, nj: int) -> np.ndarray:
        rows = ni + 1 if name.startswith("tau_") else ni
        buf = np.zeros(rows * nj)
        _check(lib().jst_batch_get_field(self._h, cid, name.encode(), _dp(buf), buf.size))
        return buf

    def residual_terms(self, cid: int) -> float:
        dt = C.c_double()
        _check(lib().jst_batch_residual_terms(self._h, cid, C.byref(dt)))
        return dt.value

    def residual_terms_noalloc(self, cid: int) -> float:
        """residual_terms without any device allocation (fixed per-device work-item array)."""
        dt = C.c_double()
        _check(lib().jst_batch_residual_terms_noalloc(self._h, cid, C.byref(dt)))
        return dt.value

    def history(self, cid: int, first: int, n: int) -> np.ndarray:
        out = np.zeros(n)
        _check(lib().jst_batch_history(self._h, cid, int(first), int(n), _dp(out)))
        return out

    def set_tiling(self, variant: int, grid: int = 0) -> None:
        _check(lib().jst_batch_set_tiling(self._h, int(variant), int(grid)))

    def segment_rows(self, cid: int) -> int:
        """Row-segment length the stage kernel runs case ``cid`` with."""
        r = C.c_int()
        _check(lib().jst_batch_segment_rows(self._h, int(cid), C.byref(r)))
        return r.value

    def tune(self, trial_steps: int = 5):
        v, g = C.c_int(), C.c_int()
        _check(lib().jst_batch_tune(self._h, int(trial_steps), C.byref(v), C.byref(g)))
        return v.value, g.value

    def time_steps(self, nsteps: int, per_stage: bool = False):
        total, stage = C.c_double(), C.c_double()
        _check(lib().jst_batch_time_steps(self._h, int(nsteps), C.byref(total),
                                          C.byref(stage) if per_stage else None))
        return total.value, (stage.value if per_stage else None)

    def time_begin(self, nsteps: int) -> None:
        """Untimed set-up of nsteps timed steps (stops off, graphs built and warmed)."""
        _check(lib().jst_batch_time_begin(self._h, int(nsteps)))

    def time_end(self) -> None:
        _check(lib().jst_batch_time_end(self._h))

    def time_kernels(self, nsteps: int):
        """Mean launch durations (ms) of RK stages 1-4, stage 5 and the step end."""
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        _check(lib().jst_batch_time_kernels(self._h, int(nsteps), C.byref(a), C.byref(b), C.byref(c)))
        return {"inner_ms": a.value, "last_ms": b.value, "end_ms": c.value}

    def counts(self):
        n, cells = C.c_int(), C.c_long()
        _check(lib().jst_batch_counts(self._h, C.byref(n), C.byref(cells)))
        return n.value, cells.value


def serve(requests, hbm_bytes: int = 16 << 30, max_cases: int = 64, steps_per_round: int = 50,
          tune: bool = False, segment_rows: int = 0):
    """Serving loop over ``requests`` = [(arrival_s, config, mesh, max_steps), ...].

    Returns (per-request results, summary) from jst_serve."""
    n = len(requests)
    keep = []
    reqs = (_Request * max(n, 1))()
    for q, (t, cfg, mesh, steps) in enumerate(requests):
        cb, mb = cfg.encode(), mesh.encode()
        keep += [cb, mb]
        reqs[q] = _Request(float(t), cb, mb, int(steps))
    res = (_RequestResult * max(n, 1))()
    summ = _ServeSummary()
    opts = _ServeOpts(int(hbm_bytes), int(max_cases), int(steps_per_round), 1 if tune else 0,
                      int(segment_rows))
    _check(lib().jst_serve(reqs, n, C.byref(opts), res, C.byref(summ)))
    out = [{"admit_s": r.admit_s, "done_s": r.done_s, "steps": r.steps, "status": r.status,
            "residual": r.residual, "cl": r.cl, "cd": r.cd} for r in res[:n]]
    return out, {"wall_s": summ.wall_s, "cell_stages": summ.cell_stages, "p50_s": summ.p50_s,
                 "p99_s": summ.p99_s, "completed": summ.completed,
                 "nonphysical": summ.nonphysical, "max_concurrent": summ.max_concurrent,
                 "tiling_switches": summ.tiling_switches, "converged": summ.converged,
                 "rejected": summ.rejected, "p50_converged_s": summ.p50_converged_s,
                 "p99_converged_s": summ.p99_converged_s}


def copy_probe(nbytes: int) -> float:
    """Device-to-device copy of nbytes (profiler byte-counter calibration); returns ms."""
    ms = C.c_double()
    _check(lib().jst_copy_probe(int(nbytes), C.byref(ms)))
    return ms.value
