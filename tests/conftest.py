"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on the CPU."""
from __future__ import annotations

import json
import math
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = ROOT / "tests" / "golden"
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))

GOLDEN_FILES = sorted(GOLDEN.glob("cyl_*.json"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


def config_text() -> str:
    return (GOLDEN / "config.json").read_text()


def config_dict() -> dict:
    return json.loads(config_text())


def mesh_text(name: str) -> str:
    return (GOLDEN / name).read_text()


def load_golden(path: Path) -> dict:
    return json.loads(path.read_text())


def field_cmp(got, want, label: str, rtol: float, atol: float = 0.0) -> None:
    """Field-normed comparison |got-want| <= atol + rtol*max|want| (jst-rs tests/golden.rs:59-79)."""
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    assert got.shape == want.shape, f"{label}: shape {got.shape} != {want.shape}"
    scale = float(np.max(np.abs(want))) if want.size else 0.0
    tol = atol + rtol * scale
    err = np.abs(got - want)
    k = int(np.argmax(err)) if err.size else 0
    assert err.size == 0 or err[k] <= tol, (
        f"{label}: mismatch at flat index {k}: got {got.flat[k]:e}, want {want.flat[k]:e} "
        f"(abs err {err[k]:e} > tol {tol:e}; field scale {scale:e}, rtol {rtol:e})")


def ellipse_mesh(rings: int, nj: int, stretch: float = 1.0, a_wall=1.0, b_wall=0.5,
                 r_far=5.0, closed=False, fmt="{:.12f} {:.12f}\n") -> str:
    """Ellipse -> circle O-mesh family of jst-rs tests/properties.rs:14-38."""
    import math
    if abs(stretch - 1.0) < 1e-12:
        svals = [i / (rings - 1) for i in range(rings)]
    else:
        w = [stretch ** k for k in range(rings - 1)]
        tot = sum(w)
        svals, acc = [0.0], 0.0
        for x in w:
            acc += x
            svals.append(acc / tot)
    per = nj + 1 if closed else nj
    out = [f"{rings} {per}\n"]
    for s in svals:
        a = a_wall + s * (r_far - a_wall)
        b = b_wall + s * (r_far - b_wall)
        for j in range(per):
            th = 2.0 * math.pi * (j % nj) / nj
            out.append(fmt.format(a * math.cos(th), b * math.sin(th)))
    return "".join(out)


def ring_mass_flux(u1, u2, tnx, tny, ni, nj):
    """Convective mass flux through each interior tau ring i = 1..NI-1 from the state: the face
    carries (rho u)_f . n with U_f the mean of the two cells (convection.rs:25-40 with the
    face average of convection.rs:56-66); returns Phi[i] for i = 1..NI-1."""
    m1 = u1.reshape(ni, nj)
    m2 = u2.reshape(ni, nj)
    nx = tnx.reshape(ni + 1, nj)[1:ni]
    ny = tny.reshape(ni + 1, nj)[1:ni]
    f = 0.5 * (m1[:-1] + m1[1:]) * nx + 0.5 * (m2[:-1] + m2[1:]) * ny
    return f.sum(axis=1)


def check_mass_conservation(fc0, u1, u2, tnx, tny, ni, nj, rtol):
    """Discrete conservation, ring by ring: the n faces cancel around each (periodic) ring, so
    the mass residual summed over rows 0..i-1 equals the flux through ring i minus the flux
    through the wall.  With the wall flux zero (mirror ghost) every partial sum equals the
    ring flux Phi_i computed from the state; a sign slip in the face assembly of any row, or
    mass leaking through the wall, breaks it."""
    fc = np.asarray(fc0).reshape(ni, nj)
    cum = np.cumsum(fc.sum(axis=1))[: ni - 1]  # rows 0..i-1 for i = 1..NI-1
    phi = ring_mass_flux(u1, u2, tnx, tny, ni, nj)
    scale = float(np.abs(fc).sum())
    err = np.abs(cum - phi)
    assert float(err.max()) <= rtol * scale, (float(err.max()), scale, int(np.argmax(err)))


def rotate_mesh(text: str, angle_rad: float) -> str:
    lines = text.splitlines()
    ca, sa = math.cos(angle_rad), math.sin(angle_rad)
    out = [lines[0]]
    for ln in lines[1:]:
        x, y = (float(v) for v in ln.split())
        out.append(f"{ca * x - sa * y:.17g} {sa * x + ca * y:.17g}")
    return "\n".join(out) + "\n"


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pyoracle
    pyoracle.build()
    return pyoracle
