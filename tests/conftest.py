"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on the CPU."""
from __future__ import annotations

import json
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


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pyoracle
    pyoracle.build()
    return pyoracle
