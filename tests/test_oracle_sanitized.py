"""The oracle under AddressSanitizer and UndefinedBehaviorSanitizer (host code only): a short
run of the golden 17x40 cylinder case through every oracle entry point the other tests use
(config, mesh parsing and its error paths, steps, residual terms, forces, the wall
distribution, field export, the default-off extensions on a state with negative nu~, a state
injection that must report NonPhysical), in a child
process that preloads the compiler's libasan.  Any invalid access, leak-free or not, or
undefined behaviour aborts the child with a report."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

import pytest

from conftest import ROOT

ORACLE = ROOT / "oracle"

SCRIPT = r"""
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
from oracle import pyoracle as po
po.LIB_PATH = po.HERE / "_build" / "libjstoracle_san.so"
golden = po.HERE.parent / "tests" / "golden"
cfg = json.loads((golden / "config.json").read_text())
text = (golden / "cyl_17x40.txt").read_text()
for bad in ("", "x 8\n", "4 8\n1 2\n", "4 8\n" + "1 zz\n" * 32, "2 4\n" + "1 2\n" * 8):
    try:
        po.parse_mesh(bad)
        raise SystemExit("bad mesh accepted: %r" % bad[:20])
    except po.MeshError:
        pass
r, t, xy = po.parse_mesh(text)
s = po.OracleSolver(cfg, text)
rep = s.run(5)
assert rep["steps"] == 5 and np.isfinite(rep["residual"])
dt = s.timestep()
s.residual_terms()
for name in ("rho", "p", "u", "v", "nut", "Fc0", "Fv1", "Fd2", "S", "vol", "tau_nx", "wall_dist"):
    assert np.all(np.isfinite(s.get(name))), name
f = s.forces()
w = s.wall()
assert np.all(np.isfinite(w["cp"])) and np.all(np.isfinite(w["cf"]))
# the default-off extensions on one short run: local and viscous time steps, the
# negative-nu~ SA branch on a state whose nu~ is negative in a band of rows
ext = json.loads(json.dumps(cfg))
ext["simulation"].update({"local_time_stepping": True, "viscous_time_step": True})
ext["spalart_allmaras"]["negative"] = True
x = po.OracleSolver(ext, text)
ux = np.stack([s.get("U%d" % k) for k in range(5)])
ux[4, 40:200] = -ux[4, 40:200]
x.set_state(ux)
assert x.run(3)["steps"] == 3 and np.all(np.isfinite(x.get("nut")))
x.timestep()
x.residual_terms()
assert np.all(np.isfinite(x.get("S"))) and np.all(np.isfinite(x.get("Fv4")))
u = np.stack([s.get("U%d" % k) for k in range(5)])
u[0, 3] = -1.0
try:
    s.set_state(u)
    raise SystemExit("rho = -1 not reported")
except po.NonPhysical:
    pass
print("sanitized oracle ok", r, t, rep["steps"], dt > 0, np.isfinite(f["cd"]))
"""


def test_oracle_runs_clean_under_asan_and_ubsan():
    libasan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                             text=True).stdout.strip()
    if not libasan or not Path(libasan).exists():
        pytest.skip("the compiler's libasan is not installed")
    mk = subprocess.run(["make", "-C", str(ORACLE), "sanitized"], capture_output=True, text=True,
                        timeout=300)
    assert mk.returncode == 0, mk.stderr
    env = dict(os.environ, LD_PRELOAD=libasan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="2")
    p = subprocess.run([sys.executable, "-c", SCRIPT, str(ROOT)], capture_output=True, text=True,
                       env=env, timeout=600)
    report = p.stdout + p.stderr
    assert p.returncode == 0, report[-4000:]
    assert "AddressSanitizer" not in report and "runtime error" not in report, report[-4000:]
    assert "sanitized oracle ok" in p.stdout
