"""jst-gpu, the command-line driver with jst-rs's CLI flags and output formats
(jst-rs/src/bin/jst.rs:10-123), on the GPU."""
from __future__ import annotations

import re
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT, config_dict, mesh_text

pytestmark = pytest.mark.gpu

CLI = ROOT / "purinecfd_amd" / "_build" / "jst-gpu"
RUST_EXP6 = r"-?\d\.\d{6}e-?\d+"


@pytest.fixture(scope="module")
def jst():
    from purinecfd_amd import jst as mod
    mod.lib()
    return mod


def run_cli(tmp_path, *args):
    p = subprocess.run([str(CLI), *args], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    return p.returncode, p.stdout, p.stderr


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cli_outputs_match_the_reference_formats(tmp_path, oracle_lib, precision):
    rc, out, err = run_cli(tmp_path, "--mesh", str(GOLDEN / "fangdata.txt"), "--config",
                           str(GOLDEN / "config.json"), "--steps", "5", "--precision", precision)
    assert rc == 0, err
    lines = out.splitlines()
    assert lines[0] == f"mesh {GOLDEN / 'fangdata.txt'}: 10 rings x 12 points -> 108 cells"
    assert re.fullmatch(r"threads: \d+", lines[1])
    steps = [ln for ln in lines if ln.startswith("step:")]
    assert len(steps) == 5
    for k, ln in enumerate(steps, 1):
        assert re.fullmatch(rf"step:{k:6d} \| residual:{RUST_EXP6}", ln), ln
    assert re.fullmatch(rf"steps: 5  final residual: {RUST_EXP6}", lines[-4])
    assert re.fullmatch(rf"totaltime \(physical\): {RUST_EXP6} s", lines[-3])
    assert re.fullmatch(r"wall clock: \d+\.\d{6} s  \(\d+\.\d{4} ms/step, \d+\.\d ns/cell/step\)", lines[-2])
    assert lines[-1] == "all data is written in result.csv"
    log = (tmp_path / "res.log").read_text().splitlines()
    assert log[0] == "step,residual" and len(log) == 6
    res = np.array([float(x.split(",")[1]) for x in log[1:]])
    assert [int(x.split(",")[0]) for x in log[1:]] == [1, 2, 3, 4, 5]
    o = oracle_lib.OracleSolver(config_dict(), mesh_text("fangdata.txt"))
    want = np.array(o.run(5)["history"])
    tol = 5e-7 if precision == "fp64" else 1e-4  # res.log keeps 7 significant digits
    assert np.all(np.abs(res - want) <= tol * np.abs(want)), (res, want)
    csv = (tmp_path / "result.csv").read_text().splitlines()
    assert csv[0] == "i,j,x,y,rho,p,T,u,v,miubl" and len(csv) == 1 + 108


def test_cli_csv_equals_the_solver_output(tmp_path, jst):
    rc, _, err = run_cli(tmp_path, "--mesh", str(GOLDEN / "fangdata.txt"), "--config",
                         str(GOLDEN / "config.json"), "--steps", "7", "--quiet")
    assert rc == 0, err
    s = jst.Solver((GOLDEN / "config.json").read_text(), mesh_text("fangdata.txt"))
    s.run(7)
    s.write_csv(tmp_path / "api.csv")
    assert (tmp_path / "result.csv").read_text() == (tmp_path / "api.csv").read_text()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cli_wall_table_equals_the_solver_output(tmp_path, jst, precision):
    rc, out, err = run_cli(tmp_path, "--mesh", str(GOLDEN / "fangdata.txt"), "--config",
                           str(GOLDEN / "config.json"), "--steps", "7", "--quiet", "--wall",
                           "wall.csv", "--precision", precision)
    assert rc == 0, err
    assert out.splitlines()[-1] == "wall data is written in wall.csv"
    rows = (tmp_path / "wall.csv").read_text().splitlines()
    assert rows[0] == "j,x,y,cp,cf" and len(rows) == 1 + 12
    assert [int(r.split(",")[0]) for r in rows[1:]] == list(range(1, 13))
    for r in rows[1:]:
        for v in r.split(",")[1:]:
            assert re.fullmatch(r"-?\d\.\d{8}e-?\d+", v), r
    prec = jst.FP64 if precision == "fp64" else jst.MIXED
    s = jst.Solver((GOLDEN / "config.json").read_text(), mesh_text("fangdata.txt"), prec)
    s.run(7)
    s.write_wall_csv(tmp_path / "api_wall.csv")
    assert (tmp_path / "wall.csv").read_text() == (tmp_path / "api_wall.csv").read_text()
    w = s.wall()
    got = np.array([[float(v) for v in r.split(",")[1:]] for r in rows[1:]])
    for c, k in enumerate(("x", "y", "cp", "cf")):
        assert np.allclose(got[:, c], w[k], rtol=1e-8, atol=0.0), k


def test_cli_forces_line_equals_the_solver_forces(tmp_path, jst):
    rc, out, err = run_cli(tmp_path, "--mesh", str(GOLDEN / "fangdata.txt"), "--config",
                           str(GOLDEN / "config.json"), "--steps", "9", "--quiet", "--forces",
                           "--no-output")
    assert rc == 0, err
    line = out.splitlines()[-1]
    m = re.fullmatch(r"CL: (-?\d\.\d{8}e-?\d+)  CD: (-?\d\.\d{8}e-?\d+)", line)
    assert m, line
    s = jst.Solver((GOLDEN / "config.json").read_text(), mesh_text("fangdata.txt"))
    s.run(9)
    f = s.forces()
    assert abs(float(m.group(1)) - f["cl"]) <= 5e-9 * max(1e-12, abs(f["cl"]))
    assert abs(float(m.group(2)) - f["cd"]) <= 5e-9 * max(1e-12, abs(f["cd"]))


def test_cli_reports_nonphysical(tmp_path):
    cfg = config_dict()
    cfg["simulation"]["CFL"] = 200.0
    (tmp_path / "bad.json").write_text(__import__("json").dumps(cfg))
    rc, out, err = run_cli(tmp_path, "--mesh", str(GOLDEN / "fangdata.txt"), "--config", "bad.json",
                           "--steps", "50", "--quiet", "--no-output")
    assert rc == 1
    assert re.match(r"Error: NonPhysical \{ i: \d+, j: \d+, rho: ", err), err
    assert not (tmp_path / "result.csv").exists()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cli_profile_times_extra_steps_after_the_outputs(tmp_path, precision):
    """--profile N times N further steps with HIP events once the outputs are written, so
    result.csv and res.log are those of the run without it."""
    common = ["--mesh", str(GOLDEN / "cyl_17x40.txt"), "--config", str(GOLDEN / "config.json"),
              "--steps", "20", "--quiet", "--precision", precision]
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    rc, out_a, err = run_cli(tmp_path / "a", *common)
    assert rc == 0, err
    rc, out_b, err = run_cli(tmp_path / "b", *common, "--profile", "10")
    assert rc == 0, err
    m = re.fullmatch(r"profile: 10 steps \(HIP events\): (\d+\.\d{4}) ms/step, stage kernel "
                     r"(\d+\.\d{2}) us, (\d\.\d{3}e[+-]\d+) cell-stage updates/s", out_b.splitlines()[-1])
    assert m, out_b.splitlines()[-1]
    ms, stage_us, rate = float(m.group(1)), float(m.group(2)), float(m.group(3))
    assert 0.0 < stage_us * 1e-3 * 5 <= 2.0 * ms and rate > 0.0
    for name in ("result.csv", "res.log"):
        assert (tmp_path / "a" / name).read_text() == (tmp_path / "b" / name).read_text(), name
