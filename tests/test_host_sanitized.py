"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer: jst_host.cpp
is rebuilt with the sanitizers on its host pass only (the device code of the other objects is
linked as built) and driven by tests/host_fuzz/mesh_parse_fuzz.cpp, a seeded fuzz of the
mesh-text parser that makes no GPU call."""
from __future__ import annotations

import os
import subprocess

import pytest

from conftest import ROOT

CSRC = ROOT / "purinecfd_amd" / "csrc"
BUILD = ROOT / "purinecfd_amd" / "_build"
HIPCC = "/opt/rocm/bin/hipcc"
CLANG = "/opt/rocm/llvm/bin/clang++"


def test_mesh_parser_fuzz_runs_clean_under_asan_and_ubsan(tmp_path):
    objs = [BUILD / f"{n}.o" for n in ("jst_fp64", "jst_mixed", "jst_batch")]
    if not (os.path.exists(HIPCC) and os.path.exists(CLANG)) or not all(o.exists() for o in objs):
        pytest.skip("the ROCm compilers or the library objects are missing (run __graft_entry__.build())")
    san = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
           "-Xarch_host", "-fno-sanitize-recover=undefined", "-Xarch_host", "-fno-omit-frame-pointer"]
    san_link = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer"]
    host_o, exe = tmp_path / "jst_host_san.o", tmp_path / "mesh_parse_fuzz"
    cc = subprocess.run([HIPCC, "--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC",
                         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", *san,
                         "-c", str(CSRC / "jst_host.cpp"), "-o", str(host_o)],
                        capture_output=True, text=True, timeout=600)
    assert cc.returncode == 0, cc.stderr[-3000:]
    # the driver and the link are host-only: the same compiler's clang++ with the same
    # sanitizers (so that one sanitizer runtime serves every object), against the HIP
    # runtime the library objects register their kernels with
    gxx = [CLANG, "-O1", "-g", "-std=c++17", *san_link]
    drv = tmp_path / "mesh_parse_fuzz.o"
    cc = subprocess.run([*gxx, "-c", str(ROOT / "tests" / "host_fuzz" / "mesh_parse_fuzz.cpp"),
                         "-o", str(drv)], capture_output=True, text=True, timeout=600)
    assert cc.returncode == 0, cc.stderr[-3000:]
    ln = subprocess.run([*gxx, str(drv), str(host_o), *map(str, objs), "-o", str(exe),
                         "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"],
                        capture_output=True, text=True, timeout=600)
    assert ln.returncode == 0, ln.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    p = subprocess.run([str(exe), "20000"], capture_output=True, text=True, env=env, timeout=600)
    report = p.stdout + p.stderr
    assert p.returncode == 0, report[-4000:]
    assert "AddressSanitizer" not in report and "runtime error" not in report, report[-4000:]
    assert "accepted" in p.stdout
