"""MI355X-native (gfx950) JST + Spalart-Allmaras residual-and-update loop.

The solver library is ``purinecfd_amd/_build/libjstgpu.so`` (HIP kernels + C ABI declared in
``include/jst_gpu.h``); ``purinecfd_amd.jst`` mirrors the jst-rs crate surface over it.
"""
from .jst import (FP64, MIXED, Config, ConfigError, JstError, Mesh, MeshError,  # noqa: F401
                  NonPhysical, RunReport, Solver, lib)
