"""kratos_amd -- MI355X-native Godunov hydro/MHD stage (PLM + HLLC/HLLD + CT
+ Heun RK2) behind a C-ABI, with a Python host layer for tests and benches.

The compute path is ``libkratos_amd.so`` (hand-written gfx950 HIP kernels);
this package only describes meshes, moves state in and out and generates
initial conditions.  There is no CPU fallback: without the built library
every entry point raises ``KratosError``.
"""
from ._capi import KratosError, LIB_PATH, load  # noqa: F401
from .mesh import Context, MergedMeshSpec, MeshSpec, Physics, RefinedMeshSpec  # noqa: F401
from . import problems  # noqa: F401

__all__ = ["Context", "MeshSpec", "Physics", "KratosError", "load", "problems", "LIB_PATH"]
