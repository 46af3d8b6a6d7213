"""MI355X (gfx950) implementation of the Mantle_Convection_Lab per-step hot path.

The Uzawa Stokes solve (Jacobi-PCG inner solver) and the SUPG temperature
update run as hand-written HIP kernels in libmcl_hip.so behind the C-ABI of
include/mcl.h; this package holds the kernels' sources (csrc/), the C host
layer (input format, quadtree mesh), and a Python mirror of the reference's
driver loop.
"""
from ._native import NativeMissing, hip_lib, host_lib  # noqa: F401
from .model import Context, Input, MclError, Mesh, Model  # noqa: F401

__all__ = ["Context", "Input", "Mesh", "Model", "MclError", "NativeMissing", "hip_lib", "host_lib"]
