"""HiPIMS 2-D shallow-water step for MI355X (gfx950).

Modules:
  SWE_CUDA    drop-in ``Godunov`` driver (reference: hipims/pythonHipims/SWE_CUDA.py)
  ops         the reference's torch-extension ops on its compact layout
  grid        structured-grid state and the native step (flux / update passes)
  generators  seeded synthetic inputs in the reference's raster format
"""
from ._lib import build, lib  # noqa: F401

__all__ = ["build", "lib"]
