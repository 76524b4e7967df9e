"""MI355X-native explicit Moelans phase-field stepper.

A drop-in for the stepping loop of jijn/phasefield_coarsening
(phasefield.c:178-258): HIP kernels for gfx950 behind the C ABI declared in
include/ppf.h, with the reference's HDF5 file contract (include/ppf_io.h).
"""
from . import _lib  # noqa: F401

__all__ = ["h5", "solver"]
