"""MI355X-native RCWA: the layer-eigensolve + S-matrix path of StableRCWA on gfx950.

The compute path is libsrcwa_amd.so (HIP kernels + C-ABI, include/srcwa.h); this package is its
Python mirror of the reference's entry points (``rcwa.solve``, ``rcwa.sweep``, ...).
"""
from . import types
from .types import (FourierMethod, Layer, Layer2D, Polarization, Result, Result2D, Stack, Stack2D, Status2D,
                    XProfile2D, jones_from_angles_2d)

__all__ = ["types", "Layer", "Stack", "Polarization", "FourierMethod", "Result", "Layer2D", "Stack2D",
           "XProfile2D", "Result2D", "Status2D", "jones_from_angles_2d"]
