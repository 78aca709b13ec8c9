"""subsweep_amd: MI355X-native sweep radiative transfer + hydrogen chemistry.

The compute path is libsubsweep_amd.so (hand-written gfx950 HIP kernels behind
the C ABI of include/subsweep_amd.h); this package is its host-side mirror of
the reference's sweep plugin interface."""

from ._native import SweepError, FACE_LOCAL, FACE_PERIODIC, FACE_BOUNDARY, TRANSPORT_FP32, TRANSPORT_FP64
from .grid import Grid, cartesian_grid, voronoi_grid, jittered_lattice, lognormal_density, assign_sources
from .sweep import (Sweep, SweepParameters, SweepPlugin, SweepResults, direction_table, random_rotation_matrix,
                    rotation_matrix)

__all__ = [
    "Grid", "Sweep", "SweepError", "SweepParameters", "SweepPlugin", "SweepResults",
    "cartesian_grid", "voronoi_grid", "jittered_lattice", "lognormal_density", "assign_sources",
    "FACE_LOCAL", "FACE_PERIODIC", "FACE_BOUNDARY", "TRANSPORT_FP32", "TRANSPORT_FP64", "direction_table",
    "rotation_matrix", "random_rotation_matrix",
]
