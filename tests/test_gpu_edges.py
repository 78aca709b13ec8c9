"""GPU edge cases against the oracle (FP64 transport, 1e-10): a single cell
with boundary faces only, a 2^3 periodic grid (every neighbour also a
periodic image), one explicit direction, no sources at all, all cells on
the deepest timestep level, and every launch shape on each."""

import numpy as np
import pytest

from subsweep_amd import Sweep
from subsweep_amd import _native as N
from subsweep_amd import workloads as W
from oracle.oracle import OracleSweep

from test_gpu_parity import _compare_fields, _compare_structure

pytestmark = pytest.mark.gpu

SHAPES = ({}, {"group_levels": 1}, {"transport_mode": N.TRANSPORT_MODE_GEOMETRIC})


def _run(pb, steps):
    levels = pb.params.num_timestep_levels
    for kw in SHAPES:
        o = OracleSweep(pb.params, pb.grid, *pb.state())
        g = Sweep(pb.params, pb.grid, *pb.state(), **kw)
        _compare_structure(g, o, levels)
        for _ in range(steps):
            assert g.run_sweeps() == o.run_sweeps()
        _compare_fields(g, o, 1e-10)
        g.close()
        o.close()


@pytest.mark.parametrize("dirs", [1, 16, 84])
def test_single_cell_boundary_only(dirs, hip_available):
    _run(W.cartesian_problem(1, dirs, levels=1, periodic=False, source_cells=(0,), steps=3), 3)


def test_two_cubed_periodic(hip_available):
    _run(W.cartesian_problem(2, 84, levels=2, periodic=True, source_cells=(0,), steps=4), 4)


def test_one_explicit_direction(hip_available):
    pb = W.cartesian_problem(6, 1, levels=1, periodic=False, source_cells=(0, 7), steps=3)
    pb = pb.with_params(directions=[[0.3, -0.2, 0.9]])
    _run(pb, 3)


def test_no_sources(hip_available):
    pb = W.cartesian_problem(5, 16, levels=2, periodic=True, source_cells=(), steps=3)
    _run(pb, 3)


def test_deepest_levels(hip_available):
    # 6 timestep levels, a short maximum timestep: the cadence walks every
    # level down, restricted sweeps of nearly empty active sets included
    pb = W.jittered_problem(6, 1.0, "2", 16, levels=6, periodic=True, steps=8)
    pb = pb.with_params(max_timestep=1e-4 * 3.15576e13)
    _run(pb, 8)
