"""GPU, full size: the 2097152-cell clustered workload (84 directions, 4
timestep levels), where the oracle would take minutes per step, checked
through properties that do not depend on size:

  * launch-shape invariance: grouped replay (default, and with 32 roots per
    workgroup), one launch per level set, the geometric replay and the
    direction sweeps give bitwise the same fields;
  * the Kahn depths form a valid order: for every local face f of cell c
    whose normal points against direction d, depth(c, d) > depth(nb, d)
    (checked for a sample of directions), and the dependency counts equal
    the number of such upwind faces per (cell, direction);
  * at full size, the dependency counts and Kahn depths of all 84 directions
    are bit-exact against the oracle's (its Kahn layering takes about a
    minute on eight host threads, far less than a step of its sweep)."""

import numpy as np
import pytest

from oracle.oracle import OracleSweep
from subsweep_amd import Sweep, direction_table
from subsweep_amd import _native as N
from subsweep_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def clustered():
    return W.clustered_2m(precision=32)


def test_clustered_launch_shapes_bitwise(clustered, hip_available):
    pb = clustered
    out = []
    for kw in ({}, {"group_block": 32}, {"group_levels": 1}, {"transport_mode": N.TRANSPORT_MODE_GEOMETRIC},
               {"direction_sweeps": True}):
        s = Sweep(pb.params, pb.grid, *pb.state(), **kw)
        for _ in range(3):
            s.run_sweeps()
        r = s.results(photon_rate=False)
        out.append((r.ionized_hydrogen_fraction, r.temperature, r.level, r.change_timescale))
        s.close()
    for k, run in enumerate(out[1:], 1):
        for u, v in zip(out[0], run):
            assert np.array_equal(u.view(np.uint8), v.view(np.uint8)), f"launch shape {k}"


def test_clustered_depths_are_a_valid_order(clustered, hip_available):
    pb = clustered
    g = pb.grid
    s = Sweep(pb.params, g, *pb.state())
    depth = s.levelsets(0)
    counts = s.counts(0)
    s.close()
    dirs = direction_table(84)
    cell = np.repeat(np.arange(g.num_cells), np.diff(g.offsets))
    loc = g.kind == 0
    c, nb, nrm = cell[loc], g.neighbour[loc].astype(np.int64), g.normal.reshape(-1, 3)[loc]
    for d in (0, 13, 29, 41, 58, 83):
        dd = (nrm[:, 0] * dirs[d, 0] + nrm[:, 1] * dirs[d, 1]) + nrm[:, 2] * dirs[d, 2]
        up = dd < 0
        assert np.all(depth[d][c[up]] > depth[d][nb[up]]), f"direction {d}: a task not after its upwind task"
        n_up = np.bincount(c[up], minlength=g.num_cells)
        assert np.array_equal(counts[d].astype(np.int64), n_up), f"direction {d}: counts"


def test_clustered_structure_against_the_oracle(clustered, hip_available):
    pb = clustered
    s = Sweep(pb.params, pb.grid, *pb.state())
    depth, counts = s.levelsets(0), s.counts(0)
    s.close()
    o = OracleSweep(pb.with_params(transport_precision=64).params, pb.grid, *pb.state(), threads=8)
    assert np.array_equal(counts, o.counts(0)), "dependency counts differ from the oracle's"
    assert np.array_equal(depth, o.levelsets(0)), "Kahn depths differ from the oracle's"
    o.close()
