"""GPU, the largest single-GPU domain: 16777216 clustered cells (sigma_ln
2.5, 256^3 field mesh), 8192 sources, 84 directions, 4 timestep levels,
periodic, threshold 1e-5 s^-1, FP32 transport. Building its grid takes about
two minutes on 16 host threads and the test about three minutes in all,
most of it without output under pytest's capture, so it runs only with
SSW_LARGE_TESTS=1 (and -s, so that the grid builder's progress lines show);
it sorts after the other GPU tests.
Checked through properties that do not depend on size: the compiled grouped
replay and the geometric replay give bitwise the same fields after two
steps, and the Kahn depths form a valid order with counts equal to the
number of upwind faces (sampled directions)."""

import os

import numpy as np
import pytest

from subsweep_amd import Sweep, direction_table
from subsweep_amd import _native as N
from subsweep_amd import workloads as W

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.environ.get("SSW_LARGE_TESTS"),
                                 reason="set SSW_LARGE_TESTS=1 (and run with -s): the 16777216-cell grid takes "
                                        "minutes to build")]


@pytest.fixture(scope="module")
def largest():
    print("building the 16777216-cell grid", flush=True)
    return W.clustered_problem(16777216, levels=4, sources=8192, sigma_ln=2.5, mesh=256, threshold=1e-5, steps=2,
                               precision=32, name="clustered_16m")


def test_largest_launch_shapes_bitwise(largest, hip_available):
    pb = largest
    out = []
    for kw in ({"transport_mode": N.TRANSPORT_MODE_COMPILED}, {"transport_mode": N.TRANSPORT_MODE_GEOMETRIC}):
        s = Sweep(pb.params, pb.grid, *pb.state(), **kw)
        for k in range(pb.steps):
            s.run_sweeps()
            print(f"{kw}: step {k + 1} done", flush=True)
        r = s.results(photon_rate=False)
        out.append((r.ionized_hydrogen_fraction, r.temperature, r.level, r.change_timescale))
        if kw["transport_mode"] == N.TRANSPORT_MODE_GEOMETRIC:
            depth, counts = s.levelsets(0), s.counts(0)
        s.close()
    for u, v in zip(out[0], out[1]):
        assert np.array_equal(u.view(np.uint8), v.view(np.uint8)), "compiled and geometric replays differ"
    g = pb.grid
    dirs = direction_table(84)
    cell = np.repeat(np.arange(g.num_cells), np.diff(g.offsets))
    loc = g.kind == 0
    c, nb, nrm = cell[loc], g.neighbour[loc].astype(np.int64), g.normal.reshape(-1, 3)[loc]
    for d in (0, 41, 83):
        dd = (nrm[:, 0] * dirs[d, 0] + nrm[:, 1] * dirs[d, 1]) + nrm[:, 2] * dirs[d, 2]
        up = dd < 0
        assert np.all(depth[d][c[up]] > depth[d][nb[up]]), f"direction {d}: a task not after its upwind task"
        assert np.array_equal(counts[d].astype(np.int64), np.bincount(c[up], minlength=g.num_cells)), d
