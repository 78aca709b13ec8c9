"""GPU: source placement on the device (ssw_assign_sources_device,
set_source_terms_system, src/source_systems.rs:48-80) against the host
placement: bitwise the same per-cell rates and the same cell for every
source, on random points, on a lattice with sources at exact ties (midpoints
and corners of lattice cells, where the lower cell index must win), with
duplicate generator points, with more cells than one search chunk holds,
and on the 2097152-cell clustered workload's 1024 sources."""

import numpy as np
import pytest

from subsweep_amd import SweepError, assign_sources
from subsweep_amd import workloads as W

pytestmark = pytest.mark.gpu


def _same(pts, pos, rates):
    hs, hc = assign_sources(pts, pos, rates)
    ds, dc = assign_sources(pts, pos, rates, device=True)
    assert np.array_equal(hc, dc), f"{int((hc != dc).sum())} sources placed differently"
    assert np.array_equal(hs.view(np.uint64), ds.view(np.uint64)), "per-cell rates differ"
    return dc


def test_random_points(hip_available):
    rng = np.random.default_rng(5)
    pts = rng.random((50000, 3))
    pos = rng.random((300, 3))
    rates = 10.0 ** rng.uniform(48, 51, 300)
    _same(pts, pos, rates)


def test_exact_ties_go_to_the_lower_index(hip_available):
    g = np.arange(8) + 0.5
    pts = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    rng = np.random.default_rng(6)
    # midpoints between lattice points and cell corners: 2 and 8 cells at the same distance
    mids = pts[rng.integers(0, len(pts), 40)] + np.array([0.5, 0.0, 0.0])
    corners = np.floor(rng.random((40, 3)) * 7) + 1.0
    pos = np.concatenate([mids, corners])
    cells = _same(pts, pos, np.ones(len(pos)))
    d = ((pts[None, :, :] - pos[:, None, :]) ** 2).sum(-1)
    for s in range(len(pos)):
        assert cells[s] == np.flatnonzero(d[s] == d[s].min())[0]


def test_duplicate_points_and_many_chunks(hip_available):
    rng = np.random.default_rng(7)
    pts = rng.random((600000, 3))  # more cells than one search chunk (2^18)
    pts[300000:300010] = pts[10]  # duplicates of a point early in the list
    pos = np.concatenate([pts[[10, 5, 599999]], rng.random((61, 3))])
    cells = _same(pts, pos, rng.random(len(pos)))
    assert cells[0] == 10


def test_clustered_workload_sources(hip_available):
    pb = W.clustered_2m(precision=32)
    pts = pb.grid.points
    rng = np.random.default_rng(8)
    pos = pts[rng.integers(0, len(pts), 1024)] + rng.normal(0.0, 1e-3, (1024, 3))
    _same(pts, pos, 10.0 ** rng.uniform(48, 52, 1024))


def test_bad_input_is_an_error(hip_available):
    pts = np.random.default_rng(9).random((100, 3))
    with pytest.raises(SweepError):
        assign_sources(pts, np.full((1, 3), np.nan), np.ones(1), device=True)
