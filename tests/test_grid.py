"""Grid construction (CPU): the Voronoi builder and the cartesian grid.

Pins: the reference's 5-point known answer (src/voronoi/mod.rs:191-244, in
tests/golden/reference_known_answers.json), volume closure of periodic grids,
exact mirror symmetry of local faces, the size formula of src/voronoi/cell.rs:186-188,
and an independent cross-check against scipy's Qhull Voronoi on a small
periodic point set (volumes, neighbour sets, face areas)."""

import json
import math
import os

import numpy as np
import pytest

from subsweep_amd.grid import cartesian_grid, jittered_lattice, uniform_points, voronoi_grid

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")))


def test_five_point_known_answer():
    g5 = GOLD["voronoi_five_points"]
    pts = np.array(g5["points"])
    g = voronoi_grid(pts, 1.0, free_space_box=(np.full(3, -10.0), np.full(3, 10.0)))
    c = g5["cell"]
    assert g.volume[c] == pytest.approx(g5["volume"], rel=1e-12)
    faces = range(g.offsets[c], g.offsets[c + 1])
    assert len(faces) == 4
    for f in faces:
        want = g5["faces"][str(g.neighbour[f])]
        assert g.kind[f] == 0
        assert g.area[f] == pytest.approx(want["area"], rel=1e-12)
        assert np.allclose(g.normal[f], want["normal"], rtol=0, atol=1e-15)


def _mirror_ok(g):
    cell = np.repeat(np.arange(g.num_cells), np.diff(g.offsets))
    loc = np.nonzero(g.kind == 0)[0]
    fwd = {(int(cell[f]), int(g.neighbour[f])): f for f in loc}
    assert len(fwd) == len(loc), "duplicate local faces"
    for (a, b), f in fwd.items():
        h = fwd[(b, a)]
        assert np.array_equal(g.normal[h], -g.normal[f])
        assert g.area[h] == g.area[f]


@pytest.mark.parametrize("n", [6, 12])
def test_periodic_jittered_grid_invariants(n):
    box = 3.0e22
    g = voronoi_grid(jittered_lattice(n, box, 0.3, 1338), box)
    assert g.dropped_faces == 0
    assert g.volume.sum() == pytest.approx(box ** 3, rel=1e-12)
    assert np.allclose(g.size, np.cbrt(3.0 * g.volume / (4.0 * math.pi)), rtol=1e-15)
    assert np.allclose(np.linalg.norm(g.normal, axis=1), 1.0, atol=1e-15)
    _mirror_ok(g)
    # every cell is closed: sum of area-weighted outward normals vanishes
    acc = np.zeros((g.num_cells, 3))
    np.add.at(acc, np.repeat(np.arange(g.num_cells), np.diff(g.offsets)), g.normal * g.area[:, None])
    scale = np.bincount(np.repeat(np.arange(g.num_cells), np.diff(g.offsets)), weights=g.area)
    assert np.all(np.linalg.norm(acc, axis=1) < 1e-9 * scale)


def test_against_qhull():
    from scipy.spatial import ConvexHull, Voronoi
    box = 1.0
    pts = uniform_points(40, box, 99)
    g = voronoi_grid(pts, box)
    shifts = np.array([(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)])
    tiled = np.concatenate([pts + s * box for s in shifts])
    v = Voronoi(tiled)
    n = len(pts)
    centre = int(np.nonzero((shifts == 0).all(axis=1))[0][0])
    ridge_map = {}
    for (a, b), rv in zip(v.ridge_points, v.ridge_vertices):
        ridge_map.setdefault(a, []).append((b, rv))
        ridge_map.setdefault(b, []).append((a, rv))
    for i in range(n):
        gi = centre * n + i
        region = v.regions[v.point_region[gi]]
        assert -1 not in region
        vol = ConvexHull(v.vertices[region]).volume
        assert g.volume[i] == pytest.approx(vol, rel=1e-9)
        want = {}
        for b, rv in ridge_map[gi]:
            j, s = b % n, shifts[b // n]
            poly = v.vertices[rv]
            cen = poly.mean(axis=0)
            nrm = tiled[b] - tiled[gi]
            nrm = nrm / np.linalg.norm(nrm)
            e1 = np.cross(nrm, [1.0, 0, 0] if abs(nrm[0]) < 0.9 else [0, 1.0, 0])
            e1 /= np.linalg.norm(e1)
            e2 = np.cross(nrm, e1)
            ang = np.arctan2((poly - cen) @ e2, (poly - cen) @ e1)
            poly = poly[np.argsort(ang)]
            area = 0.5 * np.linalg.norm(sum(np.cross(poly[k] - poly[0], poly[(k + 1) % len(poly)] - poly[0])
                                            for k in range(len(poly))))
            if area > 1e-12:
                want[(j, bool(s.any()))] = want.get((j, bool(s.any())), 0.0) + area
        got = {}
        for f in range(g.offsets[i], g.offsets[i + 1]):
            key = (int(g.neighbour[f]), bool(g.kind[f] == 1))
            got[key] = got.get(key, 0.0) + g.area[f]
        assert set(got) == set(want), i
        for key in got:
            assert got[key] == pytest.approx(want[key], rel=1e-8)


def test_cartesian_grid_semantics():
    n, box = 4, 8.0
    for periodic in (False, True):
        g = cartesian_grid(n, box, periodic)
        assert np.all(np.diff(g.offsets) == 6)
        assert np.all(g.area == (box / n) ** 2) and np.all(g.size == box / n)
        axes = np.array([[-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1]], float)
        assert np.array_equal(g.normal, np.tile(axes, (n ** 3, 1)))
        kinds = g.kind.reshape(-1, 6)
        ix = np.arange(n ** 3) // (n * n)
        assert np.all((kinds[:, 0] != 0) == (ix == 0))
        assert np.all((kinds[:, 1] != 0) == (ix == n - 1))
        assert set(np.unique(g.kind[g.kind != 0])) == ({1} if periodic else {2})
        _mirror_ok(g)


def test_source_assignment_is_nearest_cell_summed():
    # set_source_terms_system (src/source_systems.rs:48-80): each source's rate
    # goes to the nearest cell by plain Euclidean distance (no periodic
    # wrapping), rates of sources in one cell add up; brute force on random
    # points (no ties)
    import numpy as np
    from subsweep_amd import assign_sources
    rng = np.random.default_rng(7)
    pts = rng.uniform(0.0, 10.0, (500, 3))
    pos = np.vstack([rng.uniform(-1.0, 11.0, (300, 3)), pts[:5] + 1e-9])  # some outside the box, some on cells
    rates = rng.uniform(1e48, 1e50, pos.shape[0])
    src, cells = assign_sources(pts, pos, rates)
    d2 = ((pos[:, None, :] - pts[None, :, :]) ** 2).sum(axis=2)
    want = d2.argmin(axis=1)
    assert np.array_equal(np.asarray(cells), want)
    ref = np.zeros(pts.shape[0])
    np.add.at(ref, want, rates)
    np.testing.assert_allclose(src, ref, rtol=1e-15)
