"""Grids and synthetic inputs for the sweep.

A `Grid` is the reference's per-cell `grid::Cell` (src/sweep/grid/cell.rs:92-97)
flattened into CSR arrays: for cell c, faces off[c]..off[c+1] with neighbour
index, kind (FACE_LOCAL / FACE_PERIODIC / FACE_BOUNDARY, the single-rank
ParticleType variants), area and unit normal; plus cell volume and size.

Builders (host C++ in csrc/gridgen.cpp):
  cartesian_grid   the reference's cartesian constructor on one rank
  voronoi_grid     periodic 3-D Voronoi tessellation of given points
and the input generators of the workloads (jittered lattice points, lognormal
densities, nearest-cell source assignment).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N


@dataclass
class Grid:
    offsets: np.ndarray    # int64 [N+1]
    neighbour: np.ndarray  # int32 [F]
    kind: np.ndarray       # uint8 [F]
    area: np.ndarray       # float64 [F]
    normal: np.ndarray     # float64 [F, 3]
    volume: np.ndarray     # float64 [N]
    size: np.ndarray       # float64 [N]
    points: np.ndarray | None = None  # float64 [N, 3] generator positions
    dropped_faces: int = 0

    @property
    def num_cells(self) -> int:
        return int(self.volume.shape[0])

    @property
    def num_faces(self) -> int:
        return int(self.neighbour.shape[0])

    def arrays(self):
        return (
            np.ascontiguousarray(self.offsets, dtype=np.int64),
            np.ascontiguousarray(self.neighbour, dtype=np.int32),
            np.ascontiguousarray(self.kind, dtype=np.uint8),
            np.ascontiguousarray(self.area, dtype=np.float64),
            np.ascontiguousarray(self.normal, dtype=np.float64).reshape(-1),
            np.ascontiguousarray(self.volume, dtype=np.float64),
            np.ascontiguousarray(self.size, dtype=np.float64),
        )


def cartesian_grid(n: int, box: float, periodic: bool) -> Grid:
    """n^3 cells of side box/n (src/sweep/grid/cartesian.rs:231-395). Cell index
    (ix * n + iy) * n + iz; faces in the order -x, +x, -y, +y, -z, +z."""
    lib = N.grid_lib()
    ncell = n * n * n
    off = np.empty(ncell + 1, np.int64)
    nbr = np.empty(6 * ncell, np.int32)
    kind = np.empty(6 * ncell, np.uint8)
    area = np.empty(6 * ncell, np.float64)
    normal = np.empty(18 * ncell, np.float64)
    vol = np.empty(ncell, np.float64)
    size = np.empty(ncell, np.float64)
    lib.ssw_grid_cartesian(
        n, box, 1 if periodic else 0, N.ptr(off, C.c_int64), N.ptr(nbr, C.c_int32), N.ptr(kind, C.c_uint8),
        N.ptr(area, C.c_double), N.ptr(normal, C.c_double), N.ptr(vol, C.c_double), N.ptr(size, C.c_double),
    )
    h = box / n
    idx = np.arange(n, dtype=np.float64)
    ix, iy, iz = np.meshgrid(idx, idx, idx, indexing="ij")
    pts = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], axis=1) * h
    return Grid(off, nbr, kind, area, normal.reshape(-1, 3), vol, size, points=pts)


def voronoi_grid(points: np.ndarray, box, threads: int = 0, free_space_box=None) -> Grid:
    """Periodic Voronoi tessellation of `points` in [0, box)^3. Faces to a
    periodic image are FACE_PERIODIC. With `free_space_box=(lo, hi)` the
    tessellation is instead clipped to that box and its walls are
    FACE_BOUNDARY (no periodic images)."""
    lib = N.grid_lib()
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    n = pts.shape[0]
    if np.isscalar(box):
        box = (box, box, box)
    boxa = np.ascontiguousarray(box, dtype=np.float64)
    mode = 0
    lohi = None
    if free_space_box is not None:
        mode = 1
        lohi = np.ascontiguousarray(np.concatenate([free_space_box[0], free_space_box[1]]), dtype=np.float64)
    lib.ssw_grid_set_progress(1 if n >= (1 << 22) else 0)  # long builds report progress
    h = lib.ssw_voronoi_build(n, N.ptr(pts, C.c_double), N.ptr(boxa, C.c_double), mode,
                              N.ptr(lohi, C.c_double) if lohi is not None else None, threads)
    try:
        if mode == 0 and lib.ssw_voronoi_open_cells(h):
            raise RuntimeError("Voronoi construction left open cells")
        F = lib.ssw_voronoi_num_faces(h)
        off = np.empty(n + 1, np.int64)
        nbr = np.empty(F, np.int32)
        kind = np.empty(F, np.uint8)
        area = np.empty(F, np.float64)
        normal = np.empty(3 * F, np.float64)
        vol = np.empty(n, np.float64)
        size = np.empty(n, np.float64)
        lib.ssw_voronoi_get(h, N.ptr(off, C.c_int64), N.ptr(nbr, C.c_int32), N.ptr(kind, C.c_uint8),
                            N.ptr(area, C.c_double), N.ptr(normal, C.c_double), N.ptr(vol, C.c_double),
                            N.ptr(size, C.c_double))
        dropped = int(lib.ssw_voronoi_dropped_faces(h))
    finally:
        lib.ssw_voronoi_free(h)
    return Grid(off, nbr, kind, area, normal.reshape(-1, 3), vol, size, points=pts, dropped_faces=dropped)


def jittered_lattice(n: int, box: float, jitter: float, seed: int) -> np.ndarray:
    """n^3 points ((i + 1/2 + u) h), u ~ U(-jitter, jitter) per axis (SplitMix64
    stream per point), wrapped into [0, box)^3; index (iz * n + iy) * n + ix."""
    out = np.empty(3 * n ** 3, np.float64)
    N.grid_lib().ssw_points_jittered_lattice(n, box, jitter, seed, N.ptr(out, C.c_double))
    return out.reshape(-1, 3)


def uniform_points(n: int, box: float, seed: int) -> np.ndarray:
    out = np.empty(3 * n, np.float64)
    N.grid_lib().ssw_points_uniform(n, box, seed, N.ptr(out, C.c_double))
    return out.reshape(-1, 3)


def lognormal_density(n: int, rho_mean: float, sigma_ln: float, seed: int) -> np.ndarray:
    """rho = rho_mean exp(sigma g - sigma^2/2), g ~ N(0, 1) i.i.d. per cell."""
    out = np.empty(n, np.float64)
    N.grid_lib().ssw_lognormal_density(n, rho_mean, sigma_ln, seed, N.ptr(out, C.c_double))
    return out


def gaussian_random_field(mesh: int, slope: float, smoothing: float, seed: int) -> np.ndarray:
    """Seeded Gaussian random field on a periodic mesh^3 grid with power
    spectrum P(k) ~ k^slope, Gaussian-smoothed with radius `smoothing` mesh
    cells, normalised to zero mean and unit variance (numpy PCG64 stream)."""
    rng = np.random.default_rng(seed)
    white = rng.standard_normal((mesh, mesh, mesh))
    fk = np.fft.rfftn(white)
    kx = np.fft.fftfreq(mesh)[:, None, None]
    ky = np.fft.fftfreq(mesh)[None, :, None]
    kz = np.fft.rfftfreq(mesh)[None, None, :]
    k2 = kx * kx + ky * ky + kz * kz  # cycles per mesh cell
    with np.errstate(divide="ignore"):
        amp = np.where(k2 > 0, k2 ** (slope / 4.0), 0.0)  # sqrt(P) = k^(slope/2)
    amp = amp * np.exp(-0.5 * (2.0 * np.pi) ** 2 * k2 * smoothing ** 2)
    g = np.fft.irfftn(fk * amp, s=(mesh, mesh, mesh), axes=(0, 1, 2))
    g -= g.mean()
    g /= g.std()
    return g


def clustered_points(n: int, box: float, mesh: int = 128, sigma_ln: float = 2.0, slope: float = -2.0,
                     smoothing: float = 2.0, seed: int = 1340):
    """n points sampled with probability proportional to the lognormal field
    delta = exp(sigma g - sigma^2 / 2) of a Gaussian random field g (above) on a
    mesh^3 grid over the periodic box: mesh cells are drawn from the CDF of
    delta, positions uniform within the drawn cell. Returns (points [n, 3],
    delta of each point's mesh cell)."""
    g = gaussian_random_field(mesh, slope, smoothing, seed)
    delta = np.exp(sigma_ln * g - 0.5 * sigma_ln * sigma_ln).reshape(-1)
    cdf = np.cumsum(delta)
    cdf /= cdf[-1]
    rng = np.random.default_rng(seed + 1)
    cell = np.minimum(np.searchsorted(cdf, rng.random(n), side="right"), delta.size - 1)
    ix = cell // (mesh * mesh)
    iy = (cell // mesh) % mesh
    iz = cell % mesh
    h = box / mesh
    pts = (np.stack([ix, iy, iz], axis=1) + rng.random((n, 3))) * h
    pts = np.minimum(pts, np.nextafter(box, 0.0))
    return pts, delta[cell]


def assign_sources(points: np.ndarray, positions: np.ndarray, rates: np.ndarray, device: bool = False):
    """Adds each source's rate to its nearest cell (Euclidean, non-periodic),
    as set_source_terms_system does (src/source_systems.rs:48-80). Returns
    (per-cell source rate, cell index of each source). `device`: the nearest
    cells are found on the GPU (`ssw_assign_sources_device`), with bitwise the
    same results as the host search."""
    pts = np.ascontiguousarray(points, dtype=np.float64).reshape(-1, 3)
    pos = np.ascontiguousarray(positions, dtype=np.float64).reshape(-1, 3)
    rt = np.ascontiguousarray(rates, dtype=np.float64).reshape(-1)
    src = np.empty(pts.shape[0], np.float64)
    cells = np.empty(pos.shape[0], np.int64)
    if device:
        N.check(N.sweep_lib().ssw_assign_sources_device(pts.shape[0], N.ptr(pts, C.c_double), pos.shape[0],
                                                        N.ptr(pos, C.c_double), N.ptr(rt, C.c_double),
                                                        N.ptr(src, C.c_double), N.ptr(cells, C.c_int64)))
        return src, cells
    rc = N.grid_lib().ssw_assign_sources(pts.shape[0], N.ptr(pts, C.c_double), pos.shape[0],
                                         N.ptr(pos, C.c_double), N.ptr(rt, C.c_double),
                                         N.ptr(src, C.c_double), N.ptr(cells, C.c_int64))
    if rc != 0:
        raise RuntimeError("source assignment failed")
    return src, cells
