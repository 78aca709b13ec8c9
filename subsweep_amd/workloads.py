"""The synthetic workloads the sweep is measured and tested on.

  jittered_problem   n^3 jittered-lattice points (+-0.3 cell, seeded) in a
                     periodic 4000 kpc box, Voronoi grid, lognormal density
                     (mean 1e-27 g/cm^3), T = 1e4 K, x_HII = 1e-10, point
                     sources mapped to their nearest cells.
  precision_64       64^3 = 262144 cells, sigma_ln 1.5, 64 sources at uniform
                     positions with luminosities 10^U(48, 51) photons/s,
                     84 directions, 1 level, threshold 1e-5 s^-1.
  single_source_32   32^3 = 32768 cells, sigma_ln 1.0, one 1e50 photons/s
                     source at the box centre, 16 directions, 1 level.
  clustered_problem  points sampled from a lognormal field (Gaussian random
                     field, P(k) ~ k^-2, smoothed on 2 mesh cells, 128^3
                     mesh, sigma_ln 2), cell density 1e-27 g/cm^3 times the
                     field at the generator, sources on the densest cells
                     with luminosities from p(L) ~ L^-2 on [1e48, 1e52]
                     photons/s times an escape fraction of 0.01.
  clustered_2m       2097152 clustered cells, 1024 sources, 84 directions,
                     4 timestep levels, FP32 transport.
  cartesian_problem  the reference's cartesian test grid with uniform gas.
  serving_pool,      the request mix of the serving loop: grids of 8192-524288
  serving_requests   cells (log-uniform), jittered or clustered, with random
                     direction counts, timestep levels and sources.

Seeds: lattice 1338, density 1339, sources 1337, clustered field 1340/1341,
clustered sources 1342.
"""

from __future__ import annotations

from dataclasses import dataclass, replace

import numpy as np

from . import units as U
from .grid import (Grid, assign_sources, cartesian_grid, clustered_points, jittered_lattice, lognormal_density,
                   voronoi_grid)
from .sweep import SweepParameters

BOX = 4000.0 * U.KILOPARSEC
RHO_MEAN = 1e-27 * U.GRAMS_PER_CUBIC_CENTIMETERS


@dataclass
class Problem:
    name: str
    grid: Grid
    params: SweepParameters
    density: np.ndarray
    ionized_fraction: np.ndarray
    temperature: np.ndarray
    source: np.ndarray
    steps: int

    def state(self):
        return self.density, self.ionized_fraction, self.temperature, self.source

    def with_params(self, **kw) -> "Problem":
        return replace(self, params=replace(self.params, **kw))


def jittered_problem(n: int, sigma_ln: float, sources: str, num_directions: int, levels: int = 1,
                     threshold: float = 0.0, periodic: bool = True, steps: int = 10, seed: int = 1338,
                     precision: int = 64, name: str | None = None, threads: int = 0,
                     rho_mean: float = RHO_MEAN) -> Problem:
    """`rho_mean` scales the gas: the default makes a 64^3 grid's cells
    optically thick (tau ~ 200-1100 at x = 1e-10); 1e-3 of it gives tau ~ 1
    per cell on a 16^3 grid, where photons cross the box."""
    pts = jittered_lattice(n, BOX, 0.3, seed)
    grid = voronoi_grid(pts, BOX, threads=threads)
    N = grid.num_cells
    rho = lognormal_density(N, rho_mean, sigma_ln, seed + 1)
    rng = np.random.default_rng(seed - 1)
    if sources == "centre":
        pos = np.array([[BOX / 2, BOX / 2, BOX / 2]])
        rates = np.array([1e50])
    else:
        k = int(sources)
        pos = rng.uniform(0.0, BOX, size=(k, 3))
        rates = 10.0 ** rng.uniform(48.0, 51.0, size=k)
    src, _ = assign_sources(pts, pos, rates)
    params = SweepParameters(
        directions=num_directions, num_timestep_levels=levels, periodic=periodic,
        max_timestep=1.0 * U.MEGAYEARS, significant_rate_threshold=threshold,
        transport_precision=precision,
    )
    return Problem(name or f"jittered{n}", grid, params, rho, np.full(N, 1e-10), np.full(N, 1e4), src, steps)


def single_source_32(**kw) -> Problem:
    kw.setdefault("steps", 10)
    return jittered_problem(32, 1.0, "centre", 16, name="single_source_32", **kw)


def precision_64(**kw) -> Problem:
    kw.setdefault("threshold", 1e-5)
    kw.setdefault("steps", 20)
    return jittered_problem(64, 1.5, "64", 84, name="precision_64", **kw)


def power_law_luminosities(k: int, lo: float, hi: float, seed: int) -> np.ndarray:
    """k draws from p(L) ~ L^-2 on [lo, hi] by inverting the CDF
    F(L) = (1/lo - 1/L) / (1/lo - 1/hi)."""
    u = np.random.default_rng(seed).random(k)
    return 1.0 / (1.0 / lo - u * (1.0 / lo - 1.0 / hi))


def clustered_problem(n: int, num_directions: int = 84, levels: int = 4, sources: int = 1024,
                      sigma_ln: float = 2.0, mesh: int = 128, escape_fraction: float = 0.01,
                      threshold: float = 0.0, periodic: bool = True, steps: int = 10, precision: int = 32,
                      timestep_safety_factor: float = 0.1, name: str | None = None, threads: int = 0,
                      rho_mean: float = RHO_MEAN) -> Problem:
    pts, delta = clustered_points(n, BOX, mesh=mesh, sigma_ln=sigma_ln, seed=1340)
    grid = voronoi_grid(pts, BOX, threads=threads)
    N = grid.num_cells
    rho = rho_mean * delta
    # the densest cells (ties by index) each get one source
    order = np.lexsort((np.arange(N), -rho))
    src = np.zeros(N)
    src[order[:sources]] = power_law_luminosities(sources, 1e48, 1e52, 1342) * escape_fraction
    params = SweepParameters(
        directions=num_directions, num_timestep_levels=levels, periodic=periodic,
        max_timestep=1.0 * U.MEGAYEARS, significant_rate_threshold=threshold,
        timestep_safety_factor=timestep_safety_factor, transport_precision=precision,
    )
    return Problem(name or f"clustered{n}", grid, params, rho, np.full(N, 1e-10), np.full(N, 1e4), src, steps)


def clustered_2m(**kw) -> Problem:
    return clustered_problem(2097152, name="clustered_2m", **kw)


def cartesian_problem(n: int, num_directions: int, levels: int = 1, periodic: bool = False,
                      source_cells=(0,), source_rate: float = 1e50, box: float = 1e5, steps: int = 3,
                      density: float = 1e-10 * 1e-3 / 1e-6, temperature: float = 1000.0,
                      max_timestep: float = 1e-3, threshold: float = 0.0, prevent_cooling: bool = True,
                      precision: int = 64) -> Problem:
    """Uniform gas on the cartesian grid; defaults follow the reference's sweep
    test components (src/sweep/mod.rs:783-797: 1e-10 kg/cm^3, x 1e-10, 1000 K)."""
    grid = cartesian_grid(n, box, periodic)
    N = grid.num_cells
    src = np.zeros(N)
    for c in source_cells:
        src[c] += source_rate
    params = SweepParameters(
        directions=num_directions, num_timestep_levels=levels, periodic=periodic, max_timestep=max_timestep,
        significant_rate_threshold=threshold, prevent_cooling=prevent_cooling, transport_precision=precision,
    )
    return Problem(f"cartesian{n}", grid, params, np.full(N, density), np.full(N, 1e-10), np.full(N, temperature),
                   src, steps)


def serving_pool(num_grids: int, seed: int, lo: int = 8192, hi: int = 524288):
    """Grids of the serving workload: sizes log-uniform in [lo, hi] cells,
    alternately jittered-lattice Voronoi (n^3 cells, n = round(cbrt(size)))
    and clustered lognormal-field Voronoi (64^3 field mesh, sigma_ln 2), as
    (grid, generator points, field density or None)."""
    rng = np.random.default_rng(seed)
    pool = []
    for i in range(num_grids):
        size = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        if i % 2 == 0:
            n = max(2, int(round(size ** (1.0 / 3.0))))
            pts = jittered_lattice(n, BOX, 0.3, seed + 10 + i)
            delta = None
        else:
            pts, delta = clustered_points(size, BOX, mesh=64, sigma_ln=2.0, seed=seed + 10 + i)
        pool.append((voronoi_grid(pts, BOX), pts, delta))
    return pool


def serving_requests(pool, count: int, seed: int, steps: int, precision: int = 32, rho_scale: float = 1.0):
    """Request mix: a pool grid, directions from {16, 32, 64, 84}, timestep
    levels from {1, 2, 4}, 1-64 sources of 1e50 photons/s at uniform positions,
    lognormal gas (sigma_ln 1) on jittered grids and the field's density on
    clustered grids (both times `rho_scale`), `steps` steps. Each Problem
    carries `pool_index`, the index of its grid in the pool (a key for keeping
    the grid resident)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        gi = int(rng.integers(len(pool)))
        grid, pts, delta = pool[gi]
        N = grid.num_cells
        dirs = int(rng.choice([16, 32, 64, 84]))
        levels = int(rng.choice([1, 2, 4]))
        ns = int(rng.integers(1, 65))
        src, _ = assign_sources(pts, rng.uniform(0.0, BOX, size=(ns, 3)), np.full(ns, 1e50))
        rho = RHO_MEAN * delta if delta is not None else lognormal_density(N, RHO_MEAN, 1.0, seed + 1000 + k)
        rho = rho * rho_scale
        params = SweepParameters(directions=dirs, num_timestep_levels=levels, periodic=True,
                                 max_timestep=1.0 * U.MEGAYEARS, transport_precision=precision)
        pb = Problem(f"request{k}", grid, params, rho, np.full(N, 1e-10), np.full(N, 1e4), src, steps)
        pb.pool_index = gi
        out.append(pb)
    return out
