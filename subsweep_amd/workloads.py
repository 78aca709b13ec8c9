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
