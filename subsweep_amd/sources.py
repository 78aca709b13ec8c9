"""Source rates of star particles, the reference's Arepo input path
(src/arepo_postprocess/sources.rs:57-128): each star particle becomes a point
source whose rate is the BPASS ionising photon rate of its stellar
population (bpass_lookup, src/arepo_postprocess/bpass.rs:11-56) times an
escape fraction; its age runs from its formation scale factor to the
current one in a flat cosmology (src/cosmology.rs:80-107). Particles with a
formation scale factor that is not positive are wind particles and no
source. The sources then go to their nearest cells (grid.assign_sources,
src/source_systems.rs:48-80). The arithmetic is host C++ in
libsubsweep_amd.so (csrc/sources.cpp)."""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N

YEAR = 3.15576e7  # s (src/units/mod.rs:48)
SOLAR_MASS = 1.988477e30  # kg (src/units/mod.rs:55)


def bpass_source_rate(age_years, metallicity, mass_solar):
    """Ionising photons/s of stellar populations (vectorised over numpy arrays)."""
    f = N.sweep_lib().ssw_bpass_source_rate
    a, z, m = np.broadcast_arrays(np.asarray(age_years, np.float64), np.asarray(metallicity, np.float64),
                                  np.asarray(mass_solar, np.float64))
    out = np.array([f(float(x), float(y), float(w)) for x, y, w in zip(a.ravel(), z.ravel(), m.ravel())])
    return out.reshape(a.shape) if a.shape else float(out[0])


def time_between_scale_factors(a0: float, a1: float, omega_lambda: float, omega_0: float, h: float) -> float:
    """Seconds from scale factor a0 to a1 (flat cosmology; NaN otherwise)."""
    return float(N.sweep_lib().ssw_time_between_scale_factors(a0, a1, omega_lambda, omega_0, h))


def star_sources(metallicity, formation_scale_factor, mass_solar, escape_fraction: float, omega_lambda: float,
                 omega_0: float, h: float, scale_factor: float):
    """(rate, is_star) per particle: rate in photons/s (0 for wind particles)."""
    z = np.ascontiguousarray(metallicity, np.float64)
    a = np.ascontiguousarray(formation_scale_factor, np.float64)
    m = np.ascontiguousarray(mass_solar, np.float64)
    n = z.shape[0]
    if a.shape != (n,) or m.shape != (n,):
        raise N.SweepError(N.SSW_ERR_INVALID, "particle arrays must have the same length")
    rate = np.zeros(n)
    star = np.zeros(n, np.uint8)
    r = N.sweep_lib().ssw_star_sources(n, N.ptr(z, C.c_double), N.ptr(a, C.c_double), N.ptr(m, C.c_double),
                                        float(escape_fraction), float(omega_lambda), float(omega_0), float(h),
                                        float(scale_factor), N.ptr(rate, C.c_double), N.ptr(star, C.c_uint8))
    if r < 0:
        N.check(int(r))
    return rate, star.astype(bool)
