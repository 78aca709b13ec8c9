"""Comparison helpers shared by the GPU parity tests (not a test module)."""

from __future__ import annotations

import numpy as np


def rel_ok(a, b, rtol, atol_frac=0.0, name="", atol_abs=0.0):
    """|a - b| <= rtol * max(|a|, |b|) + atol_frac * max|b| + atol_abs, elementwise."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = np.maximum(np.abs(a), np.abs(b))
    atol = atol_frac * (np.abs(b).max() if b.size else 0.0) + atol_abs
    err = np.abs(a - b)
    bad = err > rtol * scale + atol
    if bad.any():
        i = np.argmax(np.where(bad, err / np.maximum(scale, 1e-300), 0))
        raise AssertionError(f"{name}: {bad.sum()} of {a.size} outside rtol {rtol}; worst at {i}: {a.flat[i]!r} vs "
                             f"{b.flat[i]!r}")


def max_rel(a, b, floor=0.0):
    """Largest relative difference where max(|a|, |b|) > floor (0 if none)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = np.maximum(np.abs(a), np.abs(b))
    keep = scale > floor
    if not keep.any():
        return 0.0
    return float((np.abs(a - b)[keep] / scale[keep]).max())


def summed_rate(sweep, source, num_directions):
    """The rate update_chemistry hands to the solver (src/sweep/mod.rs:554-558):
    sum over directions, in direction order, of (incoming + source / D) +
    periodic source, from a sweep's (or the oracle's) per-(direction, cell)
    fluxes after a step."""
    inc, _, per = sweep.fluxes()
    sod = np.asarray(source, np.float64) / float(num_directions)
    rate = np.zeros(inc.shape[1])
    for d in range(num_directions):
        rate = rate + ((inc[d] + sod) + per[d])
    return rate


def limiting_counts(info):
    return {"temperature": int(info.limiting_temperature),
            "ionization_fraction": int(info.limiting_ionization_fraction),
            "photon_rate": int(info.limiting_photon_rate)}
