"""Timestep levels and the level cadence (CPU).

Pins: the reference's own tables (tests/golden/reference_known_answers.json,
from src/sweep/timestep_level.rs:66-89 and src/sweep/timestep_state.rs:98-137)
for the oracle and for the library's host-compiled level function; and the
library's exact ceil(log2) against the oracle's libm log2 + ceil on random
ratios and on ratios a few ulps around powers of two."""

import json
import os

import numpy as np

from oracle import oracle as O
from subsweep_amd import _native as N

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")))


def test_oracle_timestep_level_table():
    for levels, desired, expect in GOLD["timestep_level_table"]["cases"]:
        assert O.timestep_level(levels, 1.0, desired) == expect, (levels, desired)


def test_library_timestep_level_table():
    lib = N.sweep_lib()
    for levels, desired, expect in GOLD["timestep_level_table"]["cases"]:
        assert lib.ssw_timestep_level(levels, 1.0, desired) == expect, (levels, desired)


def test_oracle_sweep_order_cadence():
    g = GOLD["levels_in_sweep_order"]
    assert O.levels_in_sweep_order(g["num_levels"], len(g["calls"])) == g["calls"]


def test_library_level_matches_libm_log2():
    lib = N.sweep_lib()
    rng = np.random.default_rng(7)
    ratios = list(10.0 ** rng.uniform(-5, 12, 20000))
    for k in range(-3, 40):  # around powers of two, a few ulps either side
        p = 2.0 ** k
        v = p
        for _ in range(6):
            v = np.nextafter(v, np.inf)
            ratios.append(v)
        v = p
        for _ in range(6):
            v = np.nextafter(v, -np.inf)
            ratios.append(v)
        ratios.append(p)
    ratios += [0.0, np.inf, np.nan, -1.0, 1e-320, 1e308]
    for L in (1, 2, 4, 31):
        for r in ratios:
            desired = 1.0 / r if r != 0 else np.inf
            if r == np.inf:
                desired = 0.0
            a = lib.ssw_timestep_level(L, 1.0, desired)
            b = O.timestep_level(L, 1.0, desired)
            assert a == b, (L, r, a, b)


def test_level_saturation_cases():
    lib = N.sweep_lib()
    # desired 0 -> ratio inf -> top level; desired inf -> ratio 0 -> level 0;
    # NaN -> 0 (Rust's saturating `as usize`)
    assert lib.ssw_timestep_level(5, 1.0, 0.0) == 4
    assert lib.ssw_timestep_level(5, 1.0, np.inf) == 0
    assert lib.ssw_timestep_level(5, 1.0, np.nan) == 0
    assert lib.ssw_timestep_level(5, 1.0, -1.0) == 0
