"""The sweep's time series (src/sweep/time_series.rs:61-188) and their YAML
files (src/io/time_series.rs:100-133), on hand-made cell states (CPU): the
weighted averages against sums worked by hand, the per-level counts of
active cells, and the files read back entry by entry."""

from types import SimpleNamespace

import numpy as np
import pytest

from subsweep_amd import time_series as TS


def _state():
    # two cells: volumes 1 and 3 m^3, densities 4 and 1 kg/m^3 (masses 4 and 3 kg)
    return dict(x=np.array([1.0, 0.0]), T=np.array([100.0, 200.0]), photoionization_rate=np.array([2.0, 6.0]),
                level=np.array([0, 2], np.uint8), volume=np.array([1.0, 3.0]), density=np.array([4.0, 1.0]))


def test_averages_by_hand():
    s = _state()
    v = TS.compute(**s, max_timestep=8.0, num_levels=3)
    assert v["hydrogen_ionization_mass_average"] == pytest.approx(4.0 / 7.0, rel=1e-15, abs=0)
    assert v["hydrogen_ionization_volume_average"] == pytest.approx(1.0 / 4.0, rel=1e-15, abs=0)
    assert v["temperature_mass_average"] == pytest.approx((400.0 + 600.0) / 7.0, rel=1e-15, abs=0)
    assert v["temperature_volume_average"] == pytest.approx((100.0 + 600.0) / 4.0, rel=1e-15, abs=0)
    assert v["photoionization_rate_volume_average"] == pytest.approx((2.0 + 18.0) / 4.0, rel=1e-15, abs=0)
    assert v["weighted_photoionization_rate_volume_average"] == pytest.approx(2.0 / 4.0, rel=1e-15, abs=0)
    # active cells per level: level >= 0: both, >= 1: the level-2 cell, >= 2: the level-2 cell
    assert v[TS.LEVELS] == [{"level": 0, "num": 2, "timestep": 8.0}, {"level": 1, "num": 1, "timestep": 4.0},
                            {"level": 2, "num": 1, "timestep": 2.0}]


def test_uniform_state_averages_are_the_state():
    rng = np.random.default_rng(3)
    n = 1000
    v = TS.compute(np.full(n, 0.25), np.full(n, 1e4), np.full(n, 3e-13), np.zeros(n, np.uint8),
                   rng.uniform(0.1, 10, n), rng.uniform(1e-28, 1e-24, n), 3.15576e13, 1)
    assert v["hydrogen_ionization_mass_average"] == pytest.approx(0.25, rel=1e-13, abs=0)
    assert v["temperature_volume_average"] == pytest.approx(1e4, rel=1e-13, abs=0)
    assert v["weighted_photoionization_rate_volume_average"] == pytest.approx(0.25 * 3e-13, rel=1e-13, abs=0)


def test_files_round_trip(tmp_path):
    s = _state()
    out = TS.TimeSeriesOutput(str(tmp_path), s["volume"], s["density"], 8.0, 3)
    for name in TS.NAMES:
        assert (tmp_path / f"{name}.yml").read_text() == ""  # created empty at start-up
    wrote = []
    for step, t in enumerate((1.5, 3.0)):
        res = SimpleNamespace(ionized_hydrogen_fraction=s["x"] * (step + 1) / 2, temperature=s["T"], level=s["level"])
        wrote.append(out.write(t, res, s["photoionization_rate"]))
    for name in TS.NAMES:
        entries = TS.read(str(tmp_path), name)
        assert [e["time"] for e in entries] == [1.5, 3.0]
        assert [e["val"] for e in entries] == [w[name] for w in wrote]
