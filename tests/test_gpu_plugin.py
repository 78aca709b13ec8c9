"""GPU: the plugin's step system (run_sweep_system, src/sweep/mod.rs:694-739):
the first call only clears IsFirstTime; later calls equal a direct Sweep
step by step; the ionisation time of a cell is the simulation time at the
first step that leaves its ionised fraction above 0.5 and +inf (the
component's default) until then."""

import numpy as np
import pytest

from subsweep_amd import Sweep, SweepPlugin
from subsweep_amd import workloads as W

pytestmark = pytest.mark.gpu

FIELDS = ("ionized_hydrogen_fraction", "temperature", "timestep", "photon_rate", "level", "change_timescale")


def test_plugin_step_system(hip_available):
    # a 6^3 box of 1e20 m with 1e-24 kg m^-3 gas: the ionised region grows
    # from the source cell over the six steps
    pb = W.cartesian_problem(6, 16, levels=2, periodic=False, source_cells=(0,), steps=6, max_timestep=3e12,
                             density=1e-24, box=1e20, temperature=1e4)
    plugin = SweepPlugin(Sweep(pb.params, pb.grid, *pb.state()))
    direct = Sweep(pb.params, pb.grid, *pb.state())
    assert plugin.run_sweep_system() is None  # the first call outputs the initial conditions
    t = 0.0
    first = np.full(pb.grid.num_cells, np.inf)
    for _ in range(pb.steps):
        got = plugin.run_sweep_system()
        t += direct.run_sweeps()
        want = direct.results()
        for f in FIELDS:
            assert np.array_equal(getattr(got, f), getattr(want, f)), f
        first[(want.ionized_hydrogen_fraction > 0.5) & (first == np.inf)] = t
        assert plugin.time == t
        assert np.array_equal(plugin.ionization_time, first)
    assert np.isfinite(first).any(), "no cell ionised: the test does not exercise the ionisation time"
    assert np.isinf(first).any(), "every cell ionised: the default value is not exercised"
    plugin.sweep.close()
    direct.close()


def test_plugin_optional_chemistry_outputs(hip_available):
    # the optional outputs named in optional_outputs are evaluated after each
    # step, as the plugin's sweep_optional_output_system instances
    pb = W.cartesian_problem(6, 16, levels=1, periodic=False, source_cells=(0,), steps=2, max_timestep=3e12,
                             density=1e-24, box=1e20, temperature=1e4)
    names = ("heating_rate", "photoionization_rate")
    plugin = SweepPlugin(Sweep(pb.params, pb.grid, *pb.state()), optional_outputs=names)
    direct = Sweep(pb.params, pb.grid, *pb.state())
    # the first call skips the sweep but the output systems still run: the
    # outputs of the initial state
    assert plugin.run_sweep_system() is None
    want = direct.chemistry_outputs()
    for k in names:
        assert np.array_equal(plugin.chemistry[k], want[k]), k
    assert (plugin.chemistry["photoionization_rate"] > 0).any()  # the source cell's own rate
    for _ in range(pb.steps):
        plugin.run_sweep_system()
        direct.run_sweeps()
        want = direct.chemistry_outputs()
        assert tuple(plugin.chemistry) == names
        for k in names:
            assert np.array_equal(plugin.chemistry[k], want[k]), k
    assert (plugin.chemistry["photoionization_rate"] > 0).any()
    with pytest.raises(ValueError):
        SweepPlugin(direct, optional_outputs=("cooling",))
    plugin.sweep.close()
    direct.close()


def test_plugin_time_series_against_oracle(tmp_path, hip_available):
    # the time series written after each step (src/sweep/time_series.rs:61-188)
    # equal, within 1e-10, the same series computed from the oracle's state
    from oracle.oracle import OracleSweep
    from subsweep_amd import time_series as TS
    pb = W.cartesian_problem(6, 16, levels=2, periodic=False, source_cells=(0,), steps=4, max_timestep=3e12,
                             density=1e-24, box=1e20, temperature=1e4)
    density = pb.state()[0]
    volume = pb.grid.arrays()[5]
    plugin = SweepPlugin(Sweep(pb.params, pb.grid, *pb.state()))
    out = TS.TimeSeriesOutput(str(tmp_path), volume, density, pb.params.max_timestep, pb.params.num_timestep_levels)
    o = OracleSweep(pb.params, pb.grid, *pb.state())
    assert out.write_plugin_step(plugin) is None  # nothing before the plugin's first call
    for step in range(pb.steps + 1):
        plugin.run_sweep_system()  # the first call: the initial conditions at time 0
        got = out.write_plugin_step(plugin)
        if step > 0:
            o.run_sweeps()
        ro = o.results()
        want = TS.compute(ro.ionized_hydrogen_fraction, ro.temperature, o.chemistry_outputs()["photoionization_rate"],
                          ro.level, volume, density, pb.params.max_timestep, pb.params.num_timestep_levels)
        for k in TS.AVERAGES:
            assert got[k] == pytest.approx(want[k], rel=1e-10, abs=0), k
        assert got[TS.LEVELS] == want[TS.LEVELS]
    assert got["photoionization_rate_volume_average"] > 0
    entries = TS.read(str(tmp_path), TS.LEVELS)
    assert len(entries) == pb.steps + 1 and entries[0]["time"] == 0.0
    assert entries[0]["val"][-1]["num"] == pb.grid.num_cells  # every cell starts on the deepest level
    plugin.sweep.close()
    o.close()
