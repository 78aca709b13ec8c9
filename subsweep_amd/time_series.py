"""The sweep's time series and their YAML output.

Per step the reference sends these events (src/sweep/time_series.rs:61-188)
and appends each to `<time_series_dir>/<name>.yml` as a YAML sequence of
`{time, val}` entries (src/io/time_series.rs:29-33, 100-133):

  hydrogen_ionization_mass_average          sum(m x) / sum(m)
  hydrogen_ionization_volume_average        sum(V x) / sum(V)
  temperature_mass_average                  sum(T m) / sum(m)
  temperature_volume_average                sum(T V) / sum(V)
  photoionization_rate_volume_average       sum(Gamma V) / sum(V)
  weighted_photoionization_rate_volume_average  sum(Gamma x V) / sum(V)
  num_particles_at_timestep_levels          [{level, num, timestep}] for every level,
                                            num = cells on that level or above

Gamma is the PhotoionizationRate component (src/sweep/chemistry_output.rs:25-29),
here from `Sweep.chemistry_outputs`. The cell mass is density x volume. Values
are SI floats; the reference writes diman quantities through serde, whose text
form (units, float formatting) is not reproduced.
"""

from __future__ import annotations

import os

import numpy as np
import yaml

AVERAGES = (
    "hydrogen_ionization_mass_average",
    "hydrogen_ionization_volume_average",
    "temperature_mass_average",
    "temperature_volume_average",
    "photoionization_rate_volume_average",
    "weighted_photoionization_rate_volume_average",
)
LEVELS = "num_particles_at_timestep_levels"
NAMES = AVERAGES + (LEVELS,)


def compute(x, T, photoionization_rate, level, volume, density, max_timestep: float, num_levels: int) -> dict:
    """The time-series values of one step (compute_time_series_system and
    num_particles_at_timestep_levels_system, src/sweep/time_series.rs:61-188)."""
    x, T, gam = (np.asarray(a, np.float64) for a in (x, T, photoionization_rate))
    vol = np.asarray(volume, np.float64)
    mass = np.asarray(density, np.float64) * vol
    total_mass, total_volume = mass.sum(), vol.sum()
    level = np.asarray(level)
    counts = np.bincount(level.astype(np.int64), minlength=num_levels)
    return {
        "hydrogen_ionization_mass_average": float((mass * x).sum() / total_mass),
        "hydrogen_ionization_volume_average": float((vol * x).sum() / total_volume),
        "temperature_mass_average": float((T * mass).sum() / total_mass),
        "temperature_volume_average": float((T * vol).sum() / total_volume),
        "photoionization_rate_volume_average": float((gam * vol).sum() / total_volume),
        "weighted_photoionization_rate_volume_average": float((gam * x * vol).sum() / total_volume),
        # TimestepState::iter_all_levels (src/sweep/timestep_state.rs:33-35); the count of a level is
        # Sweep::count_cells_global (src/sweep/mod.rs:234-238), the cells active on it, i.e. those
        # of level >= it (ActiveList::enumerate_active, src/sweep/active_list.rs:55-66);
        # TimestepLevel::to_timestep = max_timestep * 0.5^level (src/sweep/timestep_level.rs:42-48)
        LEVELS: [{"level": int(lv), "num": int(counts[lv:].sum()), "timestep": float(max_timestep * 0.5 ** lv)}
                 for lv in range(num_levels)],
    }


class TimeSeriesOutput:
    """Writes the time series of a `SweepPlugin` after each of its steps:
    the files are created empty at construction (initialize_output_files_system)
    and every `write` appends one entry per series (output_time_series_system)."""

    def __init__(self, directory: str, volume, density, max_timestep: float, num_levels: int):
        self.directory = directory
        self.volume = np.asarray(volume, np.float64)
        self.density = np.asarray(density, np.float64)
        self.max_timestep = float(max_timestep)
        self.num_levels = int(num_levels)
        os.makedirs(directory, exist_ok=True)
        for name in NAMES:
            open(self.path(name), "w").close()

    def path(self, name: str) -> str:
        return os.path.join(self.directory, f"{name}.yml")

    def write(self, time: float, results, photoionization_rate) -> dict:
        vals = compute(results.ionized_hydrogen_fraction, results.temperature, photoionization_rate, results.level,
                       self.volume, self.density, self.max_timestep, self.num_levels)
        for name in NAMES:
            with open(self.path(name), "a") as f:
                yaml.safe_dump([{"time": float(time), "val": vals[name]}], f, sort_keys=False)
        return vals

    def write_plugin_step(self, plugin) -> dict | None:
        """After each plugin.run_sweep_system(). The reference's series
        systems run every frame (AfterSweep, src/sweep/mod.rs:150-160), the
        first one too, whose run_sweep_system skips the sweep: that frame
        writes the initial conditions at the plugin's time, 0. None before
        the plugin's first call."""
        if plugin.last is None:
            return None
        return self.write(plugin.time, plugin.last, plugin.sweep.chemistry_outputs()["photoionization_rate"])


def read(directory: str, name: str) -> list:
    """The entries of one series file (a YAML sequence of {time, val})."""
    out = []
    with open(os.path.join(directory, f"{name}.yml")) as f:
        for doc in yaml.safe_load_all(f):
            out.extend(doc or [])
    return out
