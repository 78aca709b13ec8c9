"""Host-side mirror of the reference's sweep plugin interface over the C ABI.

Reference names and meaning are kept:

* `SweepParameters` - the `sweep` parameter section (src/sweep/parameters.rs:8-46),
  defaults as in :64-78, plus the cosmological scale factor handed to
  HydrogenOnly (src/sweep/mod.rs:685-690) and the transport precision.
* `Sweep` - the device-resident `Sweep<HydrogenOnly>` resource
  (src/sweep/mod.rs:171-191): built from the grid and the initial per-cell
  components (init_sweep_system, :634-692); `run_sweeps()` is one
  `Sweep::run_sweeps` call (:258-272) and returns the elapsed time.
* `SweepPlugin` / `run_sweep_system` - the ECS step wrapper: it skips its very
  first call (IsFirstTime, :708-714), advances the simulation time and copies
  ionised fraction, temperature, timestep, photon rate and ionisation time
  back (:715-738).

Errors raise `SweepError` with the library's message (the reference panics).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native as N
from .grid import Grid


@dataclass
class SweepParameters:
    directions: int | list = 84
    num_timestep_levels: int = 1
    periodic: bool = False
    max_timestep: float = 1e6 * 3.15576e7
    rotate_directions: bool = False
    significant_rate_threshold: float = 0.0
    timestep_safety_factor: float = 0.01 * 10.0
    chemistry_timestep_safety_factor: float = 0.01 * 10.0
    check_deadlock: bool = False
    prevent_cooling: bool = True
    num_tasks_to_solve_before_send_receive: int = 10000
    # not in the reference's section:
    scale_factor: float = 1.0
    transport_precision: int = 64

    def to_c(self):
        p = N.SweepParametersC()
        keep = None
        if isinstance(self.directions, (int, np.integer)):
            p.num_directions = int(self.directions)
            p.explicit_directions = None
        else:
            keep = np.ascontiguousarray(np.asarray(self.directions, dtype=np.float64).reshape(-1, 3))
            p.num_directions = keep.shape[0]
            p.explicit_directions = keep.ctypes.data_as(C.POINTER(C.c_double))
        if self.rotate_directions:
            raise N.SweepError(N.SSW_ERR_INVALID, "rotate_directions is not supported")
        p.num_timestep_levels = int(self.num_timestep_levels)
        p.periodic = 1 if self.periodic else 0
        p.max_timestep = float(self.max_timestep)
        p.significant_rate_threshold = float(self.significant_rate_threshold)
        p.timestep_safety_factor = float(self.timestep_safety_factor)
        p.chemistry_timestep_safety_factor = float(self.chemistry_timestep_safety_factor)
        p.prevent_cooling = 1 if self.prevent_cooling else 0
        p.transport_precision = int(self.transport_precision)
        p.scale_factor = float(self.scale_factor)
        return p, keep


def direction_table(num_directions: int) -> np.ndarray:
    """The [num_directions, 3] direction table of DirectionsSpecification::Num
    (src/sweep/direction/mod.rs:58-75), as the library uses it."""
    out = np.zeros((int(num_directions), 3), dtype=np.float64)
    N.check(N.sweep_lib().ssw_direction_table(int(num_directions), N.ptr(out, C.c_double)))
    return out


def rotation_matrix(axis, angle: float) -> np.ndarray:
    """get_rotation_matrix (src/sweep/direction/mod.rs:112-135): 3 x 3
    rotation by `angle` about the unit vector `axis`."""
    a = np.ascontiguousarray(axis, dtype=np.float64).reshape(3)
    out = np.zeros(9)
    N.check(N.sweep_lib().ssw_rotation_matrix(N.ptr(a, C.c_double), float(angle), N.ptr(out, C.c_double)))
    return out.reshape(3, 3)


def random_rotation_matrix(rng: np.random.Generator) -> np.ndarray:
    """get_random_rotation_matrix (src/sweep/direction/mod.rs:137-148) with its
    three draws (phi in [0, 2 pi), u in [0, 1), psi in [0, 2 pi)) taken from
    `rng`. The reference draws them from a StdRng seeded with 1337, whose
    stream numpy does not reproduce."""
    phi = rng.uniform(0.0, 2.0 * np.pi)
    u = rng.uniform(0.0, 1.0)
    psi = rng.uniform(0.0, 2.0 * np.pi)
    out = np.zeros(9)
    N.check(N.sweep_lib().ssw_random_rotation_matrix(phi, u, psi, N.ptr(out, C.c_double)))
    return out.reshape(3, 3)


@dataclass
class SweepResults:
    ionized_hydrogen_fraction: np.ndarray
    temperature: np.ndarray
    timestep: np.ndarray
    photon_rate: np.ndarray
    level: np.ndarray
    change_timescale: np.ndarray


class Sweep:
    """Device-resident sweep state of one problem (one HIP stream)."""

    def __init__(self, params: SweepParameters, grid: Grid, density, ionized_fraction, temperature, source,
                 transport_mode: int = N.TRANSPORT_MODE_AUTO, replay_block: int = 0, graph_replay: bool = True,
                 locality_order: bool = True, narrow_runs: bool = True, group_levels: int = 0,
                 group_block: int = 0, direction_sweeps: bool = False):
        lib = N.sweep_lib()
        self._lib = lib
        self.params = params
        pc, keep = params.to_c()
        h = C.c_void_p()
        N.check(lib.ssw_create(C.byref(pc), C.byref(h)))
        del keep
        self._h = h
        self.num_cells = grid.num_cells
        try:
            N.check(lib.ssw_set_transport_mode(h, int(transport_mode)))
            N.check(lib.ssw_set_replay_block(h, int(replay_block)))
            N.check(lib.ssw_set_graph_replay(h, 1 if graph_replay else 0))
            N.check(lib.ssw_set_narrow_runs(h, 1 if narrow_runs else 0))
            if group_levels:  # 0: the library's default
                N.check(lib.ssw_set_group_levels(h, int(group_levels)))
            N.check(lib.ssw_set_group_block(h, int(group_block)))
            N.check(lib.ssw_set_direction_sweeps(h, 1 if direction_sweeps else 0))
            if locality_order and grid.points is not None:
                pts = np.ascontiguousarray(grid.points, dtype=np.float64).reshape(-1, 3)
                if pts.shape[0] == grid.num_cells:
                    N.check(lib.ssw_set_cell_positions(h, grid.num_cells, N.ptr(pts, C.c_double)))
            off, nbr, kind, area, normal, vol, size = grid.arrays()
            N.check(lib.ssw_upload_grid(h, grid.num_cells, N.ptr(off, C.c_int64), N.ptr(nbr, C.c_int32),
                                        N.ptr(kind, C.c_uint8), N.ptr(area, C.c_double), N.ptr(normal, C.c_double),
                                        N.ptr(vol, C.c_double), N.ptr(size, C.c_double)))
            self.upload_state(density, ionized_fraction, temperature, source)
        except Exception:
            self.close()
            raise
        self.num_directions = self.info().num_directions

    def upload_state(self, density, ionized_fraction, temperature, source):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (density, ionized_fraction, temperature, source)]
        for a in arrs:
            if a.shape != (self.num_cells,):
                raise N.SweepError(N.SSW_ERR_INVALID, "state arrays must have one entry per cell")
        N.check(self._lib.ssw_upload_state(self._h, *[N.ptr(a, C.c_double) for a in arrs]))

    def run_sweeps(self) -> float:
        el = C.c_double()
        N.check(self._lib.ssw_step(self._h, C.byref(el)))
        return el.value

    @staticmethod
    def run_sweeps_packed(sweeps) -> list:
        """One run_sweeps of every handle in `sweeps`, their grouped launches
        packed into shared launches (`ssw_step_batch`); returns the elapsed
        times. Bitwise the same results as stepping each alone."""
        sweeps = list(sweeps)
        if not sweeps:
            return []
        hs = (C.c_void_p * len(sweeps))(*[s._h.value for s in sweeps])
        el = np.zeros(len(sweeps))
        N.check(sweeps[0]._lib.ssw_step_batch(hs, len(sweeps), N.ptr(el, C.c_double)))
        return [float(e) for e in el]

    @staticmethod
    def packed_graph_stats() -> dict:
        """Graphs of packed sweeps captured and graph launches so far."""
        c, r = C.c_int64(), C.c_int64()
        N.check(N.sweep_lib().ssw_packed_graph_stats(C.byref(c), C.byref(r)))
        return {"captured": c.value, "replayed": r.value}

    def results(self, photon_rate: bool = True, noalloc: bool = False) -> SweepResults:
        """The per-cell outputs (`ssw_download`; with `noalloc`,
        `ssw_download_noalloc`, which stages them through the library's fixed
        per-device buffer instead of allocating)."""
        n = self.num_cells
        x, T, ts, pr, ch = (np.empty(n) for _ in range(5))
        lev = np.empty(n, np.uint8)
        fn = self._lib.ssw_download_noalloc if noalloc else self._lib.ssw_download
        N.check(fn(self._h, N.ptr(x, C.c_double), N.ptr(T, C.c_double), N.ptr(ts, C.c_double),
                   N.ptr(pr, C.c_double) if photon_rate else None, N.ptr(lev, C.c_uint8), N.ptr(ch, C.c_double)))
        return SweepResults(x, T, ts, pr if photon_rate else None, lev, ch)

    def chemistry_outputs(self) -> dict:
        """The optional outputs of sweep_optional_output_system
        (src/sweep/chemistry_output.rs:25-67) for every cell: photoionization,
        heating, recombination and collisional ionization rates (SI), from the
        solver of Sweep::get_solver (src/sweep/mod.rs:612-632) with dt = 1 yr."""
        n = self.num_cells
        out = {k: np.empty(n) for k in ("photoionization_rate", "heating_rate", "recombination_rate",
                                         "collisional_ionization_rate")}
        N.check(self._lib.ssw_download_chemistry_outputs(self._h, *(N.ptr(v, C.c_double) for v in out.values())))
        return out

    def counts(self, sweep_level: int) -> np.ndarray:
        out = np.empty(self.num_directions * self.num_cells, np.uint32)
        N.check(self._lib.ssw_download_counts(self._h, sweep_level, N.ptr(out, C.c_uint32)))
        return out.reshape(self.num_directions, self.num_cells)

    def levelsets(self, sweep_level: int) -> np.ndarray:
        out = np.empty(self.num_directions * self.num_cells, np.uint32)
        N.check(self._lib.ssw_download_levelsets(self._h, sweep_level, N.ptr(out, C.c_uint32)))
        return out.reshape(self.num_directions, self.num_cells)

    def fluxes(self):
        m = self.num_directions * self.num_cells
        i, o, p = np.empty(m), np.empty(m), np.empty(m)
        N.check(self._lib.ssw_download_fluxes(self._h, N.ptr(i, C.c_double), N.ptr(o, C.c_double),
                                              N.ptr(p, C.c_double)))
        shape = (self.num_directions, self.num_cells)
        return i.reshape(shape), o.reshape(shape), p.reshape(shape)

    def rotate_directions(self, matrix) -> None:
        """rotate_directions_system (src/sweep/direction/mod.rs:158-175) with
        the given 3 x 3 rotation matrix: rotated directions, their dependency
        graph, and the rates remapped as the reference's remap does."""
        m = np.ascontiguousarray(matrix, dtype=np.float64).reshape(9)
        N.check(self._lib.ssw_rotate_directions(self._h, N.ptr(m, C.c_double)))

    def synchronize(self) -> None:
        N.check(self._lib.ssw_synchronize(self._h))

    def info(self) -> N.InfoC:
        inf = N.InfoC()
        N.check(self._lib.ssw_get_info(self._h, C.byref(inf)))
        return inf

    def set_transport_mode(self, mode: int) -> None:
        """Replay form (TRANSPORT_MODE_AUTO / _GEOMETRIC / _COMPILED). Only
        valid before the grid is uploaded, so pass `transport_mode` to the
        constructor; 1 and 2 (a removed persistent kernel) raise."""
        N.check(self._lib.ssw_set_transport_mode(self._h, int(mode)))

    def transport_state(self):
        """Raw X (outgoing rate / downwind area) and I in their storage
        precision and scale, shape (directions, cells)."""
        dt = np.float32 if self.params.transport_precision == 32 else np.float64
        m = self.num_directions * self.num_cells
        x, i = np.empty(m, dt), np.empty(m, dt)
        N.check(self._lib.ssw_download_transport(self._h, x.ctypes.data_as(C.c_void_p),
                                                 i.ctypes.data_as(C.c_void_p)))
        shape = (self.num_directions, self.num_cells)
        return x.reshape(shape), i.reshape(shape)

    def set_timing(self, enabled: bool) -> None:
        N.check(self._lib.ssw_set_timing(self._h, 1 if enabled else 0))

    def timing(self):
        avg, tot, chem, build = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        n = C.c_int64()
        N.check(self._lib.ssw_get_timing(self._h, C.byref(avg), C.byref(tot), C.byref(n), C.byref(chem),
                                         C.byref(build)))
        return {"transport_avg_ms": avg.value, "transport_total_ms": tot.value, "transport_launches": n.value,
                "chemistry_total_ms": chem.value, "build_total_ms": build.value}

    def save_checkpoint(self) -> bytes:
        """The dynamic state as a checkpoint (`ssw_save_checkpoint`)."""
        n = int(self._lib.ssw_checkpoint_bytes(self._h))
        buf = np.zeros(n, dtype=np.uint8)
        N.check(self._lib.ssw_save_checkpoint(self._h, buf.ctypes.data_as(C.c_void_p), n))
        return buf.tobytes()

    def load_checkpoint(self, data: bytes) -> None:
        """Replaces the dynamic state with a checkpoint of the same problem."""
        buf = np.frombuffer(data, dtype=np.uint8).copy()
        N.check(self._lib.ssw_load_checkpoint(self._h, buf.ctypes.data_as(C.c_void_p), buf.size))

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.ssw_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


CHEMISTRY_OUTPUTS = ("photoionization_rate", "heating_rate", "recombination_rate", "collisional_ionization_rate")


@dataclass
class SweepPlugin:
    """The plugin's step system over a `Sweep`: run_sweep_system
    (src/sweep/mod.rs:694-739). The first call only marks IsFirstTime and
    returns None: the components keep the initial conditions, which `last`
    then holds (the state the time series of that frame are computed from,
    src/sweep/mod.rs:150-160).

    `optional_outputs` names the optional chemistry components
    (src/components.rs:50-68) to evaluate after every call, the first one
    included, as the sweep_optional_output_system instances that the plugin
    registers for the desired output fields and that run after
    run_sweep_system each frame (src/sweep/mod.rs:162-165, 770-781); their
    values land in `chemistry`.

    `rotate_directions` (SweepParameters::rotate_directions,
    src/sweep/parameters.rs:20-22) rotates the directions by a random rotation
    after every call, the first one included, as rotate_directions_system runs
    after run_sweep_system each frame (src/sweep/mod.rs:143-148); the draws
    come from `rotation_rng` (numpy, seeded with the reference's 1337)."""

    sweep: Sweep
    time: float = 0.0
    is_first_time: bool = True
    ionization_time: np.ndarray | None = None
    last: SweepResults | None = field(default=None, repr=False)
    optional_outputs: tuple = ()
    chemistry: dict = field(default_factory=dict, repr=False)
    rotate_directions: bool = False
    rotation_rng: np.random.Generator = field(default_factory=lambda: np.random.default_rng(1337), repr=False)

    def __post_init__(self):
        unknown = [k for k in self.optional_outputs if k not in CHEMISTRY_OUTPUTS]
        if unknown:
            raise ValueError(f"unknown optional chemistry outputs {unknown}; known: {CHEMISTRY_OUTPUTS}")

    def run_sweep_system(self) -> SweepResults | None:
        res = self._run_sweep_system()
        if self.rotate_directions:
            self.sweep.rotate_directions(random_rotation_matrix(self.rotation_rng))
        return res

    def _run_sweep_system(self) -> SweepResults | None:
        if self.is_first_time:
            self.is_first_time = False
            self.last = self.sweep.results()
            self._optional_outputs()
            return None
        self.time += self.sweep.run_sweeps()
        res = self.sweep.results()
        if self.ionization_time is None:  # IonizationTime::default() is +inf (src/components.rs:79-82)
            self.ionization_time = np.full(res.ionized_hydrogen_fraction.shape, np.inf)
        newly = (res.ionized_hydrogen_fraction > 0.5) & (self.ionization_time == np.inf)
        self.ionization_time[newly] = self.time
        self.last = res
        self._optional_outputs()
        return res

    def _optional_outputs(self) -> None:
        if self.optional_outputs:
            out = self.sweep.chemistry_outputs()
            self.chemistry = {k: out[k] for k in self.optional_outputs}
