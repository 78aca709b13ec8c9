"""ctypes wrapper of the oracle (liboracle_sweep.so): the CPU restatement of
the reference's sweep + chemistry. TEST INFRASTRUCTURE ONLY - imported by
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg to compare
results against, never by the product.

`OracleSweep` has the same methods as subsweep_amd.Sweep so that a test can
drive both with the same calls. The direction bins of
DirectionsSpecification::Num come from tests/golden/reference_tables.json
(the reference's tables, src/sweep/direction/healpix.rs, read as data), not
from the product.

`threads` spreads a sweep's directions and the per-cell loops over threads;
each direction and each cell is computed exactly as with one thread, so the
results are bitwise the same for every count."""

from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle_sweep.so")
TABLES_PATH = os.path.join(os.path.dirname(HERE), "tests", "golden", "reference_tables.json")

PD = C.POINTER(C.c_double)
PI64 = C.POINTER(C.c_int64)
PI32 = C.POINTER(C.c_int32)
PU8 = C.POINTER(C.c_uint8)
PU32 = C.POINTER(C.c_uint32)
P = C.c_void_p

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build as _b
            _b.build()
        from subsweep_amd._native import SweepParametersC
        L = C.CDLL(LIB_PATH)
        sig = {
            "orc_create": (P, [C.POINTER(SweepParametersC), PD]),
            "orc_set_threads": (None, [P, C.c_int]),
            "orc_set_source": (C.c_int, [P, PD]),
            "orc_limiting_processes": (None, [P, PI64]),
            "orc_solver_terms": (None, [C.c_double] * 8 + [PD]),
            "orc_solver_try_timestep": (C.c_double, [PD] + [C.c_double] * 5 + [C.c_int] + [C.c_double] * 4 +
                                        [C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "orc_destroy": (None, [P]),
            "orc_upload_grid": (C.c_int, [P, C.c_int64, PI64, PI32, PU8, PD, PD, PD, PD]),
            "orc_upload_state": (C.c_int, [P, PD, PD, PD, PD]),
            "orc_step": (C.c_int, [P, PD]),
            "orc_download": (C.c_int, [P, PD, PD, PD, PD, PU8, PD]),
            "orc_download_counts": (C.c_int, [P, C.c_int32, PU32]),
            "orc_download_levelsets": (C.c_int, [P, C.c_int32, PU32]),
            "orc_download_fluxes": (C.c_int, [P, PD, PD, PD]),
            "orc_download_chemistry_outputs": (C.c_int, [P, PD, PD, PD, PD]),
            "orc_chemistry_outputs_cell": (None, [C.c_double] * 7 + [PD]),
            "orc_updates_last_step": (C.c_int64, [P]),
            "orc_chemistry_failures": (C.c_int64, [P]),
            "orc_num_directions": (C.c_int, [P]),
            "orc_directions": (None, [P, PD]),
            "orc_rotate_directions": (C.c_int, [P, PD]),
            "orc_last_error": (C.c_char_p, []),
            "orc_solver_rates": (None, [C.c_double, C.c_double, C.c_double, C.c_double, PD]),
            "orc_solver_perform_timestep": (C.c_double, [PD, C.c_double, C.c_double, C.c_double, C.c_double,
                                                         C.c_double, C.c_int, C.c_double, C.c_double,
                                                         C.POINTER(C.c_int), C.POINTER(C.c_int)]),
            "orc_timestep_level": (C.c_int, [C.c_int, C.c_double, C.c_double]),
            "orc_levels_in_sweep_order": (C.c_int, [C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                                    C.POINTER(C.c_int), C.c_int]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


_tables = None


def reference_tables() -> dict:
    """tests/golden/reference_tables.json: the reference's direction bins,
    BPASS table and test values, as data."""
    global _tables
    if _tables is None:
        import json
        with open(TABLES_PATH) as f:
            _tables = json.load(f)
    return _tables


def tabulated_directions(num: int) -> np.ndarray:
    """The bins of Directions::from_num(num) (src/sweep/direction/mod.rs:58-75):
    [1, 0, 0] for one direction, else the HEALPix table of the reference."""
    if num == 1:
        return np.array([[1.0, 0.0, 0.0]])
    t = reference_tables()["directions"]
    if str(num) not in t:
        raise ValueError(f"no direction table for {num} directions")
    return np.array(t[str(num)]["bins"], dtype=np.float64)


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}: {lib().orc_last_error().decode()}")


RATE_NAMES = [
    "case_b_recombination_rate", "case_b_recombination_rate_derivative",
    "collisional_ionization_rate", "collisional_ionization_rate_derivative",
    "case_b_recombination_cooling_rate", "case_b_recombination_cooling_rate_derivative",
    "collisional_ionization_cooling_rate", "collisional_ionization_cooling_rate_derivative",
    "collisional_excitation_cooling_rate", "collisional_excitation_cooling_rate_derivative",
    "bremsstrahlung_cooling_rate", "bremsstrahlung_cooling_rate_derivative",
    "compton_cooling_rate", "compton_cooling_rate_derivative",
    "cooling_rate", "cooling_rate_derivative",
]


def solver_rates(temperature, xhii=0.0, density=0.0, scale_factor=1.0) -> dict:
    out = np.empty(16)
    lib().orc_solver_rates(temperature, xhii, density, scale_factor, _p(out, C.c_double))
    return dict(zip(RATE_NAMES, out))


CHEMISTRY_OUTPUT_NAMES = ["photoionization_rate", "heating_rate", "recombination_rate", "collisional_ionization_rate"]


def chemistry_outputs_cell(x, T, density, volume, length, rate, scale_factor=1.0) -> dict:
    """The optional chemistry outputs of one cell (src/sweep/chemistry_output.rs:25-56,
    solver without floor, timestep 1 yr)."""
    out = np.empty(4)
    lib().orc_chemistry_outputs_cell(x, T, density, volume, length, rate, scale_factor, _p(out, C.c_double))
    return dict(zip(CHEMISTRY_OUTPUT_NAMES, out))


def perform_timestep(x, T, density, volume, length, rate, timestep, safety=0.1, scale_factor=1.0, floor=False):
    st = np.array([x, T], dtype=np.float64)
    proc = C.c_int()
    failed = C.c_int()
    ts = lib().orc_solver_perform_timestep(_p(st, C.c_double), density, volume, length, rate, scale_factor,
                                           1 if floor else 0, timestep, safety, C.byref(proc), C.byref(failed))
    return {"x": st[0], "T": st[1], "timescale": ts, "process": proc.value, "failed": bool(failed.value)}


TERM_NAMES = ["temperature_change", "photoheating_rate", "cooling_rate", "cooling_rate_derivative", "mu",
              "ionized_fraction_change", "photoionization_rate"]


def solver_terms(x, T, density, volume, length, rate, timestep, scale_factor=1.0) -> dict:
    """The terms of one implicit Solver update (no floor) at (x, T)."""
    out = np.empty(7)
    lib().orc_solver_terms(x, T, density, volume, length, rate, scale_factor, timestep, _p(out, C.c_double))
    return dict(zip(TERM_NAMES, out))


def try_timestep(x, T, density, volume, length, rate, timestep, safety=0.1, scale_factor=1.0, floor=None):
    """One attempt of the sub-stepping at depth 0 (clamp + try_timestep_update):
    {accepted, x, T, timescale, process}; on rejection x, T are the clamped
    starting state. floor: None or (floor_x, floor_T)."""
    st = np.array([x, T], dtype=np.float64)
    acc, proc = C.c_int(), C.c_int()
    fx, fT = floor if floor is not None else (0.0, 0.0)
    ts = lib().orc_solver_try_timestep(_p(st, C.c_double), density, volume, length, rate, scale_factor,
                                       0 if floor is None else 1, fx, fT, timestep, safety, C.byref(acc),
                                       C.byref(proc))
    return {"accepted": bool(acc.value), "x": st[0], "T": st[1], "timescale": ts, "process": proc.value}


def timestep_level(max_num_levels, max_timestep, desired):
    return lib().orc_timestep_level(max_num_levels, max_timestep, desired)


def levels_in_sweep_order(num_levels: int, calls: int):
    lowest = C.c_int(num_levels - 1)
    first = C.c_int(0)
    out = (C.c_int * 64)()
    res = []
    for _ in range(calls):
        n = lib().orc_levels_in_sweep_order(num_levels, C.byref(lowest), C.byref(first), out, 64)
        res.append([out[i] for i in range(n)])
    return res


class OracleSweep:
    def __init__(self, params, grid, density, ionized_fraction, temperature, source, threads: int = 1):
        L = lib()
        pc, keep = params.to_c()
        table = None
        if isinstance(params.directions, (int, np.integer)):
            table = np.ascontiguousarray(tabulated_directions(int(params.directions)))
        self._h = L.orc_create(C.byref(pc), _p(table, C.c_double))
        if not self._h:
            raise RuntimeError("oracle: bad parameters")
        L.orc_set_threads(self._h, int(threads))
        self.num_cells = grid.num_cells
        self.num_directions = L.orc_num_directions(self._h)
        off, nbr, kind, area, normal, vol, size = grid.arrays()
        _check(L.orc_upload_grid(self._h, grid.num_cells, _p(off, C.c_int64), _p(nbr, C.c_int32),
                                 _p(kind, C.c_uint8), _p(area, C.c_double), _p(normal, C.c_double),
                                 _p(vol, C.c_double), _p(size, C.c_double)))
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (density, ionized_fraction, temperature, source)]
        _check(L.orc_upload_state(self._h, *[_p(a, C.c_double) for a in arrs]))

    def run_sweeps(self) -> float:
        el = C.c_double()
        _check(lib().orc_step(self._h, C.byref(el)))
        return el.value

    def results(self, photon_rate=True):
        from subsweep_amd.sweep import SweepResults
        n = self.num_cells
        x, T, ts, pr, ch = (np.empty(n) for _ in range(5))
        lev = np.empty(n, np.uint8)
        _check(lib().orc_download(self._h, _p(x, C.c_double), _p(T, C.c_double), _p(ts, C.c_double),
                                  _p(pr, C.c_double), _p(lev, C.c_uint8), _p(ch, C.c_double)))
        return SweepResults(x, T, ts, pr, lev, ch)

    def counts(self, sweep_level):
        out = np.empty(self.num_directions * self.num_cells, np.uint32)
        _check(lib().orc_download_counts(self._h, sweep_level, _p(out, C.c_uint32)))
        return out.reshape(self.num_directions, self.num_cells)

    def levelsets(self, sweep_level):
        out = np.empty(self.num_directions * self.num_cells, np.uint32)
        _check(lib().orc_download_levelsets(self._h, sweep_level, _p(out, C.c_uint32)))
        return out.reshape(self.num_directions, self.num_cells)

    def fluxes(self):
        m = self.num_directions * self.num_cells
        i, o, p = np.empty(m), np.empty(m), np.empty(m)
        _check(lib().orc_download_fluxes(self._h, _p(i, C.c_double), _p(o, C.c_double), _p(p, C.c_double)))
        s = (self.num_directions, self.num_cells)
        return i.reshape(s), o.reshape(s), p.reshape(s)

    def chemistry_outputs(self) -> dict:
        out = {k: np.empty(self.num_cells) for k in CHEMISTRY_OUTPUT_NAMES}
        _check(lib().orc_download_chemistry_outputs(self._h, *(_p(v, C.c_double) for v in out.values())))
        return out

    def set_source(self, source):
        """Replaces the source rates (test hook; the reference sets sources
        once at startup)."""
        a = np.ascontiguousarray(source, dtype=np.float64)
        _check(lib().orc_set_source(self._h, _p(a, C.c_double)))

    def set_threads(self, threads: int):
        lib().orc_set_threads(self._h, int(threads))

    def limiting_processes(self) -> dict:
        """TimescaleCounter of the last step: cells per limiting process whose
        change timescale is below the maximum timestep."""
        out = np.zeros(3, np.int64)
        lib().orc_limiting_processes(self._h, _p(out, C.c_int64))
        return {"temperature": int(out[0]), "ionization_fraction": int(out[1]), "photon_rate": int(out[2])}

    def rotate_directions(self, matrix):
        """rotate_directions_system (src/sweep/direction/mod.rs:158-175) with
        the given 3 x 3 rotation matrix."""
        m = np.ascontiguousarray(matrix, dtype=np.float64).reshape(9)
        _check(lib().orc_rotate_directions(self._h, _p(m, C.c_double)))

    def directions(self):
        out = np.empty(3 * self.num_directions)
        lib().orc_directions(self._h, _p(out, C.c_double))
        return out.reshape(-1, 3)

    def updates_last_step(self):
        return lib().orc_updates_last_step(self._h)

    def chemistry_failures(self):
        return lib().orc_chemistry_failures(self._h)

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
