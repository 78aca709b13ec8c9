"""Request loop over the C ABI's server (ssw_server_*): independent solves of
different grids run concurrently on the device, each in its own handle and
HIP stream, admitted in submission order against an HBM budget.

    with Server(slots=4, hbm_budget_bytes=200 << 30) as sv:
        t = sv.submit(problem, steps=10, grid_key=7)
        r = sv.wait(t)        # RequestResult: status, times, outputs

With `max_batch` >= 2 the running requests are stepped together: one step
of each per iteration, their grouped launches packed into shared launches
(`ssw_step_batch`), with bitwise the same results.

A non-zero `grid_key` keeps the request's grid resident: its handle (grid,
dependency graph, schedule) stays cached on the device after the request,
and a later request of the same key and parameters reuses it.

The reference runs one simulation per process and has no counterpart; each
solve is one `Sweep` life cycle (init_sweep_system, `steps` x run_sweeps,
the copy-out of run_sweep_system, src/sweep/mod.rs:634-739).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .workloads import Problem


@dataclass
class RequestResult:
    ticket: int
    status: int
    error: str
    transport_mode: int
    submit_s: float
    start_s: float
    end_s: float
    updates: int
    device_bytes: int
    estimated_bytes: int
    reused: bool
    ionized_hydrogen_fraction: np.ndarray
    temperature: np.ndarray
    timestep: np.ndarray
    photon_rate: np.ndarray
    level: np.ndarray
    change_timescale: np.ndarray

    @property
    def latency_s(self) -> float:
        return self.end_s - self.submit_s


def estimate_device_bytes(problem: Problem, transport_mode: int) -> int:
    pc, keep = problem.params.to_c()
    g = problem.grid
    nper = 0
    if problem.params.periodic:
        cell_of_face = np.repeat(np.arange(g.num_cells), np.diff(g.offsets))
        nper = int(np.unique(cell_of_face[g.kind == N.FACE_PERIODIC]).size)
    return int(N.sweep_lib().ssw_estimate_device_bytes(C.byref(pc), g.num_cells, g.num_faces, nper, transport_mode))


class Server:
    def __init__(self, slots: int = 4, hbm_budget_bytes: int = 200 << 30, max_batch: int = 0):
        """`max_batch` >= 2: requests are stepped together, up to max_batch per
        packed step (`ssw_server_create_packed`); else each on its own worker."""
        self._lib = N.sweep_lib()
        h = C.c_void_p()
        N.check(self._lib.ssw_server_create_packed(int(slots), int(hbm_budget_bytes), int(max_batch), C.byref(h)))
        self._h = h
        self._pending = {}

    def submit(self, problem: Problem, steps: int, transport_mode: int = N.TRANSPORT_MODE_AUTO,
               replay_block: int = 0, grid_key: int = 0, direction_sweeps: bool = False) -> int:
        g = problem.grid
        n = g.num_cells
        arrays = list(g.arrays()) + [np.ascontiguousarray(a, dtype=np.float64) for a in problem.state()]
        out = {k: np.empty(n) for k in ("x", "T", "ts", "pr", "ch")}
        out["lev"] = np.empty(n, np.uint8)
        pc, keep = problem.params.to_c()
        r = N.RequestC()
        r.params = pc
        r.transport_mode = int(transport_mode)
        r.steps = int(steps)
        r.replay_block = int(replay_block)
        r.direction_sweeps = 1 if direction_sweeps else 0
        r.num_cells = n
        r.grid_key = int(grid_key)
        names = ["face_offsets", "face_neighbour", "face_kind", "face_area", "face_normal", "cell_volume",
                 "cell_size", "density", "ionized_fraction", "temperature", "source"]
        for name, a in zip(names, arrays):
            setattr(r, name, a.ctypes.data)
        if g.points is not None and len(g.points) == n:  # locality order from the generator positions
            pts = np.ascontiguousarray(g.points, dtype=np.float64).reshape(-1, 3)
            arrays.append(pts)
            r.cell_position = pts.ctypes.data
        r.out_ionized_fraction = out["x"].ctypes.data
        r.out_temperature = out["T"].ctypes.data
        r.out_timestep = out["ts"].ctypes.data
        r.out_photon_rate = out["pr"].ctypes.data
        r.out_level = out["lev"].ctypes.data
        r.out_change_timescale = out["ch"].ctypes.data
        t = C.c_int64()
        N.check(self._lib.ssw_server_submit(self._h, C.byref(r), C.byref(t)))
        self._pending[t.value] = (arrays, out, keep)  # buffers stay alive until wait()
        return t.value

    def wait(self, ticket: int) -> RequestResult:
        res = N.RequestResultC()
        N.check(self._lib.ssw_server_wait(self._h, int(ticket), C.byref(res)))
        _, out, _ = self._pending.pop(ticket)
        return RequestResult(ticket, res.status, res.error.decode(errors="replace"), res.transport_mode,
                             res.submit_s, res.start_s, res.end_s, res.updates, res.device_bytes,
                             res.estimated_bytes, bool(res.reused), out["x"], out["T"], out["ts"], out["pr"],
                             out["lev"], out["ch"])

    def stats(self):
        b, n = C.c_int64(), C.c_int32()
        N.check(self._lib.ssw_server_stats(self._h, C.byref(b), C.byref(n)))
        nb, ns = C.c_int64(), C.c_int64()
        N.check(self._lib.ssw_server_pack_stats(self._h, C.byref(nb), C.byref(ns)))
        return {"peak_reserved_bytes": b.value, "peak_running": n.value, "packed_steps": nb.value,
                "packed_requests_per_step": ns.value / nb.value if nb.value else 0.0}

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.ssw_server_destroy(self._h)
            self._h = None
            self._pending.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
