"""Launch-shape autotuner for one problem.

The transport's throughput on a given grid depends on how its level sets are
laid out and launched:

* direction groups - how many blocks of directions run their level sets side
  by side on separate streams (more groups overlap more of each level's
  latency, fewer keep each launch wide);
* transport mode - compiled replay (one lane per task, upwind entries stored
  per task) or geometric replay (16 lanes per task over the face lists).

`autotune` times a few steps of each candidate on the problem itself and
returns the fastest. Every candidate computes bitwise the same results, so
the choice affects speed only.
"""

from __future__ import annotations

import itertools
import time
from dataclasses import dataclass

from . import _native as N
from .sweep import Sweep
from .workloads import Problem


@dataclass
class TuneResult:
    direction_groups: int
    transport_mode: int
    ms_per_step: float
    table: list  # (groups, mode, ms per step) of every candidate

    def kwargs(self) -> dict:
        return {"direction_groups": self.direction_groups, "transport_mode": self.transport_mode}


def autotune(problem: Problem, groups=(1, 2, 4), modes=(N.TRANSPORT_MODE_COMPILED, N.TRANSPORT_MODE_GEOMETRIC),
             steps: int = 2, warmup: int = 1) -> TuneResult:
    """Times `steps` steps after `warmup` steps of every (groups, mode)
    candidate; candidates that do not fit in device memory are skipped."""
    table = []
    d = problem.params.directions
    ndir = d if isinstance(d, int) else len(d)
    for g, m in itertools.product(groups, modes):
        if g > ndir:
            continue
        try:
            s = Sweep(problem.params, problem.grid, *problem.state(), transport_mode=m, direction_groups=g)
        except N.SweepError as e:
            if e.code == N.SSW_ERR_OOM:
                continue
            raise
        try:
            for _ in range(warmup):
                s.run_sweeps()
            s.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                s.run_sweeps()
            s.synchronize()
            table.append((g, m, (time.perf_counter() - t0) / steps * 1e3))
        finally:
            s.close()
    if not table:
        raise RuntimeError("no autotuning candidate fits on the device")
    g, m, ms = min(table, key=lambda r: r[2])
    return TuneResult(g, m, ms, table)
