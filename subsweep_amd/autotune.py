"""Launch-shape autotuner for one problem.

The transport's speed on a given grid depends on how its level sets are
replayed:

* transport mode - compiled replay (one lane per task, upwind entries stored
  per task) or geometric replay (16 lanes per task over the face lists);
* replay block - the thread-block size of the compiled replay (64, 128 or
  256 tasks per block);
* graph replay - level-set launches captured as one HIP graph, or launched
  one at a time;
* narrow runs - runs of narrow level sets solved by one workgroup per run,
  or one launch per level set (compiled replay);
* group levels and group block - launches of 2-8 consecutive level sets
  (overlapping upwind cones of 32, 64 or 128 root tasks, one workgroup each), or
  one launch per level set (compiled replay);
* direction sweeps - every sweep as one launch of one workgroup per direction
  (compiled replay).

`autotune` times a few steps of each candidate on the problem itself and
returns the fastest. Every candidate computes bitwise the same results, so
the choice affects speed only.
"""

from __future__ import annotations

import time
from dataclasses import dataclass

from . import _native as N
from .sweep import Sweep
from .workloads import Problem


@dataclass
class TuneResult:
    transport_mode: int
    replay_block: int
    graph_replay: bool
    narrow_runs: bool
    group_levels: int
    group_block: int
    direction_sweeps: bool
    ms_per_step: float
    # (mode, block, graphs, narrow runs, group levels, group block, direction sweeps, ms per step) per candidate
    table: list

    def kwargs(self) -> dict:
        return {"transport_mode": self.transport_mode, "replay_block": self.replay_block,
                "graph_replay": self.graph_replay, "narrow_runs": self.narrow_runs,
                "group_levels": self.group_levels, "group_block": self.group_block,
                "direction_sweeps": self.direction_sweeps}


def candidates(blocks=(64, 128), groups=(4,), group_blocks=(32, 64, 128), direction_sweeps=True):
    C = N.TRANSPORT_MODE_COMPILED
    out = [(C, b, True, True, 1, 0, False) for b in blocks]
    out.append((C, blocks[0], False, True, 1, 0, False))
    out.append((C, blocks[0], True, False, 1, 0, False))
    out += [(C, blocks[0], True, True, m, gb, False) for m in groups for gb in group_blocks]
    if direction_sweeps:
        out.append((C, blocks[0], True, True, 1, 0, True))
    out.append((N.TRANSPORT_MODE_GEOMETRIC, 0, True, False, 1, 0, False))  # one launch shape (16 lanes per task)
    return out


def autotune(problem: Problem, cands=None, steps: int = 2, warmup: int = 2) -> TuneResult:
    """Times `steps` steps after `warmup` steps of every (mode, block, graphs,
    narrow runs, group levels, group block, direction sweeps) candidate on a
    fresh handle each; candidates that do not fit in device memory are
    skipped."""
    table = []
    for m, b, gr, nr, gl, gb, ds in (cands or candidates()):
        try:
            s = Sweep(problem.params, problem.grid, *problem.state(), transport_mode=m, replay_block=b,
                      graph_replay=gr, narrow_runs=nr, group_levels=gl, group_block=gb, direction_sweeps=ds)
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
            info = s.info()
            table.append((m, b, gr, nr, gl, int(info.group_block), bool(info.direction_sweeps),
                          (time.perf_counter() - t0) / steps * 1e3))
        finally:
            s.close()
    if not table:
        raise RuntimeError("no autotuning candidate fits on the device")
    m, b, gr, nr, gl, gb, ds, ms = min(table, key=lambda r: r[7])
    return TuneResult(m, b, gr, nr, gl, gb, ds, ms, table)
