"""Sweep benchmark: cell-direction updates per second on one MI355X.

Headline workload (one "step" = one Sweep::run_sweeps call,
src/sweep/mod.rs:258-272): the 2097152-cell clustered grid, the periodic
Voronoi tessellation of points sampled from a lognormal density field
(Gaussian random field with P(k) ~ k^-2 on a 128^3 mesh, smoothed on 2 mesh
cells, sigma_ln 2, so that cell volumes span about 1e3), 1024 sources on the
densest cells with luminosities from p(L) ~ L^-2 on [1e48, 1e52] photons/s
times an escape fraction of 0.01, 84 directions, 4 timestep levels,
timestep_safety_factor 0.1, FP32 transport with FP64 chemistry, replays
captured as HIP graphs, launch shape chosen by the autotuner. Inputs are
synthetic and seeded; grid and state are resident on the device before the
timed region starts.

`value` is the FP32-transport + FP64-chemistry throughput; the FP64-transport
throughput of the same workload is reported beside it. The roofline figure
is the algorithmic byte count of the transport per cell-direction update
(8+8 B incoming/periodic reads, 16 B outgoing read/write, 16 B per-cell
absorbed-rate accumulation, and per downwind face 36/28/12 B for
local/periodic/boundary faces in FP64, 24/16/8 B and 4+4+8 B in FP32) times
the updates of one launch of the transport kernel, divided by its average
launch duration: HIP events recorded on the library's stream around each
sweep's whole replay in extra steps run right after the timed steps (the
timed steps carry no events), divided by the launched level-set groups.
`traffic` is the HBM bytes per launch from the committed rocprofv3
FETCH_SIZE / WRITE_SIZE summary of the same kernel on the same workload.
cpu_baseline times the oracle (the single-threaded CPU restatement of the
reference's sweep) on one thread pinned to one core, on a bounded sample of
the same workload: the first step's sweep over 16 of the 84 directions.
The 262144-cell precision workload (FP32 vs FP64, with the FP32 error
against the FP64 oracle over its 20 steps from the committed GPU test
report) and the serving loop (p50/p99 latency) are reported in the same
JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
KERNELS = {3: "k_sweep_rec (geometric replay: one launch per Kahn level set, all directions, 16 lanes per task)",
           4: "k_sweep_group (compiled replay grouped: one launch per 4 consecutive Kahn level sets, one workgroup per "
              "block of 64 root tasks solving their upwind cone level set by level set; with group levels 1, "
              "k_sweep_comp: one launch per level set, one lane per task)"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


DIR_KERNEL = ("k_sweep_dir (direction sweeps: one launch per sweep, one 512-lane workgroup per direction solving its "
              "level sets in order with workgroup barriers only, recent upwind values from an LDS ring)")


def kernel_name(info) -> str:
    return DIR_KERNEL if info.direction_sweeps else KERNELS.get(info.transport_mode, "?")


def bytes_per_update(info, precision: int) -> float:
    if precision == 64:
        base, bl, bp, bb = 8 + 8 + 16 + 16, 36.0, 28.0, 12.0
    else:
        base, bl, bp, bb = 4 + 4 + 8 + 16, 24.0, 16.0, 8.0
    return base + info.downwind_local * bl + info.downwind_periodic * bp + info.downwind_boundary * bb


def occupancy(info) -> dict:
    """Waves of the dominant transport kernel per CU: launched (mean
    workgroups per launch x waves per workgroup / CUs) against what one CU
    can hold of this kernel at once (its register and LDS limits). The
    direction sweeps launch one 8-wave workgroup per direction, and their
    128 KiB LDS ring admits one workgroup per CU."""
    cus = int(info.compute_units)
    if info.direction_sweeps:
        wg = int(info.num_directions)
        launched = wg * 8 / cus if cus else None
        return {"kernel": "k_sweep_dir", "compute_units": cus, "workgroups_per_launch": wg,
                "waves_per_workgroup": 8, "launched_waves_per_cu": launched, "limit_waves_per_cu": 8,
                "frac": launched / 8 if launched else None}
    if info.group_launches <= 0:
        return None
    waves_per_wg = 2 * int(info.group_block) // 64
    wg = info.group_workgroups / info.group_launches
    limit = int(info.group_blocks_per_cu) * waves_per_wg
    launched = wg * waves_per_wg / cus if cus else None
    return {"kernel": "k_sweep_group", "compute_units": cus, "workgroups_per_launch": wg,
            "waves_per_workgroup": waves_per_wg, "launched_waves_per_cu": launched,
            "limit_waves_per_cu": limit, "frac": launched / limit if (limit and launched) else None}


def run_gpu(pb, precision: int, steps: int, warmup: int, profile_steps: int = 3, shape: dict = None,
            progress: bool = False):
    """`shape`: Sweep launch-shape keywords (transport_mode, replay_block,
    graph_replay, narrow_runs); defaults when None."""
    from subsweep_amd import Sweep
    pp = pb.with_params(transport_precision=precision)
    s = Sweep(pp.params, pp.grid, *pp.state(), **(shape or {}))
    for k in range(warmup):
        s.run_sweeps()
        if progress:
            s.synchronize()
            log(f"  warmup step {k + 1}/{warmup} done")
    s.synchronize()
    # timed region: whole steps, no per-launch events
    updates = 0
    t0 = time.perf_counter()
    for k in range(steps):
        s.run_sweeps()
        updates += s.info().updates_last_step
        if progress:
            log(f"  step {k + 1}/{steps} at {time.perf_counter() - t0:.1f} s")
    s.synchronize()
    t1 = time.perf_counter()
    # kernel timing: further steps with HIP events around each sweep's whole
    # transport replay on the library's stream
    s.set_timing(True)
    kernel_ms, launches, prof_updates = 0.0, 0, 0
    for _ in range(profile_steps):
        s.run_sweeps()
        tm = s.timing()
        kernel_ms += tm["transport_total_ms"]
        launches += tm["transport_launches"]
        prof_updates += s.info().updates_last_step
    s.set_timing(False)
    info = s.info()
    res = s.results()
    s.close()
    elapsed = t1 - t0
    bpu = bytes_per_update(info, precision)
    avg_launch_ms = kernel_ms / max(1, launches)
    upd_per_launch = prof_updates / max(1, launches)
    achieved = upd_per_launch * bpu / (avg_launch_ms * 1e-3) / 1e9
    return {
        "updates_per_s": updates / elapsed,
        "ms_per_step": elapsed / steps * 1e3,
        "updates_per_step": updates / steps,
        "avg_launch_ms": avg_launch_ms,
        "updates_per_launch": upd_per_launch,
        "launches_per_step": launches / max(1, profile_steps),
        "kernel_ms_per_step": kernel_ms / max(1, profile_steps),
        "bytes_per_update": bpu,
        "achieved_gbs": achieved,
        "kernel": kernel_name(info),
        "direction_sweeps": bool(info.direction_sweeps),
        "replay_block": int(info.replay_block),
        "graph_replay": bool(info.graph_replay),
        "narrow_runs": bool(info.narrow_runs),
        "group_levels": int(info.group_levels),
        "group_block": int(info.group_block),
        "group_tasks_per_full_sweep": int(info.group_tasks),
        "occupancy": occupancy(info),
        "transport_mode": int(info.transport_mode),
        "info": info,
        "results": res,
    }


def field_errors(a, b):
    """Relative error of each output field of `a` against `b` (max and 99th
    percentile over cells). Photon rates are compared where b's rate exceeds
    1e-6 of its maximum (relative errors of rates 1e6 below the peak, deep in
    shadow, are rounding noise in both precisions)."""
    import numpy as np
    out = {}
    for name, u, v in (("x_HII", a.ionized_hydrogen_fraction, b.ionized_hydrogen_fraction),
                       ("temperature", a.temperature, b.temperature),
                       ("photon_rate", a.photon_rate, b.photon_rate)):
        keep = np.ones(v.shape, bool) if name != "photon_rate" else np.abs(v) > 1e-6 * np.abs(v).max()
        rel = np.abs(u[keep] - v[keep]) / np.maximum(np.abs(v[keep]), 1e-300)
        out[name] = {"max": float(rel.max()) if rel.size else 0.0,
                     "p99": float(np.percentile(rel, 99)) if rel.size else 0.0}
    out["timestep_levels_equal"] = float(np.mean(a.level == b.level))
    return out


def run_largest(steps: int = 10, direction_sweeps: bool = False):
    """16777216-cell clustered grid (sigma_ln 2.5, 256^3 field mesh), 8192
    sources, 84 directions, 4 timestep levels, periodic, threshold 1e-5 s^-1,
    FP32 transport with FP64 chemistry, `steps` steps, timed whole."""
    from subsweep_amd import workloads as W
    t = time.perf_counter()
    pb = W.clustered_problem(16777216, levels=4, sources=8192, sigma_ln=2.5, mesh=256, threshold=1e-5,
                             steps=steps, precision=32, name="clustered_16m")
    log(f"largest grid: {pb.grid.num_cells} cells, {pb.grid.num_faces} faces, built in "
        f"{time.perf_counter() - t:.1f} s")
    r = run_gpu(pb, 32, steps, 0, profile_steps=1, progress=True,
                shape={"direction_sweeps": True} if direction_sweeps else None)
    log(f"largest fp32: {r['updates_per_s'] / 1e9:.3f} G updates/s, {r['ms_per_step']:.1f} ms/step, "
        f"mode {r['transport_mode']}, {r['info'].device_bytes / 2**30:.1f} GiB on the device")
    return {
        "workload": "16777216-cell clustered lognormal-field Voronoi grid (sigma_ln 2.5), 8192 sources, "
                    "84 directions, 4 timestep levels, periodic, threshold 1e-5 s^-1, FP32 transport",
        "steps": steps,
        "updates_per_s": r["updates_per_s"],
        "ms_per_step": r["ms_per_step"],
        "updates_per_step": r["updates_per_step"],
        "transport_mode": r["transport_mode"],
        "direction_sweeps": r["direction_sweeps"],
        "device_gib": r["info"].device_bytes / 2**30,
        "peak_device_gib": r["info"].peak_device_bytes / 2**30,
        "dag_depth": int(r["info"].max_depth),
    }


def run_packed(count: int = 16, steps: int = 6, packs=(1, 4)):
    """Packed steps against separate streams: `count` small problems (one
    timestep level, 8000-27000 cells, 16-84 directions, FP32 transport)
    stepped (a) each on its own handle and stream from one thread, (b) with
    one ssw_step_batch call per step (the library cuts them into packs of
    four on four streams), (c) as 4 groups of count/4, each group one
    ssw_step_batch per step from a thread of its own. Graphs of packed sweeps
    are captured on the second step. Whole steps of all problems, timed
    after 2 warmup steps; results are bitwise the same in every arrangement."""
    import threading
    from subsweep_amd import Sweep
    from subsweep_amd import workloads as W
    shapes = [(24, 84), (20, 64), (28, 32), (22, 84), (26, 16), (24, 84), (30, 32), (21, 64)]
    probs = []
    for k in range(count):
        n, dirs = shapes[k % len(shapes)]
        pb = W.jittered_problem(n, 1.0, "8", dirs, levels=1, periodic=True, steps=steps, precision=32, seed=3000 + k)
        probs.append(pb.with_params(max_timestep=1e-3 * 3.15576e13))
    out = {"problems": count, "cells": [int(p.grid.num_cells) for p in probs], "steps": steps}

    def timed(step_all, hs):
        for _ in range(2):
            step_all()
        for h in hs:
            h.synchronize()
        t = time.perf_counter()
        upd = 0
        for _ in range(steps):
            step_all()
            for h in hs:
                h.synchronize()
            upd += sum(h.info().updates_last_step for h in hs)
        dt = time.perf_counter() - t
        return {"ms_per_step": dt / steps * 1e3, "updates_per_s": upd / dt}

    hs = [Sweep(p.params, p.grid, *p.state()) for p in probs]

    def alone():
        for h in hs:
            h.run_sweeps()
    out["separate_streams"] = timed(alone, hs)
    for h in hs:
        h.close()
    for npk in packs:
        hs = [Sweep(p.params, p.grid, *p.state()) for p in probs]
        groups = [hs[i::npk] for i in range(npk)]

        def packed():
            if npk == 1:
                Sweep.run_sweeps_packed(groups[0])
                return
            th = [threading.Thread(target=Sweep.run_sweeps_packed, args=(g,)) for g in groups]
            for t in th:
                t.start()
            for t in th:
                t.join()
        out["packed_one_call" if npk == 1 else f"packed_{npk}_threads"] = timed(packed, hs)
        for h in hs:
            h.close()
    for k, v in out.items():
        if isinstance(v, dict):
            log(f"packed demo {k}: {v['ms_per_step']:.2f} ms per step of all {count}, "
                f"{v['updates_per_s'] / 1e9:.3f} G updates/s")
    return out


def run_shard(gpus: int, steps: int = 4, warmup: int = 2):
    """One GPU's share of a direction-sharded run over `gpus` GPUs: the
    2097152-cell clustered grid (4 timestep levels, 1024 sources) with the
    first ceil(84 / gpus) directions of the 84-direction table as explicit
    directions, FP32 transport. Chemistry runs on the partial rate sum of
    these directions (a sharded run would reduce the partial sums over the
    GPUs once per sweep: N doubles, 16 MB here)."""
    import math
    from subsweep_amd import direction_table
    from subsweep_amd import workloads as W
    k = math.ceil(84 / gpus)
    t = time.perf_counter()
    pb = W.clustered_2m()
    pb = pb.with_params(directions=direction_table(84)[:k].tolist())
    log(f"shard grid: {pb.grid.num_cells} cells, built in {time.perf_counter() - t:.1f} s; {k} of 84 directions")
    r = run_gpu(pb, 32, steps, warmup, profile_steps=1)
    log(f"shard of {gpus}: {r['updates_per_s'] / 1e9:.3f} G updates/s, {r['ms_per_step']:.1f} ms/step, "
        f"transport {r['kernel_ms_per_step']:.1f} ms/step")
    n = pb.grid.num_cells
    return {
        "workload": f"one GPU's share of {gpus}: 2097152-cell clustered lognormal-field Voronoi grid (sigma_ln 2), "
                    f"{k} of 84 directions, 4 timestep levels, 1024 sources, FP32 transport",
        "gpus": gpus,
        "directions": k,
        "steps": steps,
        "warmup": warmup,
        "updates_per_s": r["updates_per_s"],
        "ms_per_step": r["ms_per_step"],
        "transport_ms_per_step": r["kernel_ms_per_step"],
        "dag_depth": int(r["info"].max_depth),
        "sweeps_last_step": int(r["info"].sweeps_last_step),
        "partial_rate_bytes_per_sweep": n * 8,
    }


def tuned(pb, precision: int, warmup: int, enabled: bool, mode: int, block: int):
    """(launch shape, autotune table) for a run: the autotuner's choice, or the
    command line's."""
    if not enabled:
        return {"transport_mode": mode, "replay_block": block}, None
    from subsweep_amd.autotune import autotune
    t = time.perf_counter()
    r = autotune(pb.with_params(transport_precision=precision), warmup=warmup)
    log(f"autotune fp{precision}: mode {r.transport_mode}, block {r.replay_block}, graphs {r.graph_replay}, "
        f"narrow runs {r.narrow_runs}, group levels {r.group_levels}, group block {r.group_block}, "
        f"direction sweeps {r.direction_sweeps}, {r.ms_per_step:.2f} ms/step ({time.perf_counter() - t:.1f} s; " +
        ", ".join(f"m{m}/b{b}/g{int(gr)}/n{int(nr)}/l{gl}x{gb}/d{int(ds)} {ms:.2f}"
                  for m, b, gr, nr, gl, gb, ds, ms in r.table) + ")")
    return r.kwargs(), [list(x) for x in r.table]


def run_precision(warmup: int, steps: int = 20, autotune: bool = True):
    """The 262144-cell precision workload (64^3 jittered lattice, sigma_ln 1.5,
    64 sources, 84 directions, 1 level, threshold 1e-5 s^-1): FP32 transport
    vs FP64 transport (FP64 chemistry in both), `steps` timed steps each, and
    the FP32 error against the FP64 oracle after its 20 steps from the
    committed report of the GPU parity test that computes it."""
    from subsweep_amd import workloads as W
    t = time.perf_counter()
    pb = W.precision_64()
    log(f"precision grid: {pb.grid.num_cells} cells, {pb.grid.num_faces} faces, built in "
        f"{time.perf_counter() - t:.1f} s")
    sh32, t32 = tuned(pb, 32, warmup, autotune, 0, 0)
    r32 = run_gpu(pb, 32, steps, warmup, shape=sh32)
    log(f"precision fp32: {r32['updates_per_s'] / 1e9:.3f} G updates/s, {r32['ms_per_step']:.2f} ms/step, "
        f"kernel {r32['avg_launch_ms']:.4f} ms/launch, {r32['achieved_gbs']:.0f} GB/s")
    sh64, t64 = tuned(pb, 64, warmup, autotune, 0, 0)
    r64 = run_gpu(pb, 64, steps, warmup, shape=sh64)
    log(f"precision fp64: {r64['updates_per_s'] / 1e9:.3f} G updates/s, {r64['ms_per_step']:.2f} ms/step")
    report = None
    rpath = os.path.join(ROOT, "profiles", "precision_64_errors.json")
    if os.path.exists(rpath):
        report = json.load(open(rpath))
    return {
        "workload": "262144-cell jittered-lattice periodic Voronoi grid (sigma_ln 1.5), 84 directions, 1 timestep "
                    "level, 64 sources, threshold 1e-5 s^-1",
        "steps": steps,
        "dag_depth": int(r32["info"].max_depth),
        "fp32_updates_per_s": r32["updates_per_s"],
        "fp64_updates_per_s": r64["updates_per_s"],
        "fp32_over_fp64": r32["updates_per_s"] / r64["updates_per_s"],
        "fp32_ms_per_step": r32["ms_per_step"],
        "fp64_ms_per_step": r64["ms_per_step"],
        "fp32_roofline_frac": r32["achieved_gbs"] / PEAK_HBM_GBS,
        "fp32_avg_launch_ms": r32["avg_launch_ms"],
        "launch_shape_fp32": shape_of(r32, t32),
        "launch_shape_fp64": shape_of(r64, t64),
        "fp32_error_vs_fp64_gpu": field_errors(r32["results"], r64["results"]),
        "error_vs_fp64_oracle_after_20_steps": report,
    }


def info_is_dir(r) -> bool:
    return bool(r["info"].direction_sweeps)


def shape_of(r, table):
    return {"transport_mode": r["transport_mode"], "replay_block": r["replay_block"],
            "graph_replay": r["graph_replay"], "narrow_runs": r["narrow_runs"], "group_levels": r["group_levels"],
            "group_block": r["group_block"], "group_tasks_per_full_sweep": r["group_tasks_per_full_sweep"],
            "group_launches_per_full_sweep": int(r["info"].group_launches),
            "direction_sweeps": r["direction_sweeps"], "autotune_ms_per_step": table}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def run_cpu_baseline(pb, directions: int = 16):
    """The oracle on one thread pinned to one core: the first step of the
    workload (one sweep of every cell, then chemistry) over the first
    `directions` rows of its direction table, a bounded sample of the same
    work (updates/s is per cell-direction update either way)."""
    from oracle.oracle import OracleSweep
    from subsweep_amd import direction_table
    D = pb.params.directions if isinstance(pb.params.directions, int) else len(pb.params.directions)
    sub = pb.with_params(directions=direction_table(D)[:directions].tolist())
    cores = sorted(os.sched_getaffinity(0))
    core = cores[0]
    o = OracleSweep(sub.params, sub.grid, *sub.state(), threads=1)
    old = os.sched_getaffinity(0)
    os.sched_setaffinity(0, {core})  # this thread only (pid 0 = the calling thread)
    try:
        t0 = time.perf_counter()
        o.run_sweeps()
        dt = time.perf_counter() - t0
    finally:
        os.sched_setaffinity(0, old)
    upd = o.updates_last_step()
    o.close()
    return {"value": upd / dt, "unit": "cell-direction updates/s", "cores": 1, "kind": "port",
            "cpu_model": cpu_model(), "pinned_core": core, "seconds": dt,
            "sample": f"first step of the same workload (every cell on the top timestep level, one sweep then "
                      f"chemistry) over {directions} of its {D} directions: {upd} cell-direction updates, FP64, "
                      f"one thread pinned to core {core}"}


def run_serving(count: int, slots: int, budget_gb: float, seed: int = 2024, steps: int = 10, max_batch: int = 0,
                direction_sweeps: bool = False):
    """The serving loop on one server: first all `count` requests submitted at
    once (saturated throughput; grids become resident as they first arrive),
    then the same requests with Poisson arrivals at 0.8 x that completion rate
    (latency at 80 % load; grids resident from the first pass). max_batch >= 2:
    the running requests are stepped together in packed steps."""
    import threading
    import numpy as np
    from subsweep_amd.serve import Server
    from subsweep_amd import workloads as W
    t = time.perf_counter()
    pool = W.serving_pool(8, seed)
    reqs = W.serving_requests(pool, count, seed + 1, steps)
    log(f"serving: {len(pool)} grids ({', '.join(str(g.num_cells) for g, _, _ in pool)} cells), "
        f"{count} requests built in {time.perf_counter() - t:.1f} s")
    budget = int(budget_gb * (1 << 30))

    def summary(res, span):
        bad = [r for r in res if r.status != 0]
        if bad:
            raise RuntimeError(f"{len(bad)} serving requests failed: {bad[0].error}")
        lat = np.array([r.latency_s for r in res])
        return {"p50_latency_ms": float(np.percentile(lat, 50) * 1e3),
                "p99_latency_ms": float(np.percentile(lat, 99) * 1e3),
                "mean_service_ms": float(np.mean([r.end_s - r.start_s for r in res]) * 1e3),
                "updates_per_s": sum(r.updates for r in res) / span,
                "reused_grid_fraction": float(np.mean([r.reused for r in res]))}

    with Server(slots=slots, hbm_budget_bytes=budget, max_batch=max_batch) as sv:
        t0 = time.perf_counter()
        tickets = [sv.submit(r, steps, grid_key=r.pool_index + 1, direction_sweeps=direction_sweeps) for r in reqs]
        sat = [sv.wait(x) for x in tickets]
        span = time.perf_counter() - t0
        saturated = summary(sat, span)
        capacity = count / span
        log(f"serving{' packed' if max_batch else ''} saturated: {capacity:.1f} requests/s, "
            f"{saturated['updates_per_s'] / 1e9:.3f} G updates/s, "
            f"{saturated['reused_grid_fraction']:.2f} of the requests on resident grids")
        rate = 0.8 * capacity  # Poisson arrivals at 80 % of the measured capacity
        gaps = np.random.default_rng(seed + 2).exponential(1.0 / rate, size=count)
        arrivals = np.cumsum(gaps) - gaps[0]
        tickets = [None] * count
        start = time.perf_counter()

        def feeder():
            for i, r in enumerate(reqs):
                delay = arrivals[i] - (time.perf_counter() - start)
                if delay > 0:
                    time.sleep(delay)
                tickets[i] = sv.submit(r, steps, grid_key=r.pool_index + 1, direction_sweeps=direction_sweeps)

        th = threading.Thread(target=feeder)
        th.start()
        res = []
        for i in range(count):
            while tickets[i] is None:
                time.sleep(0.01)
            res.append(sv.wait(tickets[i]))
            if (i + 1) % 64 == 0:
                log(f"serving: {i + 1}/{count} done at {time.perf_counter() - start:.1f} s")
        th.join()
        stats = sv.stats()
    t_first = min(r.submit_s for r in res)
    t_last = max(r.end_s for r in res)
    poisson = summary(res, t_last - t_first)
    return {
        "requests": count,
        "slots": slots,
        "hbm_budget_gb": budget_gb,
        "steps_per_request": steps,
        "grids": [int(g.num_cells) for g, _, _ in pool],
        "arrival_rate_per_s": rate,
        "arrival_rule": "Poisson at 0.8 x the completion rate of the saturated pass",
        "p50_latency_ms": poisson["p50_latency_ms"],
        "p99_latency_ms": poisson["p99_latency_ms"],
        "mean_service_ms": poisson["mean_service_ms"],
        "updates_per_s": poisson["updates_per_s"],
        "reused_grid_fraction": poisson["reused_grid_fraction"],
        "saturated": saturated,
        "max_batch": max_batch,
        "direction_sweeps": direction_sweeps,
        "packed_requests_per_step": stats["packed_requests_per_step"],
        "peak_running": stats["peak_running"],
        "peak_reserved_gb": stats["peak_reserved_bytes"] / (1 << 30),
    }


def serving_passes(args):
    """The unpacked and (with --serving-max-batch >= 2) packed serving passes."""
    ds = args.serving_launch == "direction"
    serving = run_serving(args.serving_requests, args.serving_slots, 200.0, direction_sweeps=ds)
    log(f"serving: p50 {serving['p50_latency_ms']:.1f} ms, p99 {serving['p99_latency_ms']:.1f} ms, "
        f"{serving['updates_per_s'] / 1e9:.3f} G updates/s, peak {serving['peak_running']} running")
    if args.serving_max_batch >= 2:
        packed = run_serving(args.serving_requests, args.serving_packed_slots, 200.0,
                             max_batch=args.serving_max_batch, direction_sweeps=ds)
        log(f"serving packed: p50 {packed['p50_latency_ms']:.1f} ms, p99 {packed['p99_latency_ms']:.1f} ms, "
            f"{packed['updates_per_s'] / 1e9:.3f} G updates/s, {packed['packed_requests_per_step']:.2f} "
            f"requests per packed step")
        serving["packed"] = packed
    return serving


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB mebibytes (counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--precision", type=int, default=0, choices=(0, 32, 64),
                    help="profiling run: only this transport precision of --workload, no JSON line")
    ap.add_argument("--workload", default="clustered", choices=("clustered", "precision"),
                    help="workload of a --precision profiling run")
    ap.add_argument("--serving-requests", type=int, default=256,
                    help="requests of the serving-loop measurement (0: skip it)")
    ap.add_argument("--serving-slots", type=int, default=8)
    ap.add_argument("--serving-packed-slots", type=int, default=16,
                    help="slots of the packed serving pass (requests stepped together, up to --serving-max-batch)")
    ap.add_argument("--serving-max-batch", type=int, default=16,
                    help="requests per packed step (slots / max-batch steppers, at most 4); 0: skip the packed pass")
    ap.add_argument("--serving-launch", default="direction", choices=("direction", "grouped"),
                    help="launch shape of the served requests: direction sweeps or the library default")
    ap.add_argument("--serving-only", action="store_true",
                    help="only run the serving measurement and print its JSON")
    ap.add_argument("--no-precision-workload", action="store_true",
                    help="skip the 262144-cell FP32-vs-FP64 precision workload")
    ap.add_argument("--no-packed-demo", action="store_true", help="skip the packed-steps comparison")
    ap.add_argument("--packed-demo", action="store_true",
                    help="only run the packed-steps comparison (16 small problems) and print its JSON")
    ap.add_argument("--largest", action="store_true",
                    help="only run the 16777216-cell largest-domain measurement and print its JSON")
    ap.add_argument("--mode", type=int, default=0, choices=(0, 3, 4),
                    help="transport replay form: 0 automatic, 3 geometric, 4 compiled")
    ap.add_argument("--block", type=int, default=0, help="compiled replay block size (0: library default)")
    ap.add_argument("--shard", type=int, default=0, metavar="G",
                    help="measure one GPU's share of a direction-sharded run over G GPUs and print its JSON only")
    ap.add_argument("--no-autotune", action="store_true",
                    help="use --mode/--block (and --direction-sweeps) instead of the autotuner's choice")
    ap.add_argument("--direction-sweeps", action="store_true",
                    help="with --no-autotune or --largest: one launch per sweep, one workgroup per direction")
    ap.add_argument("--no-graphs", action="store_true",
                    help="profiling run: launch the replays one by one instead of as captured graphs")
    ap.add_argument("--progress", action="store_true", help="profiling run: log every step")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": n * 4, "copy_bytes_written": n * 4, "launches": 3}))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the library keeps the whole domain on a single MI355X; "
              "--gpus above 1 is not supported", file=sys.stderr)
        sys.exit(2)

    if args.packed_demo:
        print(json.dumps(run_packed()), flush=True)
        return
    if args.serving_only:
        print(json.dumps(serving_passes(args)), flush=True)
        return
    if args.largest:
        print(json.dumps(run_largest(direction_sweeps=args.direction_sweeps)), flush=True)
        return
    if args.shard:
        if not 1 < args.shard <= 84:
            print("--shard G needs 1 < G <= 84", file=sys.stderr)
            sys.exit(2)
        print(json.dumps(run_shard(args.shard)), flush=True)
        return

    from subsweep_amd import workloads as W
    t = time.perf_counter()
    pb = W.precision_64() if (args.precision and args.workload == "precision") else W.clustered_2m()
    log(f"grid: {pb.grid.num_cells} cells, {pb.grid.num_faces} faces, built in {time.perf_counter() - t:.1f} s")
    if args.precision:  # profiling run: one precision, default launch shape unless tuned, no JSON line
        sh, _ = tuned(pb, args.precision, args.warmup, not args.no_autotune, args.mode, args.block)
        if args.no_graphs:
            sh = dict(sh, graph_replay=False)
        if args.no_autotune and args.direction_sweeps:
            sh = dict(sh, direction_sweeps=True)
        r = run_gpu(pb, args.precision, args.steps, args.warmup, shape=sh, progress=args.progress)
        log(f"fp{args.precision} transport: {r['updates_per_s'] / 1e9:.3f} G updates/s, {r['ms_per_step']:.2f} "
            f"ms/step, kernel {r['avg_launch_ms']:.4f} ms/launch, {r['achieved_gbs']:.0f} GB/s")
        return

    cpu = None
    if not args.no_cpu_baseline:
        cpu = run_cpu_baseline(pb)
        log(f"cpu baseline (oracle, 1 thread on core {cpu['pinned_core']}, {cpu['cpu_model']}): "
            f"{cpu['value'] / 1e6:.2f} M updates/s in {cpu['seconds']:.1f} s")

    tune = not args.no_autotune
    sh32, tune32 = tuned(pb, 32, args.warmup, tune, args.mode, args.block)
    fp32 = run_gpu(pb, 32, args.steps, args.warmup, shape=sh32)
    log(f"fp32 transport: {fp32['updates_per_s'] / 1e9:.3f} G updates/s, {fp32['ms_per_step']:.2f} ms/step, "
        f"kernel {fp32['avg_launch_ms']:.4f} ms/launch, {fp32['achieved_gbs']:.0f} GB/s")
    sh64, tune64 = tuned(pb, 64, args.warmup, tune, args.mode, args.block)
    fp64 = run_gpu(pb, 64, args.steps, args.warmup, shape=sh64)
    log(f"fp64 transport: {fp64['updates_per_s'] / 1e9:.3f} G updates/s, {fp64['ms_per_step']:.2f} ms/step, "
        f"kernel {fp64['avg_launch_ms']:.4f} ms/launch, {fp64['achieved_gbs']:.0f} GB/s")

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "transport_traffic.json")
    if os.path.exists(tpath):
        try:
            key = ("k_sweep_dir_" if info_is_dir(fp32) else "") + "clustered_2m_fp32_bytes_per_launch"
            traffic = json.load(open(tpath)).get(key)
        except Exception:
            traffic = None

    precision = None
    if not args.no_precision_workload:
        precision = run_precision(args.warmup, autotune=tune)

    serving = None
    if args.serving_requests > 0:
        serving = serving_passes(args)

    packed_demo = None
    if not args.no_packed_demo:
        packed_demo = run_packed()

    info = fp32["info"]
    line = {
        "metric": "sweep cell-direction updates/s (FP32 transport + FP64 chemistry)",
        "value": fp32["updates_per_s"],
        "unit": "cell-direction updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fp32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 transport, f64 chemistry",
        "data": "synthetic (seeded clustered lognormal-field Voronoi grid, sources on the densest cells)",
        "config": {
            "workload": "clustered batched sweep: 2097152-cell clustered lognormal-field periodic Voronoi grid "
                        "(sigma_ln 2, cell volumes spanning ~1e3), 1024 sources on the densest cells (p(L) ~ L^-2 "
                        "on 1e48-1e52 photons/s, escape fraction 0.01), 84 directions, 4 timestep levels, "
                        "timestep safety factor 0.1, FP32 transport + FP64 chemistry, graph-captured replays, "
                        "autotuned launch shape",
            "cells": int(info.num_cells),
            "faces": int(info.num_faces),
            "directions": int(info.num_directions),
            "timestep_levels": 4,
            "dag_depth": int(info.max_depth),
            "updates_per_step": fp32["updates_per_step"],
            "sweeps_per_step": int(info.sweeps_last_step),
            "parallelism": "single GPU, whole domain resident",
        },
        "fp64_transport_value": fp64["updates_per_s"],
        "fp64_ms_per_step": fp64["ms_per_step"],
        "fp32_over_fp64": fp32["updates_per_s"] / fp64["updates_per_s"],
        "fp32_error_vs_fp64": field_errors(fp32["results"], fp64["results"]),
        "roofline": {
            "bound": "hbm",
            "achieved": fp32["achieved_gbs"],
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": fp32["achieved_gbs"] / PEAK_HBM_GBS,
            "traffic": traffic,
            "kernel": fp32["kernel"],
            "avg_launch_ms": fp32["avg_launch_ms"],
            "launches_per_step": fp32["launches_per_step"],
            "kernel_ms_per_step": fp32["kernel_ms_per_step"],
            "updates_per_launch": fp32["updates_per_launch"],
            "bytes_per_update": fp32["bytes_per_update"],
            "fp64_achieved": fp64["achieved_gbs"],
            "fp64_bytes_per_update": fp64["bytes_per_update"],
            "occupancy": fp32["occupancy"],
        },
        "launch_shape": {"fp32": shape_of(fp32, tune32), "fp64": shape_of(fp64, tune64)},
        "cpu_baseline": cpu,
        "precision_262k": precision,
        "serving": serving,
        "packed_demo": packed_demo,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
