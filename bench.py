"""Sweep benchmark: cell-direction updates per second on one MI355X.

Workload (one "step" = one Sweep::run_sweeps call, src/sweep/mod.rs:258-272):
262144 cells, the periodic Voronoi tessellation of a 64^3 jittered lattice in
a 4000 kpc box, lognormal density (sigma_ln 1.5, mean 1e-27 g/cm^3), 64 point
sources with luminosities 10^U(48, 51) photons/s, 84 directions, 1 timestep
level, significant_rate_threshold 1e-5 s^-1, max timestep 1 Myr. Inputs are
synthetic and seeded; grid and state are resident on the device before the
timed region starts.

`value` is the FP32-transport + FP64-chemistry throughput; the FP64-transport
throughput of the same workload is reported beside it. The roofline figure is
the algorithmic byte count of the transport per cell-direction update (the
per-update byte model: 8+8 B incoming/periodic reads, 16 B outgoing
read/write, 16 B per-cell absorbed-rate accumulation, and per downwind face
36/28/12 B for local/periodic/boundary faces in FP64, 24/16/8 B and 4+4+8 B
in FP32) times the updates of one launch of the transport kernel, divided by
its average duration from HIP events recorded on the library's stream around
every launch, in extra steps run right after the timed steps (the timed steps
themselves carry no per-launch events). cpu_baseline times the oracle (the single-threaded CPU
restatement of the reference's sweep) on one step of the same workload.
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bytes_per_update(info, precision: int) -> float:
    if precision == 64:
        base, bl, bp, bb = 8 + 8 + 16 + 16, 36.0, 28.0, 12.0
    else:
        base, bl, bp, bb = 4 + 4 + 8 + 16, 24.0, 16.0, 8.0
    return base + info.downwind_local * bl + info.downwind_periodic * bp + info.downwind_boundary * bb


def run_gpu(pb, precision: int, steps: int, warmup: int, profile_steps: int = 3):
    from subsweep_amd import Sweep
    pp = pb.with_params(transport_precision=precision)
    s = Sweep(pp.params, pp.grid, *pp.state())
    for _ in range(warmup):
        s.run_sweeps()
    s.synchronize()
    # timed region: whole steps, no per-launch events
    updates = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        s.run_sweeps()
        updates += s.info().updates_last_step
    s.synchronize()
    t1 = time.perf_counter()
    # kernel timing: the same steps again with HIP events around every launch
    # of the transport kernel on the library's stream
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
        "launches_per_step": launches / max(1, profile_steps),
        "kernel_ms_per_step": kernel_ms / max(1, profile_steps),
        "bytes_per_update": bpu,
        "achieved_gbs": achieved,
        "info": info,
    }


def run_cpu_baseline(pb, budget_s: float = 30.0):
    from oracle.oracle import OracleSweep
    o = OracleSweep(pb.params, pb.grid, *pb.state())
    t0 = time.perf_counter()
    o.run_sweeps()
    dt = time.perf_counter() - t0
    upd = o.updates_last_step()
    o.close()
    return {"value": upd / dt, "unit": "cell-direction updates/s", "cores": 1, "kind": "port",
            "sample": f"1 step ({upd} cell-direction updates) of the same workload, FP64, single thread"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB mebibytes (counter calibration for rocprofv3 --pmc)")
    ap.add_argument("--precision", type=int, default=0, choices=(0, 32, 64),
                    help="run only this transport precision (0: both)")
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
        print("bench.py measures one GPU: this build keeps the whole domain on a single MI355X; "
              "--gpus above 1 is not supported", file=sys.stderr)
        sys.exit(2)

    from subsweep_amd import workloads as W
    t = time.perf_counter()
    pb = W.precision_64()
    log(f"grid: {pb.grid.num_cells} cells, {pb.grid.num_faces} faces, built in {time.perf_counter() - t:.1f} s")

    if args.precision in (0, 32):
        fp32 = run_gpu(pb, 32, args.steps, args.warmup)
        log(f"fp32 transport: {fp32['updates_per_s'] / 1e9:.3f} G updates/s, {fp32['ms_per_step']:.2f} ms/step, "
            f"kernel {fp32['avg_launch_ms']:.3f} ms/launch, {fp32['achieved_gbs']:.0f} GB/s")
    if args.precision in (0, 64):
        fp64 = run_gpu(pb, 64, args.steps, args.warmup)
        log(f"fp64 transport: {fp64['updates_per_s'] / 1e9:.3f} G updates/s, {fp64['ms_per_step']:.2f} ms/step, "
            f"kernel {fp64['avg_launch_ms']:.3f} ms/launch, {fp64['achieved_gbs']:.0f} GB/s")
    if args.precision != 0:  # profiling runs: no JSON line
        return

    traffic = None
    tpath = os.path.join(ROOT, "profiles", "transport_traffic.json")
    if os.path.exists(tpath):
        try:
            traffic = json.load(open(tpath)).get("fp32_bytes_per_launch")
        except Exception:
            traffic = None

    cpu = None
    if not args.no_cpu_baseline:
        cpu = run_cpu_baseline(pb)
        log(f"cpu baseline (oracle, 1 thread): {cpu['value'] / 1e6:.2f} M updates/s")

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
        "data": "synthetic (seeded jittered-lattice Voronoi grid, lognormal gas, random sources)",
        "config": {
            "workload": "precision comparison: 64^3 = 262144-cell jittered-lattice periodic Voronoi grid, "
                        "84 directions, 1 timestep level, 64 sources, threshold 1e-5 s^-1",
            "cells": int(info.num_cells),
            "faces": int(info.num_faces),
            "directions": int(info.num_directions),
            "timestep_levels": 1,
            "dag_depth": int(info.max_depth),
            "updates_per_step": fp32["updates_per_step"],
            "parallelism": "single GPU, whole domain resident",
        },
        "fp64_transport_value": fp64["updates_per_s"],
        "fp64_ms_per_step": fp64["ms_per_step"],
        "fp32_over_fp64": fp32["updates_per_s"] / fp64["updates_per_s"],
        "roofline": {
            "bound": "hbm",
            "achieved": fp32["achieved_gbs"],
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": fp32["achieved_gbs"] / PEAK_HBM_GBS,
            "traffic": traffic,
            "kernel": "k_sweep_rec (one launch per Kahn level set, all directions, graph replay)",
            "avg_launch_ms": fp32["avg_launch_ms"],
            "launches_per_step": fp32["launches_per_step"],
            "kernel_ms_per_step": fp32["kernel_ms_per_step"],
            "bytes_per_update": fp32["bytes_per_update"],
            "fp64_achieved": fp64["achieved_gbs"],
            "fp64_bytes_per_update": fp64["bytes_per_update"],
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
