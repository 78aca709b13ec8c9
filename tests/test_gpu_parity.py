"""GPU parity: the HIP sweep (through the C ABI) against the oracle on the
same grids and inputs.

Bars:
  * dependency counts, level-set (Kahn) depths and timestep levels: bit-exact;
  * FP64 transport: per-cell ionised fraction, temperature, photon rate and
    per-(cell, direction) fluxes within 1e-10 relative (summation order only);
  * FP32 transport + FP64 chemistry: ionised fraction within 1e-5 relative,
    the error of every field reported.
"""

import numpy as np
import pytest

from subsweep_amd import Sweep, SweepError
from subsweep_amd import workloads as W
from oracle.oracle import OracleSweep

pytestmark = pytest.mark.gpu

NONE = np.uint32(0xFFFFFFFF)


def _pair(pb):
    g = Sweep(pb.params, pb.grid, *pb.state())
    o = OracleSweep(pb.params, pb.grid, *pb.state())
    return g, o


def _rel_ok(a, b, rtol, atol_frac=0.0, name="", atol_abs=0.0):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = np.maximum(np.abs(a), np.abs(b))
    atol = atol_frac * (np.abs(b).max() if b.size else 0.0) + atol_abs
    err = np.abs(a - b)
    bad = err > rtol * scale + atol
    if bad.any():
        i = np.argmax(np.where(bad, err / np.maximum(scale, 1e-300), 0))
        raise AssertionError(f"{name}: {bad.sum()} of {a.size} outside rtol {rtol}; worst at {i}: {a.flat[i]!r} vs {b.flat[i]!r}")


def _compare_structure(g, o, levels):
    for lvl in range(levels):
        cg, co = g.counts(lvl), o.counts(lvl)
        assert np.array_equal(cg, co), f"counts differ at sweep level {lvl}"
        dg, do = g.levelsets(lvl), o.levelsets(lvl)
        assert np.array_equal(dg, do), f"level sets differ at sweep level {lvl}"


# The timescales are dt / rel with rel a relative change: |(R - R_prev) / R| for
# the photon rate (src/chemistry/mod.rs:73-77) or |change / value| / safety in
# the solver (src/chemistry/hydrogen_only/mod.rs:453-458). rel is what the
# reference computes; its natural scale is 1 and its rounding error is absolute
# (~1e-16 per summation-order difference of R), so rel = dt / timescale is
# compared with an absolute 1e-9 plus 1e-6 relative. Cells whose summed rate R
# is below 1e-30 of the largest (and below 1e-250 s^-1 in any case) are left
# out unless both codes give exactly R = 0: such rates come from chains of
# attenuation factors exp(-tau) near exp's underflow threshold (tau ~ 745 per
# cell), where the device's exp and the host libm differ in returning a
# denormal or 0, so R and its relative change are rounding noise. The integer
# levels derived from the timescales must agree exactly for every cell.
R_FLOOR = 1e-250
R_REL_FLOOR = 1e-30


def _summed_rate(sweep):
    inc, _, per = sweep.fluxes()
    return (inc + per).sum(axis=0)


def _compare_timescale(a, b, level, max_ts, name, keep, rate):
    dt = max_ts * 0.5 ** level.astype(np.float64)
    try:
        _rel_ok((dt / a)[keep], (dt / b)[keep], 1e-6, name=name + " (as dt / timescale)", atol_abs=1e-9)
    except AssertionError as e:
        bad = np.nonzero(keep)[0]
        raise AssertionError(f"{e}; summed rates of compared cells span {rate[bad].min()!r}..{rate[bad].max()!r}")


def _compare_fields(g, o, rtol, flux_atol_frac=1e-12, max_ts=None):
    rg, ro = g.results(), o.results()
    assert np.array_equal(rg.level, ro.level), "timestep levels differ"
    _rel_ok(rg.ionized_hydrogen_fraction, ro.ionized_hydrogen_fraction, rtol, name="x_HII")
    _rel_ok(rg.temperature, ro.temperature, rtol, name="T")
    max_ts = max_ts if max_ts is not None else g.params.max_timestep
    rate, rate_g = _summed_rate(o), _summed_rate(g)
    floor = max(R_FLOOR, R_REL_FLOOR * rate.max())
    keep = ((rate > floor) & (rate_g > floor)) | ((rate == 0.0) & (rate_g == 0.0))
    _compare_timescale(rg.timestep, ro.timestep, ro.level, max_ts, "species timestep", keep, rate)
    _compare_timescale(rg.change_timescale, ro.change_timescale, ro.level, max_ts, "change timescale", keep, rate)
    _rel_ok(rg.photon_rate, ro.photon_rate, rtol, flux_atol_frac, name="photon rate")
    ig, og, pg = g.fluxes()
    io, oo, po = o.fluxes()
    _rel_ok(ig, io, rtol, flux_atol_frac, name="incoming")
    _rel_ok(og, oo, rtol, flux_atol_frac, name="outgoing")
    _rel_ok(pg, po, rtol, flux_atol_frac, name="periodic")


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("dirs", [16, 84])
def test_cartesian_fp64(periodic, dirs, hip_available):
    pb = W.cartesian_problem(8, dirs, levels=1, periodic=periodic, source_cells=(0, 200, 511), steps=3)
    g, o = _pair(pb)
    _compare_structure(g, o, 1)
    for _ in range(pb.steps):
        eg, eo = g.run_sweeps(), o.run_sweeps()
        assert eg == eo
        _compare_fields(g, o, 1e-10)


@pytest.mark.parametrize("levels", [1, 4])
def test_voronoi_fp64_levels(levels, hip_available):
    pb = W.jittered_problem(12, 1.0, "8", 16, levels=levels, periodic=True, steps=6)
    pb = pb.with_params(max_timestep=1e-3 * 3.15576e13, timestep_safety_factor=0.1)
    g, o = _pair(pb)
    _compare_structure(g, o, levels)
    for s in range(pb.steps):
        eg, eo = g.run_sweeps(), o.run_sweeps()
        assert eg == eo
        _compare_fields(g, o, 1e-10)
        _compare_structure(g, o, levels)


def test_voronoi_84_nonperiodic(hip_available):
    pb = W.jittered_problem(10, 1.5, "4", 84, levels=2, periodic=False, threshold=1e-5, steps=3)
    g, o = _pair(pb)
    _compare_structure(g, o, 2)
    for _ in range(pb.steps):
        assert g.run_sweeps() == o.run_sweeps()
        _compare_fields(g, o, 1e-10)


def test_fp32_transport_error(hip_available):
    pb = W.jittered_problem(12, 1.0, "8", 16, levels=1, periodic=True, steps=4)
    g32 = Sweep(pb.params.__class__(**{**pb.params.__dict__, "transport_precision": 32}), pb.grid, *pb.state())
    o = OracleSweep(pb.params, pb.grid, *pb.state())
    for _ in range(pb.steps):
        g32.run_sweeps()
        o.run_sweeps()
    r32, ro = g32.results(), o.results()
    _rel_ok(r32.ionized_hydrogen_fraction, ro.ionized_hydrogen_fraction, 1e-5, name="x_HII fp32")
    _rel_ok(r32.temperature, ro.temperature, 1e-5, name="T fp32")


def test_cycle_is_reported(hip_available):
    # three cells whose local faces form a cycle for direction (1, 0, 0):
    # normals are mirror pairs but inconsistent with any ordering.
    from subsweep_amd.grid import Grid
    s = np.sqrt(0.5)
    # cell 0 -> 1 -> 2 -> 0 all point along +x for their forward faces
    nrm = {(0, 1): (1.0, 0.0, 0.0), (1, 2): (s, s, 0.0), (2, 0): (s, -s, 0.0)}
    faces = {0: [], 1: [], 2: []}
    for (a, b), n in nrm.items():
        faces[a].append((b, n))
        faces[b].append((a, tuple(-v for v in n)))
    off = [0]
    nbr, kind, area, normal = [], [], [], []
    for c in range(3):
        for b, n in faces[c]:
            nbr.append(b)
            kind.append(0)
            area.append(1.0)
            normal.append(n)
        off.append(len(nbr))
    grid = Grid(np.array(off, np.int64), np.array(nbr, np.int32), np.array(kind, np.uint8), np.array(area),
                np.array(normal), np.ones(3), np.ones(3))
    from subsweep_amd import SweepParameters
    with pytest.raises(SweepError) as ei:
        Sweep(SweepParameters(directions=[[1.0, 0.0, 0.0]]), grid, np.ones(3), np.full(3, 1e-10),
              np.full(3, 1e4), np.zeros(3))
    assert ei.value.code == -4


@pytest.mark.parametrize("periodic", [True, False])
def test_single_source_32_full_size(periodic, hip_available):
    """The 32^3 = 32768-cell, 16-direction, single-source workload at its full
    size and step count (10 steps), FP64 transport against the oracle, with
    periodic faces carrying fluxes and acting as boundary."""
    pb = W.single_source_32(periodic=periodic)
    g, o = _pair(pb)
    _compare_structure(g, o, 1)
    for _ in range(pb.steps):
        assert g.run_sweeps() == o.run_sweeps()
    _compare_fields(g, o, 1e-10)


def _field_errors(rg, ro, og, oo):
    """Max and 99th-percentile relative error of each output field of rg
    against ro (rates where the oracle's exceeds 1e-6 of its largest: deeper
    in shadow both precisions carry rounding noise of the attenuation)."""
    out = {}
    fields = {"x_HII": (rg.ionized_hydrogen_fraction, ro.ionized_hydrogen_fraction),
              "temperature": (rg.temperature, ro.temperature),
              "photon_rate": (rg.photon_rate, ro.photon_rate)}
    for k in og:
        fields[k] = (og[k], oo[k])
    for name, (u, v) in fields.items():
        keep = np.abs(v) > 1e-6 * np.abs(v).max() if name not in ("x_HII", "temperature") else np.ones(v.shape, bool)
        rel = np.abs(u[keep] - v[keep]) / np.maximum(np.maximum(np.abs(u[keep]), np.abs(v[keep])), 1e-300)
        out[name] = {"max": float(rel.max()) if rel.size else 0.0,
                     "p99": float(np.percentile(rel, 99)) if rel.size else 0.0,
                     "compared_cells": int(keep.sum())}
    return out


def test_precision_64_all_steps_against_oracle(hip_available):
    """The 262144-cell, 84-direction, 64-source workload at its full step
    count (20 steps): FP64 transport within 1e-10 of the oracle after every
    step; FP32 transport's ionised fraction and temperature within 1e-5 and
    its timestep levels equal after every step. The per-field error of both
    precisions against the oracle after the last step (ionised fraction,
    temperature, photon rate and the optional per-cell chemistry rates) is
    printed, and written as JSON into $SSW_REPORT_DIR when that is set."""
    import json
    import os
    pb = W.precision_64()
    g64 = Sweep(pb.params, pb.grid, *pb.state())
    g32 = Sweep(pb.with_params(transport_precision=32).params, pb.grid, *pb.state())
    o = OracleSweep(pb.params, pb.grid, *pb.state(), threads=8)
    for step in range(pb.steps):
        e = o.run_sweeps()
        assert g64.run_sweeps() == e
        assert g32.run_sweeps() == e
        r64, r32, ro = g64.results(), g32.results(), o.results()
        assert np.array_equal(r64.level, ro.level), f"step {step}: fp64 levels differ"
        assert np.array_equal(r32.level, ro.level), f"step {step}: fp32 levels differ"
        _rel_ok(r64.ionized_hydrogen_fraction, ro.ionized_hydrogen_fraction, 1e-10, name=f"step {step} x_HII fp64")
        _rel_ok(r64.temperature, ro.temperature, 1e-10, name=f"step {step} T fp64")
        _rel_ok(r32.ionized_hydrogen_fraction, ro.ionized_hydrogen_fraction, 1e-5, name=f"step {step} x_HII fp32")
        _rel_ok(r32.temperature, ro.temperature, 1e-5, name=f"step {step} T fp32")
    _compare_fields(g64, o, 1e-10)
    oo = o.chemistry_outputs()
    assert (ro.photon_rate > 0).sum() > 1000, "photons should have spread from the sources"
    report = {"workload": "262144-cell jittered-lattice grid, 84 directions, 64 sources, threshold 1e-5 s^-1, "
                          "20 steps",
              "fp32_vs_oracle": _field_errors(r32, ro, g32.chemistry_outputs(), oo),
              "fp64_vs_oracle": _field_errors(r64, ro, g64.chemistry_outputs(), oo),
              "cells_with_photons": int((ro.photon_rate > 0).sum())}
    print(json.dumps(report))
    if os.environ.get("SSW_REPORT_DIR"):
        with open(os.path.join(os.environ["SSW_REPORT_DIR"], "precision_64_errors.json"), "w") as f:
            json.dump(report, f, indent=1)
    assert report["fp32_vs_oracle"]["x_HII"]["max"] < 1e-5
    for x in (g64, g32, o):
        x.close()


@pytest.mark.parametrize("precision", [64, 32])
def test_ionisation_front_against_oracle(precision, hip_available):
    """An ionisation front crossing a 6^3 box (1e20 m, 1e-24 kg m^-3, 16
    directions, 2 timestep levels, 6 steps): FP64 transport within 1e-10 of
    the oracle in every field; FP32 transport's ionised fraction and
    temperature within 1e-5 relative, timestep levels equal."""
    pb = W.cartesian_problem(6, 16, levels=2, periodic=False, source_cells=(0,), steps=6, max_timestep=3e12,
                             density=1e-24, box=1e20, temperature=1e4)
    o = OracleSweep(pb.params, pb.grid, *pb.state())
    g = Sweep(pb.with_params(transport_precision=precision).params, pb.grid, *pb.state())
    for _ in range(pb.steps):
        assert g.run_sweeps() == o.run_sweeps()
    ro = o.results()
    assert (ro.ionized_hydrogen_fraction > 0.5).sum() > 10, "the front should have crossed several cells"
    if precision == 64:
        _compare_fields(g, o, 1e-10)
    else:
        rg = g.results()
        _rel_ok(rg.ionized_hydrogen_fraction, ro.ionized_hydrogen_fraction, 1e-5, name="x_HII fp32")
        _rel_ok(rg.temperature, ro.temperature, 1e-5, name="T fp32")
        assert np.array_equal(rg.level, ro.level)
    g.close()
    o.close()


def test_chemistry_outputs_against_oracle(hip_available):
    """The optional chemistry outputs (src/sweep/chemistry_output.rs:25-67)
    after the ionisation front of the test above (FP64 transport): every
    cell's photoionisation, heating, recombination and collisional
    ionisation rate within 1e-10 of the oracle's (the heating, a difference
    of photoheating and cooling, with 1e-10 of its largest magnitude as the
    absolute floor)."""
    pb = W.cartesian_problem(6, 16, levels=2, periodic=False, source_cells=(0,), steps=6, max_timestep=3e12,
                             density=1e-24, box=1e20, temperature=1e4)
    g, o = _pair(pb)
    for _ in range(pb.steps):
        assert g.run_sweeps() == o.run_sweeps()
    og, oo = g.chemistry_outputs(), o.chemistry_outputs()
    assert (oo["photoionization_rate"] > 0).sum() > 10, "the front should photoionise several cells"
    for k in ("photoionization_rate", "recombination_rate", "collisional_ionization_rate"):
        _rel_ok(og[k], oo[k], 1e-10, name=k)
    _rel_ok(og["heating_rate"], oo["heating_rate"], 1e-10, atol_frac=1e-10, name="heating_rate")
    g.close()
    o.close()


def test_chemistry_outputs_fp32_transport(hip_available):
    """The optional outputs with FP32 transport (incoming rates stored in a
    power-of-two scale and converted back): within 1e-5 of the oracle's FP64
    values, like the ionised fraction, after the same ionisation front."""
    pb = W.cartesian_problem(6, 16, levels=2, periodic=False, source_cells=(0,), steps=6, max_timestep=3e12,
                             density=1e-24, box=1e20, temperature=1e4)
    o = OracleSweep(pb.params, pb.grid, *pb.state())
    g = Sweep(pb.with_params(transport_precision=32).params, pb.grid, *pb.state())
    for _ in range(pb.steps):
        assert g.run_sweeps() == o.run_sweeps()
    og, oo = g.chemistry_outputs(), o.chemistry_outputs()
    assert (oo["photoionization_rate"] > 0).sum() > 10
    for k in ("photoionization_rate", "recombination_rate", "collisional_ionization_rate"):
        _rel_ok(og[k], oo[k], 1e-5, atol_frac=1e-7, name=k + " fp32")
    _rel_ok(og["heating_rate"], oo["heating_rate"], 1e-5, atol_frac=1e-5, name="heating_rate fp32")
    g.close()
    o.close()
