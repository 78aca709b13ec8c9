"""GPU parity on optically thin gas, where photons cross the box.

The headline workloads' gas is optically thick (tau ~ 200-1100 per cell), so
their fluxes are mostly zero a few cells from a source. Here the mean density
is scaled down until a cell's optical depth is of order one: on the 16^3
jittered grid in the 4000 kpc box with mean density 1e-30 g/cm^3 more than
99 % of the (cell, direction) pairs carry photons from the second step on, so
the compared fluxes run through every face kind, the plane entries and the
overflow entries of the compiled schedule, and overlapping cones of the
grouped replay.

Bars (the transport in pull form, src/sweep/mod.rs:423-513, against the
oracle's push form):
  * FP64 transport: every flux, ionised fraction, temperature and photon rate
    within 1e-10 relative in every launch shape. Fluxes are compared down to
    1e-20 of the largest: in this gas every flux only grows from sweep to sweep
    (the ionised fraction rises, the attenuation falls), so the oracle's
    corrections never cancel and both forms are products and sums of
    positive terms with relative rounding errors of a few ulps per
    attenuation factor, far below 1e-10 even 60 factors deep; no value comes
    near the double underflow.
  * the species timescale within 1e-10 for the cells whose update is not
    near cancellation (its relative sensitivity to the summed rate is below
    100, measured on the oracle's own solver), after a one-level step;
  * FP32 transport: ionised fraction and temperature within 1e-5 relative,
    timestep levels equal.
  * the TimescaleCounter of every step equal to the oracle's.
"""

import numpy as np
import pytest

from oracle import oracle as O
from oracle.oracle import OracleSweep
from parity_util import limiting_counts, max_rel, rel_ok, summed_rate
from subsweep_amd import Sweep
from subsweep_amd import _native as N
from subsweep_amd import workloads as W

pytestmark = pytest.mark.gpu

MYR = 3.15576e13
THIN = 1e-3 * W.RHO_MEAN
SHAPES = {
    "grouped": {},
    "per level set": {"group_levels": 1},
    "geometric": {"transport_mode": N.TRANSPORT_MODE_GEOMETRIC},
    "direction sweeps": {"direction_sweeps": True},
}
FLUX_FLOOR = 1e-20
ORACLE_THREADS = 8


@pytest.fixture(scope="module")
def thin16():
    return W.jittered_problem(16, 1.0, "8", 84, levels=4, periodic=True, steps=4, rho_mean=THIN).with_params(
        max_timestep=1e-3 * MYR)


@pytest.fixture(scope="module")
def thin_clustered():
    # 20000 clustered cells (volumes spanning ~1e3), mean density 1e-31 g/cm^3:
    # the timestep levels spread over all four by the fourth step, so the
    # restricted sweeps replay filtered cones
    return W.clustered_problem(20000, 84, levels=4, sources=16, mesh=32, steps=4, precision=64,
                               rho_mean=1e-4 * W.RHO_MEAN).with_params(max_timestep=1e-3 * MYR)


def _oracle_run(pb):
    o = OracleSweep(pb.params, pb.grid, *pb.state(), threads=ORACLE_THREADS)
    states = []
    for _ in range(pb.steps):
        el = o.run_sweeps()
        r = o.results()
        states.append((el, r, o.fluxes(), o.limiting_processes()))
    o.close()
    return states


def _compare_step(g, want, step, what):
    el, ro, (io, oo, po), lim = want
    rg = g.results()
    assert np.array_equal(rg.level, ro.level), f"{what} step {step}: timestep levels differ"
    rel_ok(rg.ionized_hydrogen_fraction, ro.ionized_hydrogen_fraction, 1e-10, name=f"{what} step {step} x_HII")
    rel_ok(rg.temperature, ro.temperature, 1e-10, name=f"{what} step {step} T")
    ig, og, pg = g.fluxes()
    for name, u, v in (("incoming", ig, io), ("outgoing", og, oo), ("periodic", pg, po), ("photon rate",
                                                                                           rg.photon_rate,
                                                                                           ro.photon_rate)):
        rel_ok(u, v, 1e-10, atol_frac=FLUX_FLOOR, name=f"{what} step {step} {name}")
    got = limiting_counts(g.info())
    tot = sum(lim.values())
    for k in lim:
        assert abs(got[k] - lim[k]) <= max(2, tot // 1000), f"{what} step {step}: limiting processes {got} vs {lim}"
    return io


@pytest.mark.parametrize("shape", list(SHAPES))
def test_thin_jittered_fp64_every_shape(thin16, shape, hip_available):
    pb = thin16
    want = _oracle_run(pb)
    g = Sweep(pb.params, pb.grid, *pb.state(), **SHAPES[shape])
    info = g.info()
    DN = pb.grid.num_cells * info.num_directions
    if shape != "geometric":
        assert info.overflow_entries > 0, "the grid should exercise the overflow entries"
    if shape == "grouped":
        assert info.group_levels == 4 and info.group_tasks > DN, "cones of different workgroups should overlap"
        assert info.group_launches > 0 and info.group_workgroups >= info.group_launches
        assert info.group_blocks_per_cu >= 1, "the grouped kernel's occupancy limit should be known"
        assert info.compute_units >= 1
    assert info.direction_sweeps == (1 if shape == "direction sweeps" else 0)
    for step in range(pb.steps):
        assert g.run_sweeps() == want[step][0]
        io = _compare_step(g, want[step], step, shape)
        if step >= 1:
            frac = float((io > FLUX_FLOOR * io.max()).mean())
            assert frac > 0.5, f"only {frac:.3f} of the cell-directions carry photons"
    g.close()


def test_thin_clustered_fp64_levels(thin_clustered, hip_available):
    pb = thin_clustered
    want = _oracle_run(pb)
    assert len(np.unique(want[-1][1].level)) == 4, "the levels should spread over all four"
    for shape in ("grouped", "geometric", "direction sweeps"):
        g = Sweep(pb.params, pb.grid, *pb.state(), **SHAPES[shape])
        for step in range(pb.steps):
            assert g.run_sweeps() == want[step][0]
            io = _compare_step(g, want[step], step, f"clustered {shape}")
        assert (io > 0).mean() > 0.5
        g.close()


@pytest.mark.parametrize("which", ["jittered", "clustered"])
def test_thin_fp32_transport(which, thin16, thin_clustered, hip_available):
    pb = thin16 if which == "jittered" else thin_clustered
    want = _oracle_run(pb)
    g = Sweep(pb.with_params(transport_precision=32).params, pb.grid, *pb.state())
    for step in range(pb.steps):
        assert g.run_sweeps() == want[step][0]
        rg, ro = g.results(), want[step][1]
        rel_ok(rg.ionized_hydrogen_fraction, ro.ionized_hydrogen_fraction, 1e-5, name=f"fp32 step {step} x_HII")
        rel_ok(rg.temperature, ro.temperature, 1e-5, name=f"fp32 step {step} T")
        assert np.array_equal(rg.level, ro.level), f"fp32 step {step}: timestep levels differ"
    g.close()


def test_thin_species_timescale_tight(hip_available):
    pb = W.jittered_problem(16, 1.0, "8", 84, levels=1, periodic=True, steps=3, rho_mean=THIN).with_params(
        max_timestep=1e-3 * MYR)
    o = OracleSweep(pb.params, pb.grid, *pb.state(), threads=ORACLE_THREADS)
    g = Sweep(pb.params, pb.grid, *pb.state())
    for _ in range(pb.steps - 1):
        o.run_sweeps()
        g.run_sweeps()
    before = o.results()
    o.run_sweeps()
    g.run_sweeps()
    ro, rg = o.results(), g.results()
    D = pb.params.directions
    rate = summed_rate(o, pb.source, D)
    rho, vol, size = pb.density, pb.grid.volume, pb.grid.size
    dt, s = pb.params.max_timestep, pb.params.chemistry_timestep_safety_factor
    keep = np.zeros(pb.grid.num_cells, bool)
    delta = 1e-12
    for c in range(pb.grid.num_cells):
        args = (before.ionized_hydrogen_fraction[c], before.temperature[c], rho[c], vol[c], size[c])
        base = O.perform_timestep(*args, rate[c], dt, s, floor=True)["timescale"]
        # the inputs are the oracle's own: its single-cell solver reproduces its timescale bit for bit
        assert base == ro.timestep[c], c
        hi = O.perform_timestep(*args, rate[c] * (1 + delta), dt, s, floor=True)["timescale"]
        lo = O.perform_timestep(*args, rate[c] * (1 - delta), dt, s, floor=True)["timescale"]
        kappa = max(abs(hi - base), abs(lo - base)) / abs(base) / delta
        keep[c] = kappa <= 100.0
    assert keep.mean() > 0.5, f"only {keep.mean():.3f} of the cells are well conditioned"
    rel_ok(rg.timestep[keep], ro.timestep[keep], 1e-10, name="species timestep (well-conditioned cells)")
    assert max_rel(rg.ionized_hydrogen_fraction, ro.ionized_hydrogen_fraction) < 1e-10
    o.close()
    g.close()


def test_fp32_transport_applies_the_rate_threshold(hip_available):
    # significant_rate_threshold zeroes the outgoing rate of a task whose
    # incoming rate is below it (src/chemistry/hydrogen_only/mod.rs:81-82).
    # FP32 transport carries rates divided by a power of two near the largest
    # source (here ~2^160), where the threshold itself is far below the float
    # range; the comparison is made in double, so both precisions zero the
    # same tasks, up to those whose incoming rate lies within FP32 rounding of
    # the threshold. Non-periodic, so the incoming rate is I + S/D exactly.
    thr = 1e42
    pb = W.jittered_problem(16, 1.0, "8", 84, levels=1, periodic=False, steps=2, threshold=thr,
                            rho_mean=THIN).with_params(max_timestep=1e-3 * MYR)
    g64 = Sweep(pb.params, pb.grid, *pb.state())
    g32 = Sweep(pb.with_params(transport_precision=32).params, pb.grid, *pb.state())
    sod = pb.source / 84.0
    for step in range(pb.steps):
        g64.run_sweeps()
        g32.run_sweeps()
        i64, o64, _ = g64.fluxes()
        _, o32, _ = g32.fluxes()
        zero64, zero32 = o64 == 0.0, o32 == 0.0
        assert 0.05 < zero64.mean() < 0.95, f"step {step}: the threshold zeroes {zero64.mean():.3f} of the tasks"
        near = np.abs((i64 + sod) - thr) < 1e-4 * thr
        differ = zero64 != zero32
        assert not (differ & ~near).any(), (f"step {step}: {int((differ & ~near).sum())} tasks away from the "
                                            f"threshold are zeroed in one precision only")
    g64.close()
    g32.close()
