"""GPU: rotate_directions (src/sweep/direction/mod.rs:158-205) against the
oracle. Between steps the directions are rotated, by small rotations (the
last direction keeps its rates, all others are zeroed by the reference's
remap, so the rotated state carries non-zero rates whose accumulation the
next sweeps continue) and by random ones; after every rotation the dependency
counts and level sets of the rotated directions are bit-exact and, after
every step, the fields match the oracle: FP64 transport within 1e-10 (fluxes
with an absolute floor of 1e-12 of the largest, since the rotated rates are
continued as offset + upwind sum, which cancels where the remapped rate and
the rotated upwind sum nearly agree), FP32 transport's ionised fraction and
temperature within 1e-5. Every launch shape."""

import numpy as np
import pytest

from oracle.oracle import OracleSweep
from parity_util import rel_ok
from subsweep_amd import Sweep, SweepPlugin, random_rotation_matrix, rotation_matrix
from subsweep_amd import _native as N
from subsweep_amd import workloads as W

pytestmark = pytest.mark.gpu

MYR = 3.15576e13
SHAPES = {
    "grouped": {},
    "per level set": {"group_levels": 1},
    "geometric": {"transport_mode": N.TRANSPORT_MODE_GEOMETRIC},
    "direction sweeps": {"direction_sweeps": True},
}


def _problem(levels=2, dirs=16):
    return W.jittered_problem(12, 1.0, "6", dirs, levels=levels, periodic=True, steps=4,
                              rho_mean=1e-3 * W.RHO_MEAN).with_params(max_timestep=1e-3 * MYR)


def _matrices(n, seed=7):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        if k % 2 == 0:  # small: the last direction stays the best aligned with itself
            axis = rng.normal(size=3)
            out.append(rotation_matrix(axis / np.linalg.norm(axis), 0.05))
        else:
            out.append(random_rotation_matrix(rng))
    return out


def _compare(g, o, what, rtol=1e-10):
    rg, ro = g.results(), o.results()
    assert np.array_equal(rg.level, ro.level), f"{what}: timestep levels differ"
    rel_ok(rg.ionized_hydrogen_fraction, ro.ionized_hydrogen_fraction, rtol, name=f"{what} x_HII")
    rel_ok(rg.temperature, ro.temperature, rtol, name=f"{what} T")
    if rtol <= 1e-10:
        rel_ok(rg.photon_rate, ro.photon_rate, rtol, atol_frac=1e-12, name=f"{what} photon rate")
        for name, u, v in zip(("incoming", "outgoing", "periodic"), g.fluxes(), o.fluxes()):
            rel_ok(u, v, rtol, atol_frac=1e-12, name=f"{what} {name}")


@pytest.mark.parametrize("shape", list(SHAPES))
def test_rotation_fp64_against_oracle(shape, hip_available):
    pb = _problem()
    g = Sweep(pb.params, pb.grid, *pb.state(), **SHAPES[shape])
    o = OracleSweep(pb.params, pb.grid, *pb.state(), threads=8)
    kept_rates = 0
    for step, m in enumerate(_matrices(pb.steps)):
        assert g.run_sweeps() == o.run_sweeps()
        _compare(g, o, f"{shape} step {step}")
        g.rotate_directions(m)
        o.rotate_directions(m)
        assert np.array_equal(g.counts(0), o.counts(0)), f"step {step}: counts of the rotated directions"
        assert np.array_equal(g.counts(1), o.counts(1)), f"step {step}: counts at sweep level 1"
        assert np.array_equal(g.levelsets(0), o.levelsets(0)), f"step {step}: level sets of the rotated directions"
        _compare(g, o, f"{shape} after rotation {step}")
        kept_rates += int((o.fluxes()[1] != 0).any())
    assert kept_rates >= 2, "the small rotations should keep the last direction's rates"
    for _ in range(2):
        assert g.run_sweeps() == o.run_sweeps()
        _compare(g, o, f"{shape} after the rotations")
    g.close()
    o.close()


def test_rotation_fp32_transport(hip_available):
    pb = _problem()
    g = Sweep(pb.with_params(transport_precision=32).params, pb.grid, *pb.state())
    o = OracleSweep(pb.params, pb.grid, *pb.state(), threads=8)
    for step, m in enumerate(_matrices(pb.steps, seed=11)):
        assert g.run_sweeps() == o.run_sweeps()
        _compare(g, o, f"fp32 step {step}", rtol=1e-5)
        g.rotate_directions(m)
        o.rotate_directions(m)
    assert g.run_sweeps() == o.run_sweeps()
    _compare(g, o, "fp32 after the rotations", rtol=1e-5)
    g.close()
    o.close()


def test_plugin_rotates_after_every_call(hip_available):
    # rotate_directions_system runs after run_sweep_system every frame, the
    # first (skipped) one included (src/sweep/mod.rs:143-148)
    pb = _problem(levels=1, dirs=16)
    g = Sweep(pb.params, pb.grid, *pb.state())
    plugin = SweepPlugin(g, rotate_directions=True, rotation_rng=np.random.default_rng(5))
    o = OracleSweep(pb.params, pb.grid, *pb.state(), threads=8)
    rng = np.random.default_rng(5)
    assert plugin.run_sweep_system() is None
    o.rotate_directions(random_rotation_matrix(rng))
    for step in range(3):
        res = plugin.run_sweep_system()
        o.run_sweeps()
        ro = o.results()
        rel_ok(res.ionized_hydrogen_fraction, ro.ionized_hydrogen_fraction, 1e-10, name=f"plugin step {step} x")
        o.rotate_directions(random_rotation_matrix(rng))
    assert np.array_equal(g.counts(0), o.counts(0))
    with pytest.raises(N.SweepError):
        g.save_checkpoint()
    g.close()
    o.close()
