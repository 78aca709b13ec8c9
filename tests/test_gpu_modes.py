"""GPU: the two replay forms of the transport and the build-while-solving
sweep evaluate the same arithmetic in the same order, so their results are
bitwise equal:

  * geometric replay (task records, face geometry re-read every sweep) and
    compiled replay (per-task upwind entries in level-position planes, with
    overflow entries for tasks with more than 12 upwind faces);
  * the first sweep of an active set (which builds the level sets while
    solving) and a replay of the same sweep (schedule built beforehand
    without transport, through the level-set parity hook).

Results are also checked against the oracle in the parity tests; here the bar
is bit-identity of the raw FP32/FP64 transport state and of x, T and levels.
"""

import numpy as np
import pytest

from subsweep_amd import Sweep
from subsweep_amd import _native as N
from subsweep_amd import workloads as W

pytestmark = pytest.mark.gpu


def _run(pb, mode, steps, prebuild=False, groups=0):
    s = Sweep(pb.params, pb.grid, *pb.state(), transport_mode=mode, direction_groups=groups)
    info = s.info()
    if prebuild:
        for lvl in range(pb.params.num_timestep_levels):
            s.levelsets(lvl)
    states = []
    for _ in range(steps):
        s.run_sweeps()
        x, i = s.transport_state()
        r = s.results(photon_rate=False)
        states.append((x, i, r.ionized_hydrogen_fraction, r.temperature, r.level))
    s.close()
    return info, states


def _assert_same(a, b, what):
    for step, (sa, sb) in enumerate(zip(a, b)):
        for name, u, v in zip(("X", "I", "x_HII", "T", "level"), sa, sb):
            assert u.dtype == v.dtype and np.array_equal(u.view(np.uint8), v.view(np.uint8)), \
                f"{what}: {name} differs bitwise after step {step + 1}"


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("levels", [1, 4])
def test_geometric_and_compiled_bitwise(precision, levels, hip_available):
    pb = W.jittered_problem(12, 1.0, "8", 84, levels=levels, periodic=True, steps=5, precision=precision)
    pb = pb.with_params(max_timestep=1e-3 * 3.15576e13)
    ig, sg = _run(pb, N.TRANSPORT_MODE_GEOMETRIC, pb.steps)
    ic, sc = _run(pb, N.TRANSPORT_MODE_COMPILED, pb.steps)
    assert ig.transport_mode == N.TRANSPORT_MODE_GEOMETRIC
    assert ic.transport_mode == N.TRANSPORT_MODE_COMPILED
    assert ic.overflow_entries > 0, "the grid should exercise the overflow entries"
    _assert_same(sg, sc, "geometric vs compiled")


@pytest.mark.parametrize("mode", [N.TRANSPORT_MODE_GEOMETRIC, N.TRANSPORT_MODE_COMPILED])
def test_build_equals_replay(mode, hip_available):
    pb = W.jittered_problem(12, 1.0, "8", 16, levels=3, periodic=True, steps=4, precision=32)
    pb = pb.with_params(max_timestep=1e-3 * 3.15576e13)
    _, built = _run(pb, mode, pb.steps)
    _, replayed = _run(pb, mode, pb.steps, prebuild=True)
    _assert_same(built, replayed, "build-while-solving vs replay")


@pytest.mark.parametrize("mode", [N.TRANSPORT_MODE_GEOMETRIC, N.TRANSPORT_MODE_COMPILED])
def test_direction_groups_bitwise(mode, hip_available):
    # 84 directions in 1, 3 (uneven blocks of 28) and 5 (16/17 directions) groups
    pb = W.jittered_problem(10, 1.0, "6", 84, levels=3, periodic=True, steps=4, precision=32)
    pb = pb.with_params(max_timestep=1e-3 * 3.15576e13)
    i1, s1 = _run(pb, mode, pb.steps, groups=1)
    assert i1.direction_groups == 1
    for groups in (3, 5):
        ig, sg = _run(pb, mode, pb.steps, groups=groups)
        assert ig.direction_groups == groups
        assert ig.max_depth == i1.max_depth
        _assert_same(s1, sg, f"1 vs {groups} direction groups")
    # the level-set parity hook is independent of the grouping too
    for groups in (1, 4):
        s = Sweep(pb.params, pb.grid, *pb.state(), transport_mode=mode, direction_groups=groups)
        d = s.levelsets(0)
        s.close()
        if groups == 1:
            d1 = d
        else:
            assert np.array_equal(d, d1)


def test_removed_modes_are_rejected(hip_available):
    pb = W.cartesian_problem(4, 16, levels=1, periodic=False, source_cells=(0,), steps=1)
    with pytest.raises(N.SweepError):
        Sweep(pb.params, pb.grid, *pb.state(), transport_mode=1)
    s = Sweep(pb.params, pb.grid, *pb.state())
    with pytest.raises(N.SweepError):
        s.set_transport_mode(N.TRANSPORT_MODE_GEOMETRIC)  # after the grid upload
    s.close()


def test_autotune_picks_a_measured_candidate(hip_available):
    from subsweep_amd.autotune import autotune
    pb = W.jittered_problem(8, 1.0, "4", 16, levels=1, periodic=True, steps=2, precision=32)
    r = autotune(pb, groups=(1, 2, 32), modes=(N.TRANSPORT_MODE_GEOMETRIC, N.TRANSPORT_MODE_COMPILED), steps=1,
                 warmup=1)
    assert len(r.table) == 4  # 32 groups exceed the 16 directions and are skipped
    assert (r.direction_groups, r.transport_mode, r.ms_per_step) in [tuple(x) for x in r.table]
    assert r.ms_per_step == min(x[2] for x in r.table)
