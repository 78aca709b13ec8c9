"""GPU: the two replay forms of the transport evaluate the same arithmetic in
the same order, so their results are bitwise equal:

  * geometric replay (task records, face geometry re-read every sweep) and
    compiled replay (per-task upwind entries in level-position planes, with
    overflow entries for tasks with more than 12 upwind faces), with one and
    several timestep levels (restricted sweeps replay the full schedule
    through active-set lists);
  * the compiled replay at every block size.

Results are also checked against the oracle in the parity tests; here the bar
is bit-identity of the raw FP32/FP64 transport state and of x, T and levels.
"""

import numpy as np
import pytest

from subsweep_amd import Sweep
from subsweep_amd import _native as N
from subsweep_amd import workloads as W

pytestmark = pytest.mark.gpu


def _run(pb, mode, steps, prebuild=False, block=0, group_levels=0):
    s = Sweep(pb.params, pb.grid, *pb.state(), transport_mode=mode, replay_block=block, group_levels=group_levels)
    info = s.info()
    states = []
    for _ in range(steps):
        if prebuild:  # parity-hook layerings between the steps
            for lvl in range(pb.params.num_timestep_levels):
                s.levelsets(lvl)
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
    ic, sc = _run(pb, N.TRANSPORT_MODE_COMPILED, pb.steps)  # grouped replay (default)
    assert ic.group_levels == 4 and ig.group_levels == 1
    _assert_same(sg, sc, "geometric vs compiled (grouped)")
    ic, sc = _run(pb, N.TRANSPORT_MODE_COMPILED, pb.steps, group_levels=1)
    assert ig.transport_mode == N.TRANSPORT_MODE_GEOMETRIC
    assert ic.transport_mode == N.TRANSPORT_MODE_COMPILED
    assert ic.overflow_entries > 0, "the grid should exercise the overflow entries"
    _assert_same(sg, sc, "geometric vs compiled")


def test_graph_replay_bitwise(hip_available):
    # one handle at a time: graphs are captured only while a single handle lives
    pb = W.jittered_problem(10, 1.0, "6", 84, levels=3, periodic=True, steps=5, precision=32)
    pb = pb.with_params(max_timestep=1e-3 * 3.15576e13)
    runs = []
    for graphs in (True, False):
        s = Sweep(pb.params, pb.grid, *pb.state(), graph_replay=graphs)
        states = []
        for _ in range(pb.steps):
            s.run_sweeps()
            states.append(s.transport_state())
        info = s.info()
        assert info.graph_replay == (1 if graphs else 0)
        assert info.graph_capture_failures == 0
        s.close()
        runs.append(states)
    for step, (a, b) in enumerate(zip(*runs)):
        for u, v in zip(a, b):
            assert np.array_equal(u.view(np.uint8), v.view(np.uint8)), f"step {step + 1}"


def test_replay_blocks_bitwise(hip_available):
    pb = W.jittered_problem(10, 1.0, "6", 84, levels=3, periodic=True, steps=4, precision=32)
    pb = pb.with_params(max_timestep=1e-3 * 3.15576e13)
    # one launch per level set (no groups), where the replay block applies
    i64, s64 = _run(pb, N.TRANSPORT_MODE_COMPILED, pb.steps, block=64, group_levels=1)
    assert i64.replay_block == 64
    for block in (128, 256):
        ib, sb = _run(pb, N.TRANSPORT_MODE_COMPILED, pb.steps, block=block, group_levels=1)
        assert ib.replay_block == block
        _assert_same(s64, sb, f"block 64 vs {block}")


@pytest.mark.parametrize("mode", [N.TRANSPORT_MODE_GEOMETRIC, N.TRANSPORT_MODE_COMPILED])
def test_levelset_hook_leaves_the_replay_alone(mode, hip_available):
    # the parity hook builds a restricted layering in scratch memory between
    # steps; the replay schedule and its active-set lists must be unaffected
    pb = W.jittered_problem(12, 1.0, "8", 16, levels=3, periodic=True, steps=4, precision=32)
    pb = pb.with_params(max_timestep=1e-3 * 3.15576e13)
    _, plain = _run(pb, mode, pb.steps)
    _, hooked = _run(pb, mode, pb.steps, prebuild=True)
    _assert_same(plain, hooked, "with and without level-set hook calls")


def test_removed_modes_are_rejected(hip_available):
    pb = W.cartesian_problem(4, 16, levels=1, periodic=False, source_cells=(0,), steps=1)
    with pytest.raises(N.SweepError):
        Sweep(pb.params, pb.grid, *pb.state(), transport_mode=1)
    s = Sweep(pb.params, pb.grid, *pb.state())
    with pytest.raises(N.SweepError):
        s.set_transport_mode(N.TRANSPORT_MODE_GEOMETRIC)  # after the grid upload
    s.close()


def test_autotune_picks_a_measured_candidate(hip_available):
    from subsweep_amd.autotune import autotune, candidates
    pb = W.jittered_problem(8, 1.0, "4", 16, levels=1, periodic=True, steps=2, precision=32)
    r = autotune(pb, cands=candidates((64, 256), groups=(2,), group_blocks=(64,)), steps=1, warmup=2)
    assert len(r.table) == 7
    assert any(x[6] for x in r.table), "the direction sweeps are a candidate"
    assert (r.transport_mode, r.replay_block, r.graph_replay, r.narrow_runs, r.group_levels, r.group_block,
            r.direction_sweeps, r.ms_per_step) in [tuple(x) for x in r.table]
    assert r.ms_per_step == min(x[7] for x in r.table)


@pytest.mark.parametrize("mode", [N.TRANSPORT_MODE_GEOMETRIC, N.TRANSPORT_MODE_COMPILED])
def test_locality_order_bitwise(mode, hip_available):
    # cells stored in the Morton order of their positions vs the caller's order:
    # every output, in the caller's order, bitwise equal
    pb = W.jittered_problem(10, 1.0, "6", 16, levels=3, periodic=True, steps=4, precision=32)
    pb = pb.with_params(max_timestep=1e-3 * 3.15576e13)
    out = []
    for order in (True, False):
        s = Sweep(pb.params, pb.grid, *pb.state(), transport_mode=mode, locality_order=order)
        assert s.info().locality_order == (1 if order else 0)
        seq = []
        for _ in range(pb.steps):
            s.run_sweeps()
            r = s.results()
            seq.append((*s.transport_state(), *s.fluxes(), r.ionized_hydrogen_fraction, r.temperature,
                        r.photon_rate, r.level, r.change_timescale, s.counts(1), s.levelsets(2)))
        s.close()
        out.append(seq)
    for step, (a, b) in enumerate(zip(*out)):
        for k, (u, v) in enumerate(zip(a, b)):
            assert np.array_equal(u.view(np.uint8), v.view(np.uint8)), f"step {step + 1}, field {k}"


@pytest.mark.parametrize("mode", [N.TRANSPORT_MODE_GEOMETRIC, N.TRANSPORT_MODE_COMPILED])
def test_device_memory_within_estimate(mode, hip_available):
    # the handle's peak over upload, steps and the result download stays
    # within the admission estimate; the layering scratch of the diagnostic
    # hooks is held only while they run
    from subsweep_amd.serve import estimate_device_bytes
    pb = W.jittered_problem(10, 1.0, "6", 84, levels=3, periodic=True, steps=2, precision=32)
    s = Sweep(pb.params, pb.grid, *pb.state(), transport_mode=mode)
    for _ in range(2):
        s.run_sweeps()
    s.results()
    info = s.info()
    held = info.device_bytes
    assert held <= info.peak_device_bytes <= estimate_device_bytes(pb, mode)
    s.levelsets(1)
    s.counts(0)
    assert s.info().device_bytes == held
    s.close()


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("levels", [1, 4])
def test_narrow_runs_bitwise(precision, levels, hip_available):
    # runs of narrow level sets in one workgroup vs one launch per level set;
    # with 4 levels the restricted sweeps replay through active-set lists,
    # whose level sets are mostly narrow
    pb = W.jittered_problem(12, 1.0, "8", 84, levels=levels, periodic=True, steps=5, precision=precision)
    pb = pb.with_params(max_timestep=1e-3 * 3.15576e13)
    out = []
    for narrow in (True, False):
        s = Sweep(pb.params, pb.grid, *pb.state(), transport_mode=N.TRANSPORT_MODE_COMPILED, narrow_runs=narrow,
                  group_levels=1)
        assert s.info().narrow_runs == (1 if narrow else 0)
        seq = []
        for _ in range(pb.steps):
            s.run_sweeps()
            r = s.results(photon_rate=False)
            seq.append((*s.transport_state(), r.ionized_hydrogen_fraction, r.temperature, r.level))
        s.close()
        out.append(seq)
    _assert_same(out[1], out[0], "narrow runs off vs on")


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("levels", [1, 4])
def test_grouped_full_sweeps_bitwise(precision, levels, hip_available):
    # groups of 2 to 8 level sets per launch (overlapping upwind cones of 64 or
    # 128 roots, one workgroup each) vs one launch per level set; with 4
    # timestep levels the restricted sweeps replay filtered cones
    pb = W.jittered_problem(12, 1.0, "8", 84, levels=levels, periodic=True, steps=5, precision=precision)
    pb = pb.with_params(max_timestep=1e-3 * 3.15576e13)
    ref = None
    for m, gb in ((1, 0), (2, 128), (3, 64), (4, 128), (6, 64), (8, 128), (4, 32), (8, 32)):
        s = Sweep(pb.params, pb.grid, *pb.state(), transport_mode=N.TRANSPORT_MODE_COMPILED, group_levels=m,
                  group_block=gb)
        info = s.info()
        assert info.group_levels == m and info.group_fallback == 0
        assert info.group_block == (gb or 64)
        if m > 1:
            assert info.group_tasks >= pb.grid.num_cells * info.num_directions
        seq = []
        for _ in range(pb.steps):
            s.run_sweeps()
            r = s.results(photon_rate=False)
            seq.append((*s.transport_state(), r.ionized_hydrogen_fraction, r.temperature, r.level))
        s.close()
        if ref is None:
            ref = seq
        else:
            _assert_same(ref, seq, f"groups of {m} (block {gb}) vs per level set")
