"""GPU: packed steps (ssw_step_batch). Uneven problems - cell counts,
direction counts, timestep levels, transport precision, group block, launch
shapes without groups - stepped together, their grouped launches packed into
shared launches, give bitwise the results of stepping each alone: the fluxes
X and I, the ionised fraction, the temperature and the timestep levels after
every step, and the elapsed times. The packed path launches the same
workgroups with the same inputs as the lone launches, only more of them per
launch, so any difference is an indexing mistake of the packing."""

import numpy as np
import pytest

from subsweep_amd import Sweep, SweepError
from subsweep_amd import _native as N
from subsweep_amd import workloads as W

pytestmark = pytest.mark.gpu

MYR = 3.15576e13
THIN = 1e-3 * W.RHO_MEAN


def _problems():
    """(problem, launch-shape keywords): optically thin gas, so that fluxes
    cross every problem, and at least two timestep levels where given, so that
    restricted sweeps replay filtered cones."""
    out = []
    for n, dirs, levels, prec in ((8, 16, 1, 64), (12, 84, 4, 64), (10, 32, 2, 32), (12, 84, 4, 32), (9, 64, 2, 64)):
        pb = W.jittered_problem(n, 1.0, "4", dirs, levels=levels, periodic=True, steps=4, precision=prec,
                                rho_mean=THIN, seed=1338 + n)
        out.append((pb.with_params(max_timestep=1e-3 * MYR), {}))
    pc = W.clustered_problem(6000, 32, levels=4, sources=8, mesh=32, steps=4, precision=64, rho_mean=1e-4 * W.RHO_MEAN)
    out.append((pc.with_params(max_timestep=1e-3 * MYR), {}))
    base = out[1][0]
    out.append((base, {"group_block": 32}))
    out.append((base, {"group_block": 128}))
    out.append((base, {"group_levels": 1}))
    out.append((base, {"transport_mode": N.TRANSPORT_MODE_GEOMETRIC}))
    return out


def _state(s):
    r = s.results(photon_rate=False)
    return (*s.transport_state(), r.ionized_hydrogen_fraction, r.temperature, r.level)


def _assert_same(a, b, what):
    for name, u, v in zip(("X", "I", "x_HII", "T", "level"), a, b):
        assert u.dtype == v.dtype and np.array_equal(u.view(np.uint8), v.view(np.uint8)), f"{what}: {name} differs"


@pytest.fixture(scope="module")
def problems():
    return _problems()


def _alone(problems, steps):
    out = []
    for pb, kw in problems:
        s = Sweep(pb.params, pb.grid, *pb.state(), **kw)
        seq = []
        for _ in range(steps):
            el = s.run_sweeps()
            seq.append((el, _state(s)))
        s.close()
        out.append(seq)
    return out


def test_packed_steps_bitwise_equal_alone(problems, hip_available):
    steps = 4
    alone = _alone(problems, steps)
    hs = [Sweep(pb.params, pb.grid, *pb.state(), **kw) for pb, kw in problems]
    assert any(h.info().group_levels > 1 for h in hs) and any(h.info().group_levels == 1 for h in hs)
    for step in range(steps):
        els = Sweep.run_sweeps_packed(hs)
        for k, h in enumerate(hs):
            assert els[k] == alone[k][step][0], f"problem {k} step {step}: elapsed time"
            _assert_same(_state(h), alone[k][step][1], f"problem {k} step {step}")
    for h in hs:
        h.close()


def test_packed_more_problems_than_one_launch_holds(hip_available):
    # 11 problems of one precision and group block: two packed launches per
    # plan item (8 + 3), and problems whose plans end at different items
    probs = []
    for k in range(11):
        pb = W.jittered_problem(6 + k % 4, 1.0, "2", (16, 32, 64)[k % 3], levels=1 + k % 3, periodic=True, steps=3,
                                precision=64, rho_mean=THIN, seed=2000 + k)
        probs.append((pb.with_params(max_timestep=1e-3 * MYR), {}))
    alone = _alone(probs, 3)
    hs = [Sweep(pb.params, pb.grid, *pb.state(), **kw) for pb, kw in probs]
    for step in range(3):
        els = Sweep.run_sweeps_packed(hs)
        for k, h in enumerate(hs):
            assert els[k] == alone[k][step][0]
            _assert_same(_state(h), alone[k][step][1], f"problem {k} step {step}")
    # packed and lone steps mix on the same handles
    hs[0].run_sweeps()
    Sweep.run_sweeps_packed(hs[1:])
    for h in hs:
        h.close()


def test_packed_graph_replays_bitwise(hip_available):
    # one timestep level: the plans stay, so the packed sweeps are captured on
    # the second step and replayed from the third; every replay is checked
    # against the lone steps, with downloads (eager work) between replays
    probs = []
    for k, (n, dirs, prec) in enumerate(((8, 84, 32), (10, 32, 32), (7, 16, 64), (9, 64, 64), (8, 84, 32))):
        pb = W.jittered_problem(n, 1.0, "3", dirs, levels=1, periodic=True, steps=5, precision=prec, rho_mean=THIN,
                                seed=4000 + k)
        probs.append((pb.with_params(max_timestep=1e-3 * MYR), {"group_block": 32} if k == 4 else {}))
    alone = _alone(probs, 5)
    hs = [Sweep(pb.params, pb.grid, *pb.state(), **kw) for pb, kw in probs]
    before = Sweep.packed_graph_stats()
    for step in range(5):
        Sweep.run_sweeps_packed(hs)
        for k, h in enumerate(hs):
            _assert_same(_state(h), alone[k][step][1], f"problem {k} step {step} (graphs)")
    after = Sweep.packed_graph_stats()
    assert after["captured"] > before["captured"] and after["replayed"] - before["replayed"] >= 3
    for h in hs:
        h.close()


def test_packed_step_rejects_bad_batches(problems, hip_available):
    pb, kw = problems[0]
    a = Sweep(pb.params, pb.grid, *pb.state())
    with pytest.raises(SweepError):
        Sweep.run_sweeps_packed([a, a])
    assert Sweep.run_sweeps_packed([]) == []
    a.close()
