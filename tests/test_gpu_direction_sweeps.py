"""GPU: direction sweeps (ssw_set_direction_sweeps). One workgroup per
direction solves its direction's level sets in order, so every task is solved
from the same inputs with the same arithmetic as in the grouped replay, only
in another launch shape: the transport state (X, I), ionised fraction,
temperature and timestep levels are bitwise those of the default replay after
every step, in FP32 and FP64 transport, with restricted sweeps (four timestep
levels spread over the clustered grid's cells: full sweeps by direction,
restricted sweeps grouped) in between, and inside packed steps. The oracle comparison of this shape is in
test_gpu_thin.py."""

import numpy as np
import pytest

from subsweep_amd import Sweep
from subsweep_amd import workloads as W

pytestmark = pytest.mark.gpu

MYR = 3.15576e13
THIN = 1e-3 * W.RHO_MEAN


def _state(s):
    r = s.results(photon_rate=False)
    return (*s.transport_state(), r.ionized_hydrogen_fraction, r.temperature, r.level)


def _same(a, b, what):
    for name, u, v in zip(("X", "I", "x_HII", "T", "level"), a, b):
        assert u.dtype == v.dtype and np.array_equal(u.view(np.uint8), v.view(np.uint8)), f"{what}: {name} differs"


@pytest.mark.parametrize("precision", [32, 64])
@pytest.mark.parametrize("which", ["jittered", "clustered"])
def test_direction_sweeps_bitwise_equal_default(precision, which, hip_available):
    if which == "jittered":
        pb = W.jittered_problem(14, 1.0, "8", 84, levels=4, periodic=True, steps=5, precision=precision,
                                rho_mean=THIN).with_params(max_timestep=1e-3 * MYR)
    else:
        # the levels of this workload spread over all four by the fourth step
        pb = W.clustered_problem(20000, 84, levels=4, sources=16, mesh=32, steps=5, precision=precision,
                                 rho_mean=1e-4 * W.RHO_MEAN).with_params(max_timestep=1e-3 * MYR)
    a = Sweep(pb.params, pb.grid, *pb.state())
    b = Sweep(pb.params, pb.grid, *pb.state(), direction_sweeps=True)
    assert a.info().direction_sweeps == 0 and b.info().direction_sweeps == 1
    levels_seen = set()
    for step in range(pb.steps):
        assert a.run_sweeps() == b.run_sweeps()
        sa = _state(a)
        _same(sa, _state(b), f"{which} fp{precision} step {step}")
        levels_seen |= set(np.unique(sa[4]).tolist())
    if which == "clustered":
        assert len(levels_seen) > 1, "restricted sweeps should take part"
    a.close()
    b.close()


def test_direction_sweeps_in_packed_steps(hip_available):
    probs = [W.jittered_problem(n, 1.0, "4", dirs, levels=lv, periodic=True, steps=3, precision=prec, rho_mean=THIN,
                                seed=5000 + n).with_params(max_timestep=1e-3 * MYR)
             for n, dirs, lv, prec in ((8, 84, 1, 32), (10, 32, 2, 64), (9, 16, 3, 32))]
    alone = []
    for pb in probs:
        s = Sweep(pb.params, pb.grid, *pb.state(), direction_sweeps=True)
        seq = []
        for _ in range(3):
            seq.append((s.run_sweeps(), _state(s)))
        s.close()
        alone.append(seq)
    hs = [Sweep(pb.params, pb.grid, *pb.state(), direction_sweeps=(k != 1)) for k, pb in enumerate(probs)]
    for step in range(3):
        els = Sweep.run_sweeps_packed(hs)
        for k, h in enumerate(hs):
            assert els[k] == alone[k][step][0]
            _same(_state(h), alone[k][step][1], f"packed problem {k} step {step}")
    for h in hs:
        h.close()
