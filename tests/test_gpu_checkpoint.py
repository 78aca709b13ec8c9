"""GPU: checkpoints of the dynamic state. A run continued after a checkpoint
and a fresh handle that loads it and continues give bitwise the same
results, also when the second handle stores cells in another order or uses
the other transport mode; a checkpoint of another problem is refused."""

import numpy as np
import pytest

from subsweep_amd import Sweep
from subsweep_amd import _native as N
from subsweep_amd import workloads as W

pytestmark = pytest.mark.gpu


def _state(s):
    r = s.results()
    return (*s.transport_state(), r.ionized_hydrogen_fraction, r.temperature, r.timestep, r.photon_rate, r.level,
            r.change_timescale)


@pytest.mark.parametrize("precision", [64, 32])
def test_checkpoint_continue_bitwise(precision, hip_available):
    pb = W.jittered_problem(10, 1.0, "6", 32, levels=3, periodic=True, steps=3, precision=precision)
    pb = pb.with_params(max_timestep=1e-3 * 3.15576e13)
    a = Sweep(pb.params, pb.grid, *pb.state())
    for _ in range(3):
        a.run_sweeps()
    ckpt = a.save_checkpoint()
    for _ in range(3):
        a.run_sweeps()
    ref = _state(a)
    a.close()
    for kw in ({}, {"locality_order": False}, {"transport_mode": N.TRANSPORT_MODE_GEOMETRIC},
               {"group_levels": 1}, {"direction_sweeps": True}):
        b = Sweep(pb.params, pb.grid, *pb.state(), **kw)
        b.load_checkpoint(ckpt)
        for _ in range(3):
            b.run_sweeps()
        got = _state(b)
        b.close()
        for k, (u, v) in enumerate(zip(ref, got)):
            assert np.array_equal(u.view(np.uint8), v.view(np.uint8)), f"{kw}: field {k}"


def test_checkpoint_of_another_problem_is_refused(hip_available):
    pa = W.jittered_problem(8, 1.0, "4", 16, levels=2, periodic=True, steps=1)
    pb = W.jittered_problem(9, 1.0, "4", 16, levels=2, periodic=True, steps=1)
    a = Sweep(pa.params, pa.grid, *pa.state())
    a.run_sweeps()
    ckpt = a.save_checkpoint()
    a.close()
    b = Sweep(pb.params, pb.grid, *pb.state())
    with pytest.raises(N.SweepError):
        b.load_checkpoint(ckpt)
    with pytest.raises(N.SweepError):
        b.load_checkpoint(ckpt[:10])
    b.close()
