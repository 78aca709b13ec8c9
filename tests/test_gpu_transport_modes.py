"""GPU: a persistent (one launch per sweep) transport kernel gives bitwise the
same fluxes as the level-synchronous kernels (the pull form makes every task's
result independent of the order in which tasks run), for FP64 and FP32
transport, with one and several timestep levels."""

import numpy as np
import pytest

from subsweep_amd import Sweep
from subsweep_amd import workloads as W

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="the persistent transport kernel did not terminate on its first two MI355X runs; "
                            "it has been removed and ssw_set_transport_mode rejects modes 1 and 2"),
]


def _run(pb, mode, steps):
    s = Sweep(pb.params, pb.grid, *pb.state())
    s.set_transport_mode(mode)
    for _ in range(steps):
        s.run_sweeps()
    r = s.results()
    i, o, p = s.fluxes()
    s.close()
    return r, (i, o, p)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("levels", [1, 3])
def test_persistent_matches_level_synchronous(precision, levels, hip_available):
    pb = W.jittered_problem(16, 1.0, "8", 84, levels=levels, periodic=True, steps=5, precision=precision)
    pb = pb.with_params(max_timestep=1e-2 * 3.15576e13)
    r0, f0 = _run(pb, 0, pb.steps)
    for mode in (1, 2):
        r, f = _run(pb, mode, pb.steps)
        assert np.array_equal(r.level, r0.level)
        for a, b in zip(f, f0):
            assert np.array_equal(a, b), f"mode {mode}: fluxes differ from the level-synchronous kernels"
        assert np.array_equal(r.ionized_hydrogen_fraction, r0.ionized_hydrogen_fraction)
        assert np.array_equal(r.temperature, r0.temperature)
