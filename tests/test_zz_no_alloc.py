"""GPU: ssw_download_noalloc, the download that allocates nothing, gives
bitwise the outputs of ssw_download. Two calls in a row through the C ABI,
the second on a smaller grid (the shared staging buffer keeps the first call's
contents beyond the second's cells), with the cells stored in position order
(outputs permuted into the caller's order on the device) and in the caller's
order, FP64 and FP32 transport."""

import ctypes as C

import numpy as np
import pytest

from subsweep_amd import Sweep
from subsweep_amd import _native as N
from subsweep_amd import workloads as W

pytestmark = pytest.mark.gpu

MYR = 3.15576e13
FIELDS = ("ionized_hydrogen_fraction", "temperature", "timestep", "photon_rate", "level", "change_timescale")


def _problem(n, precision):
    return W.jittered_problem(n, 1.0, "4", 16, levels=2, periodic=True, steps=3, precision=precision,
                              rho_mean=1e-3 * W.RHO_MEAN).with_params(max_timestep=1e-3 * MYR)


@pytest.mark.parametrize("locality_order", [True, False])
@pytest.mark.parametrize("precision", [64, 32])
def test_download_noalloc_matches_download(locality_order, precision, hip_available):
    for n in (14, 9):  # 2744 cells, then 729
        pb = _problem(n, precision)
        s = Sweep(pb.params, pb.grid, *pb.state(), locality_order=locality_order)
        assert bool(s.info().locality_order) == locality_order
        for _ in range(pb.steps):
            s.run_sweeps()
        want = s.results()
        got = s.results(noalloc=True)
        for f in FIELDS:
            a, b = getattr(got, f), getattr(want, f)
            assert a.dtype == b.dtype and a.shape == b.shape, f
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8)), f"{n}^3 cells: {f} differs"
        assert (want.photon_rate > 0).mean() > 0.5, "the photon rates should be non-zero"
        # only some outputs requested: the others stay untouched
        x = np.full(pb.grid.num_cells, -1.0)
        lvl = np.full(pb.grid.num_cells, 7, np.uint8)
        N.check(N.sweep_lib().ssw_download_noalloc(s._h, N.ptr(x, C.c_double), None, None, None,
                                                   N.ptr(lvl, C.c_uint8), None))
        assert np.array_equal(x, want.ionized_hydrogen_fraction)
        assert np.array_equal(lvl, want.level)
        s.close()
