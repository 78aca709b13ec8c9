"""GPU: the request loop (ssw_server_*).

  * each result is the same alone and in the mix: problems of different grids,
    direction counts, timestep levels and precisions are solved one at a time
    on a one-slot server and all at once on a four-slot server, and every
    output field is bitwise equal (the kernels' arithmetic does not depend on
    scheduling, the replay form or what else runs);
  * a one-slot solve equals a direct `Sweep` life cycle of the same problem;
  * admission: requests whose estimates do not fit the budget together never
    overlap, a request above the whole budget fails with SSW_ERR_BUDGET, and
    the estimate bounds the device memory a handle holds.
"""

import numpy as np
import pytest

from subsweep_amd import Sweep
from subsweep_amd import _native as N
from subsweep_amd import workloads as W
from subsweep_amd.serve import Server, estimate_device_bytes

pytestmark = pytest.mark.gpu

STEPS = 3


def _problems():
    ms = 1e-3 * 3.15576e13
    return [
        W.jittered_problem(8, 1.0, "4", 16, levels=1, periodic=True).with_params(max_timestep=ms),
        W.jittered_problem(10, 1.0, "8", 32, levels=2, periodic=False, precision=32).with_params(max_timestep=ms),
        W.clustered_problem(3000, num_directions=16, levels=4, sources=4, mesh=16, precision=32).with_params(
            max_timestep=ms),
        W.cartesian_problem(8, 84, levels=2, periodic=True, source_cells=(0, 100)),
        W.clustered_problem(2000, num_directions=64, levels=1, sources=2, mesh=16, precision=64).with_params(
            max_timestep=ms),
        W.jittered_problem(9, 1.5, "2", 84, levels=4, periodic=True, precision=32).with_params(max_timestep=ms),
    ]


FIELDS = ("ionized_hydrogen_fraction", "temperature", "timestep", "photon_rate", "level", "change_timescale")


def _same(a, b, what):
    for f in FIELDS:
        u, v = getattr(a, f), getattr(b, f)
        assert np.array_equal(u.view(np.uint8), v.view(np.uint8)), f"{what}: {f} differs"


def test_alone_equals_mix(hip_available):
    probs = _problems()
    with Server(slots=1) as sv:
        alone = [sv.wait(sv.submit(pb, STEPS)) for pb in probs]
    with Server(slots=4) as sv:
        tickets = [sv.submit(pb, STEPS) for pb in probs]
        mix = [sv.wait(t) for t in tickets]
        stats = sv.stats()
    assert stats["peak_running"] > 1, "the four-slot server never ran two requests at once"
    for i, (a, m) in enumerate(zip(alone, mix)):
        assert a.status == 0 and m.status == 0, (a.error, m.error)
        assert a.updates == m.updates > 0
        _same(a, m, f"request {i} alone vs in the mix")
        assert m.device_bytes <= m.estimated_bytes, "estimate below the memory the handle held"


def test_server_equals_direct_sweep(hip_available):
    pb = _problems()[1]
    with Server(slots=2) as sv:
        r = sv.wait(sv.submit(pb, STEPS, transport_mode=N.TRANSPORT_MODE_GEOMETRIC))
    assert r.status == 0, r.error
    s = Sweep(pb.params, pb.grid, *pb.state(), transport_mode=N.TRANSPORT_MODE_COMPILED)
    for _ in range(STEPS):
        s.run_sweeps()
    d = s.results()
    s.close()
    for f in FIELDS:
        assert np.array_equal(getattr(r, f), getattr(d, f)), f


def test_budget_admission(hip_available):
    pb = _problems()[0]
    est = estimate_device_bytes(pb, N.TRANSPORT_MODE_GEOMETRIC)
    # room for one request at a time
    with Server(slots=4, hbm_budget_bytes=est + est // 2) as sv:
        ts = [sv.submit(pb, STEPS, transport_mode=N.TRANSPORT_MODE_GEOMETRIC) for _ in range(3)]
        rs = [sv.wait(t) for t in ts]
        assert sv.stats()["peak_running"] == 1
    assert all(r.status == 0 for r in rs)
    spans = sorted((r.start_s, r.end_s) for r in rs)
    for (s0, e0), (s1, e1) in zip(spans, spans[1:]):
        assert s1 >= e0, "requests overlapped although the budget holds only one"
    with Server(slots=2, hbm_budget_bytes=est // 2) as sv:
        r = sv.wait(sv.submit(pb, STEPS, transport_mode=N.TRANSPORT_MODE_GEOMETRIC))
    assert r.status == N.SSW_ERR_BUDGET and "budget" in r.error
