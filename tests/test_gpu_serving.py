"""GPU: the request loop (ssw_server_*).

  * resident grids: requests with a grid key reuse the cached handle of an
    earlier request of the same key and give bitwise its results;
  * the serving mix at its real sizes (grids of 8192-524288 cells, 16-84
    directions, 1-4 timestep levels, 1-64 sources, FP32 transport): 32
    requests alone and in the mix are bitwise equal, and three small requests
    served in the mix (FP32 and FP64 transport) match the oracle (FP64 within
    1e-10, FP32's ionised fraction and temperature within 1e-5);

  * each result is the same alone and in the mix: problems of different grids,
    direction counts, timestep levels and precisions are solved one at a time
    on a one-slot server and all at once on a four-slot server, and every
    output field is bitwise equal (the kernels' arithmetic does not depend on
    scheduling, the replay form or what else runs);
  * packed stepping (max_batch 8: one stepper packs up to 8 running requests
    per step; max_batch 2: several steppers, each packing 2, side by side):
    the same results, bitwise, as the one-slot server;
  * a one-slot solve equals a direct `Sweep` life cycle of the same problem;
  * admission: requests whose estimates do not fit the budget together never
    overlap, a request above the whole budget fails with SSW_ERR_BUDGET, and
    the estimate bounds the device memory a handle holds.
"""

import dataclasses

import numpy as np
import pytest

from oracle.oracle import OracleSweep
from parity_util import rel_ok
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


@pytest.mark.parametrize("max_batch", [0, 8, 2])
def test_alone_equals_mix(max_batch, hip_available):
    probs = _problems()
    with Server(slots=1) as sv:
        alone = [sv.wait(sv.submit(pb, STEPS)) for pb in probs]
    with Server(slots=4, max_batch=max_batch) as sv:
        tickets = [sv.submit(pb, STEPS) for pb in probs]
        mix = [sv.wait(t) for t in tickets]
        stats = sv.stats()
    assert stats["peak_running"] > 1, "the four-slot server never ran two requests at once"
    if max_batch == 8:  # one stepper: the requests set up while it steps join its next packed step
        assert stats["packed_requests_per_step"] > 1, "no step was packed"
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


def test_served_direction_sweeps_equal_direct_sweep(hip_available):
    # requests asking for direction sweeps (one of them twice, so that the
    # second is served on the resident grid) give bitwise the grouped replay's
    # results of a direct Sweep
    probs = _problems()[:3]
    with Server(slots=2, max_batch=0) as sv:
        tickets = [sv.submit(pb, STEPS, grid_key=k + 1, direction_sweeps=True) for k, pb in enumerate(probs)]
        tickets.append(sv.submit(probs[0], STEPS, grid_key=1, direction_sweeps=True))
        rs = [sv.wait(t) for t in tickets]
    for r, pb in zip(rs, probs + [probs[0]]):
        assert r.status == 0, r.error
        s = Sweep(pb.params, pb.grid, *pb.state())
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


def test_resident_grid_reuse_is_bitwise(hip_available):
    probs = _problems()
    with Server(slots=4) as sv:
        first = [sv.wait(sv.submit(pb, STEPS, grid_key=i + 1)) for i, pb in enumerate(probs)]
        tickets = [sv.submit(pb, STEPS, grid_key=i + 1) for i, pb in enumerate(probs)]
        again = [sv.wait(t) for t in tickets]
        # a different state on a resident grid: a fresh handle's results
        pb = probs[0]
        other = W.Problem(pb.name, pb.grid, pb.params, pb.density * 0.5, pb.ionized_fraction, pb.temperature,
                          pb.source * 3.0, pb.steps)
        r_other = sv.wait(sv.submit(other, STEPS, grid_key=1))
    assert not any(r.reused for r in first)
    assert all(r.reused for r in again) and r_other.reused
    for i, (a, b) in enumerate(zip(first, again)):
        assert a.status == 0 and b.status == 0, (a.error, b.error)
        _same(a, b, f"request {i} fresh vs on its resident grid")
    with Server(slots=1) as sv:
        fresh = sv.wait(sv.submit(other, STEPS))
    assert not fresh.reused
    _same(r_other, fresh, "a new state on a resident grid vs a fresh handle")


@pytest.fixture(scope="module")
def serving_mix():
    pool = W.serving_pool(6, 2024)
    return pool, W.serving_requests(pool, 32, 2025, 4)


@pytest.fixture(scope="module")
def serving_alone(serving_mix):
    _, reqs = serving_mix
    with Server(slots=1) as sv:
        return [sv.wait(sv.submit(r, r.steps)) for r in reqs]


@pytest.mark.parametrize("max_batch", [0, 8, 2])
def test_serving_mix_real_sizes_alone_equals_mix(max_batch, serving_mix, serving_alone, hip_available):
    pool, reqs = serving_mix
    sizes = [g.num_cells for g, _, _ in pool]
    assert min(sizes) >= 8192 and max(sizes) <= 524288 and max(sizes) > 100000
    assert {r.params.directions for r in reqs} >= {16, 84} and {r.params.num_timestep_levels for r in reqs} >= {1, 4}
    alone = serving_alone
    with Server(slots=8 if max_batch else 4, max_batch=max_batch) as sv:
        tickets = [sv.submit(r, r.steps, grid_key=r.pool_index + 1) for r in reqs]
        mix = [sv.wait(t) for t in tickets]
        st = sv.stats()
        assert st["peak_running"] > 1
        if max_batch == 8:
            assert st["packed_requests_per_step"] > 1
    assert any(m.reused for m in mix), "some requests should run on resident grids"
    for i, (a, m) in enumerate(zip(alone, mix)):
        assert a.status == 0 and m.status == 0, (a.error, m.error)
        _same(a, m, f"request {i} ({reqs[i].grid.num_cells} cells) alone vs in the mix")


@pytest.mark.parametrize("precision,max_batch", [(64, 0), (32, 0), (64, 8)])
def test_served_small_requests_match_the_oracle(precision, max_batch, hip_available):
    # optically thin gas (1e-3 of the mix's density): the serving mix's own
    # gas leaves most rates a few cells from a source near the double (FP64)
    # or float (FP32) underflow, where the relative rate change that picks a
    # timestep level is rounding noise in both codes
    pool = W.serving_pool(3, 77, lo=8192, hi=20000)
    reqs = W.serving_requests(pool, 3, 78, 3, precision=precision, rho_scale=1e-3)
    big = W.serving_requests(W.serving_pool(1, 79, lo=200000, hi=300000), 2, 80, 3)
    with Server(slots=4, max_batch=max_batch) as sv:
        tickets = [sv.submit(r, r.steps, grid_key=10 + i) for i, r in enumerate(big)]
        tickets += [sv.submit(r, r.steps, grid_key=r.pool_index + 1) for r in reqs]
        out = [sv.wait(t) for t in tickets][len(big):]
    for r, g in zip(reqs, out):
        assert g.status == 0, g.error
        o = OracleSweep(dataclasses.replace(r.params, transport_precision=64), r.grid, *r.state(), threads=8)
        for _ in range(r.steps):
            o.run_sweeps()
        ro = o.results()
        o.close()
        assert (ro.photon_rate > 0).mean() > 0.5, "photons should reach most cells"
        assert np.array_equal(g.level, ro.level)
        tol = 1e-10 if precision == 64 else 1e-5
        rel_ok(g.ionized_hydrogen_fraction, ro.ionized_hydrogen_fraction, tol, name=f"{r.name} x_HII")
        rel_ok(g.temperature, ro.temperature, tol, name=f"{r.name} T")
        if precision == 64:
            rel_ok(g.photon_rate, ro.photon_rate, tol, atol_frac=1e-12, name=f"{r.name} photon rate")
