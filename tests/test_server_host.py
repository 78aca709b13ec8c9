"""The request loop's host logic (CPU, no device work): argument checks of
the server constructors, and a request whose handle cannot be created (no
device in this process) coming back with its error instead of blocking,
on the one-request-per-worker server and on the packed server."""

import ctypes as C

import pytest

from subsweep_amd import _native as N
from subsweep_amd import workloads as W
from subsweep_amd.serve import Server


def _no_device():
    try:
        import torch
        return not torch.cuda.is_available()
    except ImportError:
        return True


def test_server_constructors_reject_bad_arguments():
    lib = N.sweep_lib()
    h = C.c_void_p()
    assert lib.ssw_server_create(0, 1 << 30, C.byref(h)) == N.SSW_ERR_INVALID
    assert lib.ssw_server_create(4, 0, C.byref(h)) == N.SSW_ERR_INVALID
    assert lib.ssw_server_create_packed(4, 1 << 30, -1, C.byref(h)) == N.SSW_ERR_INVALID
    assert lib.ssw_server_create_packed(4, 1 << 30, 65, C.byref(h)) == N.SSW_ERR_INVALID
    assert lib.ssw_server_pack_stats(None, None, None) == N.SSW_ERR_INVALID


@pytest.mark.parametrize("max_batch", [0, 1, 8])
def test_server_starts_and_stops_without_requests(max_batch):
    with Server(slots=3, hbm_budget_bytes=1 << 30, max_batch=max_batch) as sv:
        st = sv.stats()
    assert st["peak_running"] == 0 and st["packed_steps"] == 0


@pytest.mark.skipif(not _no_device(), reason="needs a process without a device")
@pytest.mark.parametrize("max_batch", [0, 8])
def test_failed_setup_returns_the_error(max_batch):
    pb = W.cartesian_problem(3, 16, levels=1)
    with Server(slots=2, hbm_budget_bytes=1 << 34, max_batch=max_batch) as sv:
        rs = [sv.wait(sv.submit(pb, 2)) for _ in range(3)]
    for r in rs:
        assert r.status != 0 and r.error
