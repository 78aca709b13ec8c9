"""CPU: a host-thread model of the ticket/done protocol of the persistent
transport kernel that was removed after it did not terminate on the device
(git history: commits f0715e8 and f1de03b). The model
(`tests/protocol_model/ticket_model.cpp`) rebuilds the kernel's per-slot chunk
lists the way its list-building kernel did and runs its workgroup loop on host
threads: tickets from a per-slot counter, a wait until the direction's done
counter reaches the tasks of the earlier level sets, done += chunk size.
Workgroups start in index order, at most `resident` at a time, as the device
dispatches them.

The schedules are the oracle's Kahn level sets of the workload of the test
that did not terminate (16^3 jittered Voronoi cells, 84 directions, periodic,
three timestep levels, after three steps so that the restricted active sets
of sweep levels 1 and 2 differ from the full set). Every run must end with
every task solved once, no chunk released before the earlier level sets of
its direction were complete, and no wait longer than its bound. Planted
mistakes in the model (a wait that includes the chunk's own level set; lists
with pairs of levels swapped) must be caught as waits that never end."""

import shutil
import subprocess

import numpy as np
import pytest

from oracle.oracle import OracleSweep
from subsweep_amd import workloads as W

MODEL = __import__("pathlib").Path(__file__).parent / "protocol_model" / "ticket_model.cpp"
NONE = 0xFFFFFFFF


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    exe = tmp_path_factory.mktemp("model") / "ticket_model"
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", str(MODEL), "-o", str(exe)], check=True)
    return exe


@pytest.fixture(scope="module")
def schedules():
    pb = W.jittered_problem(16, 1.0, "8", 84, levels=3, periodic=True, steps=5, precision=64)
    pb = pb.with_params(max_timestep=1e-2 * 3.15576e13)
    o = OracleSweep(pb.params, pb.grid, *pb.state(), threads=8)
    full = o.levelsets(0)
    kmax = int(full.max()) + 1
    for _ in range(3):
        o.run_sweeps()
    out = []
    for lvl in range(3):
        dep = o.levelsets(lvl).astype(np.int64)
        w = np.zeros((dep.shape[0], kmax), np.int64)
        for d in range(dep.shape[0]):
            act = dep[d] != NONE
            w[d] = np.bincount(dep[d][act], minlength=kmax)
        out.append(w)
    o.close()
    assert out[1].sum() < out[0].sum(), "the restricted active set should be smaller than the full set"
    return pb.grid.num_cells, out


def _run(model, n, w, nslots, workgroups, resident, seed, wait_ms=5000, mistake=0):
    text = f"{w.shape[0]} {w.shape[1]} {n}\n" + "\n".join(" ".join(map(str, r)) for r in w) + "\n"
    args = [str(model), str(nslots), str(workgroups), str(resident), str(seed), "16", str(wait_ms)]
    if mistake:
        args.append(str(mistake))
    r = subprocess.run(args, input=text, capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout.strip()


# (slots, workgroups, resident): the device's 8 XCD slots with 2048 workgroups
# (256 CUs x 8) and several residency limits, one slot, and fewer workgroups
# than slots would need
SHAPES = [(8, 2048, 64), (8, 2048, 1), (8, 2048, 256), (1, 2048, 16), (3, 5, 2), (8, 9, 3)]


@pytest.mark.parametrize("lvl", [0, 1, 2])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_protocol_terminates_on_the_kernel_schedules(model, schedules, lvl, shape):
    n, ws = schedules
    for seed in (1, 2):
        rc, out = _run(model, n, ws[lvl], *shape, seed)
        assert rc == 0 and out.startswith("ok"), out
        assert "dropped=0 " in out and out.endswith("short_directions=0"), out


def test_slots_without_workgroups_are_left_to_the_host(model, schedules):
    # 4 workgroups for 8 slots: the directions of slots 4-7 (10 each of the
    # 84) are never swept in the launch; the host's check re-runs the sweep
    n, ws = schedules
    rc, out = _run(model, n, ws[0], 8, 4, 4, 3)
    assert rc == 0 and out.endswith("short_directions=40"), out


@pytest.mark.parametrize("mistake", [1, 2])
def test_model_catches_planted_mistakes(model, schedules, mistake):
    n, ws = schedules
    rc, out = _run(model, n, ws[0], 8, 2048, 1, 1, wait_ms=300, mistake=mistake)
    assert rc == 1 and out.startswith("FAIL wait not satisfied"), out
