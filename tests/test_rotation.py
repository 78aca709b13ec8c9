"""rotate_directions (CPU): the rotation matrices of the library's C ABI and
the oracle's rotation of directions and remap of the rates.

src/sweep/direction/mod.rs:
  * get_rotation_matrix (:112-135), get_random_rotation_matrix (:137-148);
    the reference's own test rotation_matrix_has_determinant_1 (:218-225);
  * rotate_directions_system (:158-175): every direction d becomes M . d;
  * remap (:190-205): `for i { for j { values[i] = old[j] * kernel[i][j] } }`
    overwrites values[i] for every j, so only j = D-1 survives:
    values[i] = old[D-1] * kernel[i][D-1], with kernel[i][j] = 1 for the last
    new direction j most aligned with old direction i (kernel_f :177-188).
A rotation by the identity therefore keeps direction D-1 and zeroes every
other one, which pins the overwrite (a remap that moved each direction's
value to its best-aligned new direction would keep them all).
"""

import numpy as np
import pytest

from oracle.oracle import OracleSweep
from subsweep_amd import random_rotation_matrix, rotation_matrix
from subsweep_amd import workloads as W

MYR = 3.15576e13


def test_rotation_matrix_has_determinant_1():
    rng = np.random.default_rng(1337)
    for _ in range(100):
        m = random_rotation_matrix(rng)
        assert abs(np.linalg.det(m) - 1.0) < 1e-12
        assert np.allclose(m @ m.T, np.eye(3), rtol=0, atol=1e-14)


@pytest.mark.parametrize("angle", [0.0, 0.3, 2.0, -1.1])
def test_rotation_matrix_closed_forms(angle):
    axis = np.array([1.0, 2.0, -2.0]) / 3.0
    m = rotation_matrix(axis, angle)
    assert np.allclose(m @ axis, axis, rtol=0, atol=1e-15)  # the axis stays put
    assert abs(np.trace(m) - (1.0 + 2.0 * np.cos(angle))) < 1e-14
    # a vector orthogonal to the axis turns by the angle, counter-clockwise about the axis
    u = np.cross(axis, [0.0, 0.0, 1.0])
    u /= np.linalg.norm(u)
    v = m @ u
    assert abs(np.dot(u, v) - np.cos(angle)) < 1e-14
    assert np.allclose(np.cross(u, v), np.sin(angle) * axis, rtol=0, atol=1e-14)


def test_random_rotation_matrix_axis_and_angle():
    # the draws phi, u, psi: axis of polar angle acos(2u - 1), azimuth phi; rotation by psi
    import ctypes as C
    from subsweep_amd import _native as N
    phi, u, psi = 1.2, 0.8, 0.7
    out = np.zeros(9)
    N.check(N.sweep_lib().ssw_random_rotation_matrix(phi, u, psi, N.ptr(out, C.c_double)))
    theta = np.arccos(2 * u - 1)
    axis = np.array([np.cos(phi) * np.sin(theta), np.sin(phi) * np.sin(theta), np.cos(theta)])
    assert np.allclose(out.reshape(3, 3), rotation_matrix(axis, psi), rtol=0, atol=1e-15)


def _problem():
    return W.jittered_problem(6, 1.0, "3", 16, levels=1, periodic=True, steps=2,
                              rho_mean=1e-3 * W.RHO_MEAN).with_params(max_timestep=1e-3 * MYR)


def _run(pb, steps):
    o = OracleSweep(pb.params, pb.grid, *pb.state())
    for _ in range(steps):
        o.run_sweeps()
    return o


def test_identity_rotation_keeps_only_the_last_direction():
    pb = _problem()
    o = _run(pb, 2)
    before = [a.copy() for a in o.fluxes()]
    assert (before[1] > 0).mean() > 0.5, "the fluxes should be non-zero"
    o.rotate_directions(np.eye(3))
    after = o.fluxes()
    for b, a in zip(before, after):
        assert np.array_equal(a[-1], b[-1])
        assert np.all(a[:-1] == 0.0)
    o.close()


def test_rotation_moves_directions_and_remaps_by_the_last_direction():
    pb = _problem()
    o = _run(pb, 2)
    dirs = o.directions().copy()
    before = [a.copy() for a in o.fluxes()]
    m = rotation_matrix(np.array([0.0, 0.6, 0.8]), 0.9)
    o.rotate_directions(m)
    new = o.directions()
    assert np.allclose(new, dirs @ m.T, rtol=0, atol=1e-15)
    # kernel[i][D-1]: 1 where the last new direction is the last maximum of 1 + old_i . new_j
    D = dirs.shape[0]
    kept = np.zeros(D)
    for i in range(D):
        v = 1.0 + new @ dirs[i]
        kept[i] = 1.0 if np.flatnonzero(v == v.max())[-1] == D - 1 else 0.0
    for b, a in zip(before, o.fluxes()):
        assert np.array_equal(a, b[-1][None, :] * kept[:, None])
    o.close()


def test_counts_follow_the_rotated_directions():
    pb = _problem()
    o = _run(pb, 1)
    m = rotation_matrix(np.array([1.0, 0.0, 0.0]), 0.4)
    o.rotate_directions(m)
    dirs = o.directions()
    g = pb.grid
    counts = o.counts(0)
    want = np.zeros_like(counts)
    for c in range(g.num_cells):
        for f in range(g.offsets[c], g.offsets[c + 1]):
            if g.kind[f] != 0:
                continue
            dd = (g.normal[f, 0] * dirs[:, 0] + g.normal[f, 1] * dirs[:, 1]) + g.normal[f, 2] * dirs[:, 2]
            want[:, c] += dd < 0.0
    assert np.array_equal(counts, want)
    o.close()
