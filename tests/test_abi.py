"""The C ABI (CPU, no device work): libsubsweep_amd.so loads, exports every
function include/subsweep_amd.h declares, fills the reference defaults, and
rejects invalid parameters before touching a device."""

import ctypes as C
import os
import re

from subsweep_amd import _native as N

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "subsweep_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ssw_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 18
    lib = C.CDLL(N.SWEEP_LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(N.SWEEP_SIGNATURES), set(names) - set(N.SWEEP_SIGNATURES)


def test_default_parameters_are_the_reference_defaults():
    p = N.SweepParametersC()
    N.sweep_lib().ssw_default_parameters(C.byref(p))
    # src/sweep/parameters.rs:64-78
    assert p.significant_rate_threshold == 0.0
    assert p.timestep_safety_factor == 0.01 * 10.0
    assert p.chemistry_timestep_safety_factor == 0.01 * 10.0
    assert p.prevent_cooling == 1
    assert p.transport_precision == 64


def test_invalid_parameters_are_rejected():
    lib = N.sweep_lib()
    p = N.SweepParametersC()
    lib.ssw_default_parameters(C.byref(p))
    h = C.c_void_p()
    p.num_timestep_levels = 0
    assert lib.ssw_create(C.byref(p), C.byref(h)) == N.SSW_ERR_INVALID
    p.num_timestep_levels = 1
    p.transport_precision = 16
    assert lib.ssw_create(C.byref(p), C.byref(h)) == N.SSW_ERR_INVALID
    p.transport_precision = 64
    p.num_directions = 17
    assert lib.ssw_create(C.byref(p), C.byref(h)) == N.SSW_ERR_INVALID
    assert b"num_directions" in lib.ssw_last_error()
    assert not h.value


def test_direction_tables_match_the_reference():
    # host-only call: the tables the library uses for DirectionsSpecification::Num
    # against the reference's own (src/sweep/direction/healpix.rs, read as data
    # into tests/golden/reference_tables.json), number for number; the oracle
    # takes its bins from the same data
    import numpy as np
    import pytest
    from oracle.oracle import OracleSweep, reference_tables
    from subsweep_amd import direction_table
    from subsweep_amd import workloads as W
    t = reference_tables()["directions"]
    assert sorted(int(k) for k in t) == [16, 21, 32, 64, 84]
    for num in (16, 21, 32, 64, 84):
        ref = np.array(t[str(num)]["bins"])
        assert ref.shape == (num, 3)
        assert np.array_equal(direction_table(num), ref), num
        pb = W.cartesian_problem(2, num, levels=1, periodic=False, source_cells=(0,))
        o = OracleSweep(pb.params, pb.grid, *pb.state())
        assert np.array_equal(o.directions(), ref), num
        o.close()
    # one direction is [1, 0, 0] (src/sweep/direction/mod.rs:62)
    assert np.array_equal(direction_table(1), [[1.0, 0.0, 0.0]])
    with pytest.raises(N.SweepError):
        direction_table(5)


def test_device_source_placement_rejects_bad_arguments_before_the_device():
    lib = N.sweep_lib()
    one = (C.c_double * 3)(0.0, 0.0, 0.0)
    out = (C.c_double * 1)()
    assert lib.ssw_assign_sources_device(0, one, 1, one, one, out, None) == N.SSW_ERR_INVALID
    assert lib.ssw_assign_sources_device(1, one, -1, one, one, out, None) == N.SSW_ERR_INVALID
    assert lib.ssw_assign_sources_device(1, None, 1, one, one, out, None) == N.SSW_ERR_INVALID
    # no sources: every cell's rate is zero, nothing touches the device
    out[0] = 5.0
    assert lib.ssw_assign_sources_device(1, one, 0, None, None, out, None) == N.SSW_OK
    assert out[0] == 0.0
