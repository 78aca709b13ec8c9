"""Star-particle sources (CPU): the BPASS lookup and the star -> source
conversion of the reference's Arepo input path
(src/arepo_postprocess/bpass.rs:7-56, src/arepo_postprocess/sources.rs:57-128,
src/cosmology.rs:80-107), against tests/golden/reference_tables.json (the
reference's table and test values, read from its source as data):

  * the reference's own tests: bpass check_is_close (bpass.rs:893-913) and the
    cosmology time difference (cosmology.rs:212-220);
  * at every interior grid point of the table the lookup returns the table
    value exactly, and between grid points it is linear in age and in
    metallicity (midpoints are the means of their neighbours);
  * ages and metallicities outside the table clamp to 1.00001 inside its
    ends; the rate scales with the mass;
  * the cosmic time difference equals the integral of da / (a H(a)) for a
    flat cosmology (scipy quadrature), which a dropped factor would fail.
"""

import json
import math
import os

import numpy as np
import pytest
from scipy.integrate import quad

from subsweep_amd import sources as S

HERE = os.path.dirname(os.path.abspath(__file__))
TABLES = json.load(open(os.path.join(HERE, "golden", "reference_tables.json")))
BP = TABLES["bpass"]
AGES = np.array(BP["age_bins_yr"])
METS = np.array(BP["metallicity_bins"])
TABLE = np.array(BP["photons_per_s_per_solar_mass"])
COSMO = TABLES["cosmology"]
HUBBLE = 3.2407789e-18


def _close(x, y, tol):
    # assert_float_is_close of the reference's bpass test
    return abs((x - y) / (abs(x) + abs(y))) < tol


def test_reference_bpass_known_answers():
    assert TABLE.shape == (51, 13)
    for k in BP["known_answers"]:
        got = S.bpass_source_rate(k["age_yr"], k["metallicity"], 1.0)  # Mass::solar(1.0)
        assert _close(got, k["photons_per_s_per_solar_mass"], BP["known_answer_rel_tol"]), (k, got)


def test_table_points_are_exact():
    for i in range(1, len(AGES) - 1):
        for j in range(1, len(METS) - 1):
            assert S.bpass_source_rate(AGES[i], METS[j], 1.0) == TABLE[i, j], (i, j)


def test_linear_between_table_points():
    for i in range(0, len(AGES) - 1, 5):
        for j in range(0, len(METS) - 1, 3):
            a = 0.5 * (AGES[i] + AGES[i + 1])
            z = 0.5 * (METS[j] + METS[j + 1])
            if i == 0 or j == 0:  # keep inside the clamped range
                continue
            want_age = 0.5 * (TABLE[i, j] + TABLE[i + 1, j])
            assert math.isclose(S.bpass_source_rate(a, METS[j], 1.0), want_age, rel_tol=1e-13)
            corner = 0.25 * (TABLE[i, j] + TABLE[i + 1, j] + TABLE[i, j + 1] + TABLE[i + 1, j + 1])
            assert math.isclose(S.bpass_source_rate(a, z, 1.0), corner, rel_tol=1e-13)


def test_clamping_and_mass_scaling():
    lo_age, hi_age = AGES[0] * 1.00001, AGES[-1] / 1.00001
    assert S.bpass_source_rate(1.0, 0.02, 1.0) == S.bpass_source_rate(lo_age, 0.02, 1.0)
    assert S.bpass_source_rate(1e13, 0.02, 1.0) == S.bpass_source_rate(hi_age, 0.02, 1.0)
    assert S.bpass_source_rate(1e8, 1.0, 1.0) == S.bpass_source_rate(1e8, METS[-1] / 1.00001, 1.0)
    assert S.bpass_source_rate(1e8, 0.0, 1.0) == S.bpass_source_rate(1e8, METS[0] * 1.00001, 1.0)
    base = S.bpass_source_rate(3e7, 0.01, 1.0)
    assert S.bpass_source_rate(3e7, 0.01, 1e4) == base * 1e4


def _time_integral(a0, a1, ol, om, h):
    f = lambda a: 1.0 / (a * math.sqrt(om / a ** 3 + ol))
    return quad(f, a0, a1, epsabs=0, epsrel=1e-13, limit=200)[0] / (HUBBLE * h)


def test_cosmology_time_difference():
    ol, om, h = COSMO["omega_lambda"], COSMO["omega_0"], COSMO["little_h"]
    for case in COSMO["time_difference"]:
        got = S.time_between_scale_factors(case["a0"], case["a1"], ol, om, h)
        want = case["gigayears"] * 1e9 * S.YEAR
        assert abs(got - want) <= case["abs_tol_years"] * S.YEAR, (case, got)
    for a0, a1 in ((0.01, 1.0), (0.1, 0.5), (0.5, 0.9)):
        got = S.time_between_scale_factors(a0, a1, ol, om, h)
        assert math.isclose(got, _time_integral(a0, a1, ol, om, h), rel_tol=1e-9)
    assert math.isnan(S.time_between_scale_factors(0.5, 1.0, 0.5, 0.3, h))  # not flat


def test_star_sources():
    ol, om, h = COSMO["omega_lambda"], COSMO["omega_0"], COSMO["little_h"]
    z = np.array([0.02, 0.001, 0.01, 0.03])
    a_form = np.array([0.95, -0.3, 0.5, 0.99])  # the second is a wind particle
    m = np.array([1e5, 1e5, 2e6, 3e4])
    rate, star = S.star_sources(z, a_form, m, 0.1, ol, om, h, 1.0)
    assert star.tolist() == [True, False, True, True]
    assert rate[1] == 0.0
    for i in (0, 2, 3):
        age = S.time_between_scale_factors(a_form[i], 1.0, ol, om, h) / S.YEAR
        assert rate[i] == S.bpass_source_rate(age, z[i], m[i]) * 0.1
    with pytest.raises(Exception):
        S.star_sources(z, a_form[:2], m, 0.1, ol, om, h, 1.0)
