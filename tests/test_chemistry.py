"""Hydrogen chemistry (CPU).

The oracle (oracle/oracle_sweep.cpp, restating src/chemistry/hydrogen_only/mod.rs)
is pinned by:
  * the reference's own derivative-consistency tests (mod.rs:491-591),
  * the textbook case-B recombination coefficient at 1e4 K,
  * the collisional equilibrium x* = beta / (alpha + beta), the fixed point of
    the implicit ionised-fraction update (mod.rs:334-354),
  * the reference's no-panic extreme states (mod.rs:948-984) and the clamp
    bounds (mod.rs:356-369).
The library's single-cell chemistry (the code the device kernel runs,
compiled for the host) must then agree with the oracle bit for bit on random
states: both are transcriptions of the same expressions with the same libm.
"""

import ctypes as C
import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as O
from subsweep_amd import _native as N

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")))
MYR = 1e6 * 3.15576e7
GRAMS_PER_CM3 = 1e-3 / 1e-6
KPC = 1000 * 3.0857e16


def test_rate_derivatives_match_finite_differences():
    g = GOLD["chemistry_derivative_check"]
    for fname, dname in g["pairs"]:
        for T in g["temperatures"]:
            a = O.solver_rates(T, 0.0, 0.0, 1.0)
            b = O.solver_rates(T + g["delta"], 0.0, 0.0, 1.0)
            analytic = a[dname]
            numeric = (b[fname] - a[fname]) / g["delta"]
            rel = abs(analytic - numeric) / (abs(analytic) + abs(numeric) + 1e-50)
            assert rel < g["epsilon"], (fname, T, analytic, numeric)


def test_a_plausible_sign_error_fails_the_derivative_check():
    # the check above is not vacuous: a derivative with the wrong sign fails it
    g = GOLD["chemistry_derivative_check"]
    T = 1e4
    a = O.solver_rates(T)
    b = O.solver_rates(T + g["delta"])
    numeric = (b["case_b_recombination_rate"] - a["case_b_recombination_rate"]) / g["delta"]
    wrong = -a["case_b_recombination_rate_derivative"]
    assert abs(wrong - numeric) / (abs(wrong) + abs(numeric)) > g["epsilon"]


def test_case_b_recombination_textbook_value():
    # alpha_B(1e4 K) = 2.59e-13 cm^3 s^-1 (Osterbrock & Ferland); the Hui & Gnedin
    # fit the reference uses reproduces it to about 1 %.
    alpha = O.solver_rates(1e4)["case_b_recombination_rate"] / 1e-6
    assert alpha == pytest.approx(2.59e-13, rel=0.02)


def test_collisional_equilibrium_fixed_point():
    # Without photons the implicit update dx = dt (c - x (c + d)) / (1 - j dt)
    # (mod.rs:334-354, c = beta n_e, d = alpha n_e) vanishes at x* = beta / (alpha + beta)
    # and moves x towards it from either side. A short step keeps the temperature
    # change (applied first, mod.rs:376-382) negligible.
    T = 1e5
    rho = 1e-24 * GRAMS_PER_CM3
    r = O.solver_rates(T)
    x_star = r["collisional_ionization_rate"] / (r["collisional_ionization_rate"] + r["case_b_recombination_rate"])
    dt = 1e-4 * MYR
    at = O.perform_timestep(x_star, T, rho, 1.0, 1.0, 0.0, dt, 0.1)
    below = O.perform_timestep(0.99 * x_star, T, rho, 1.0, 1.0, 0.0, dt, 0.1)
    far_below = O.perform_timestep(0.5, T, rho, 1.0, 1.0, 0.0, dt, 0.1)
    assert abs(at["x"] - x_star) < 1e-6 * abs(below["x"] - 0.99 * x_star)
    assert below["x"] > 0.99 * x_star
    assert far_below["x"] > 0.5
    # a hotter gas has the higher collisional equilibrium
    r2 = O.solver_rates(2e5)
    assert r2["collisional_ionization_rate"] / (r2["collisional_ionization_rate"] + r2["case_b_recombination_rate"]) > x_star


def test_extreme_states_do_not_fail_and_stay_clamped():
    e = GOLD["chemistry_extreme_states"]
    for x0 in e["ionized_fractions"]:
        out = O.perform_timestep(x0, e["temperature_K"], e["density_g_per_cm3"] * GRAMS_PER_CM3, e["volume_m3"],
                                 e["length_kpc"] * KPC, e["rate_per_s"], MYR, 0.1, e["scale_factor"], floor=False)
        assert math.isfinite(out["x"]) and math.isfinite(out["T"]) and out["timescale"] > 0
        assert 1e-10 <= out["x"] <= 1 - 1e-10


def _lib_step(x, T, rho, vol, length, rate, a, floor, dt, safety):
    st = np.array([x, T], np.float64)
    proc, failed = C.c_int32(), C.c_int32()
    ts = N.sweep_lib().ssw_chemistry_timestep(st.ctypes.data_as(C.POINTER(C.c_double)), rho, vol, length, rate, a,
                                              1 if floor else 0, dt, safety, C.byref(proc), C.byref(failed))
    return {"x": st[0], "T": st[1], "timescale": ts, "process": proc.value, "failed": bool(failed.value)}


def test_library_chemistry_matches_oracle_bitwise():
    rng = np.random.default_rng(11)
    n = 3000
    mismatches = []
    for i in range(n):
        T = 10 ** rng.uniform(1.5, 7)
        x = 10 ** rng.uniform(-10, 0) * (1 - 1e-10)
        rho = 10 ** rng.uniform(-30, -22) * GRAMS_PER_CM3
        size = 10 ** rng.uniform(-1, 2) * KPC
        vol = 4.0 / 3.0 * math.pi * size ** 3
        rate = 0.0 if rng.random() < 0.2 else 10 ** rng.uniform(30, 52)
        dt = 10 ** rng.uniform(-3, 1) * MYR
        floor = bool(rng.random() < 0.5)
        a = 10 ** rng.uniform(0, 1)
        o = O.perform_timestep(x, T, rho, vol, size, rate, dt, 0.1, a, floor)
        l = _lib_step(x, T, rho, vol, size, rate, a, floor, dt, 0.1)
        if (o["x"], o["T"], o["timescale"], o["process"], o["failed"]) != (l["x"], l["T"], l["timescale"],
                                                                          l["process"], l["failed"]):
            mismatches.append((i, o, l))
    assert not mismatches, mismatches[:3]


def test_library_extreme_states():
    e = GOLD["chemistry_extreme_states"]
    for x0 in e["ionized_fractions"]:
        out = _lib_step(x0, e["temperature_K"], e["density_g_per_cm3"] * GRAMS_PER_CM3, e["volume_m3"],
                        e["length_kpc"] * KPC, e["rate_per_s"], e["scale_factor"], False, MYR, 0.1)
        assert math.isfinite(out["x"]) and math.isfinite(out["T"])
        assert 1e-10 <= out["x"] <= 1 - 1e-10
