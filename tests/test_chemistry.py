"""Hydrogen chemistry (CPU).

The oracle (oracle/oracle_sweep.cpp, restating src/chemistry/hydrogen_only/mod.rs)
is pinned by:
  * the reference's own derivative-consistency tests (mod.rs:491-591),
  * the textbook case-B recombination coefficient at 1e4 K,
  * the collisional equilibrium x* = beta / (alpha + beta), the fixed point of
    the implicit ionised-fraction update (mod.rs:334-354),
  * the reference's no-panic extreme states (mod.rs:948-984) and the clamp
    bounds (mod.rs:356-369),
  * for the optional chemistry outputs (src/sweep/chemistry_output.rs:25-56):
    recombinations balance collisional ionisations at x*, heating is minus
    the cooling without photons and the photoheating per photoionisation is
    E_photon - E_Rydberg, every photon is absorbed in the optically thick
    limit and the rate stops depending on the density in the thin one.
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


PROTON_MASS = 1.67262192369e-27
EV = 1.602176634e-19


def test_chemistry_outputs_pinned():
    T = 2e4
    rho = 1e-24 * GRAMS_PER_CM3
    size = 0.5 * KPC
    vol = 4.0 / 3.0 * math.pi * size ** 3
    r = O.solver_rates(T)
    beta, alpha = r["collisional_ionization_rate"], r["case_b_recombination_rate"]
    x_star = beta / (alpha + beta)
    # no photons, collisional equilibrium: recombinations balance collisional
    # ionisations, nothing is photoionised and the heating is minus the cooling
    out = O.chemistry_outputs_cell(x_star, T, rho, vol, size, 0.0)
    assert out["recombination_rate"] > 0
    assert out["recombination_rate"] == pytest.approx(out["collisional_ionization_rate"], rel=1e-12, abs=0)
    assert out["photoionization_rate"] == 0.0
    assert out["heating_rate"] == pytest.approx(-O.solver_rates(T, x_star, rho)["cooling_rate"], rel=1e-12, abs=0)
    # optically thick (tau ~ 3e3 here): every photon of the rate R ionises one
    # of the n_HI V neutral atoms, Gamma n_HI V = R; each photoionisation
    # heats by E_photon - E_Rydberg = 18.0284 - 13.6569 eV
    x, R = 0.3, 1e48
    out = O.chemistry_outputs_cell(x, T, rho, vol, size, R)
    n_hi = rho / PROTON_MASS * (1 - x)
    assert out["photoionization_rate"] * n_hi * vol == pytest.approx(R, rel=1e-12, abs=0)
    cool = O.solver_rates(T, x, rho)["cooling_rate"]
    per_ionisation = (out["heating_rate"] + cool) / (out["photoionization_rate"] * n_hi)
    assert per_ionisation == pytest.approx((18.028356312818811 - 13.65693) * EV, rel=1e-9, abs=0)
    # optically thin (tau ~ 6e-7, where 1 - exp(-tau) still holds ~10 digits):
    # Gamma = R sigma L / V does not depend on the density (to ~tau / 2);
    # sigma is the number-weighted cross-section, ~2.96e-18 cm^2
    thin = [O.chemistry_outputs_cell(x, T, d, vol, size, R)["photoionization_rate"] for d in (3e-31, 6e-31)]
    assert thin[0] == pytest.approx(thin[1], rel=1e-6, abs=0)
    sigma = thin[0] * vol / (R * size)
    assert sigma == pytest.approx(2.958e-18 * 1e-4, rel=1e-3, abs=0)


def test_a_dropped_factor_in_the_outputs_fails_the_pins():
    # the balance above tells apart the recombination rate with and without its x
    T = 2e4
    rho = 1e-24 * GRAMS_PER_CM3
    r = O.solver_rates(T)
    x_star = r["collisional_ionization_rate"] / (r["collisional_ionization_rate"] + r["case_b_recombination_rate"])
    out = O.chemistry_outputs_cell(x_star, T, rho, 1.0, 1.0, 0.0)
    planted = out["recombination_rate"] / x_star  # alpha n_e without x
    assert planted != pytest.approx(out["collisional_ionization_rate"], rel=1e-3, abs=0)
