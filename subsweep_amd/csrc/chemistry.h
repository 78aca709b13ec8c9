// FP64 hydrogen-only chemistry for one cell (device and host), after
// HydrogenOnly / Solver in src/chemistry/hydrogen_only/mod.rs:68-461.
//
// Rates are the Rosdahl et al. (2015) fits the reference uses; operation order
// follows the reference expression by expression (no FMA contraction: the
// library is compiled with -ffp-contract=off). The recursive sub-stepping of
// perform_timestep_internal (:394-424) is run as an explicit depth-first walk:
// a rejected step of depth k is replaced by two steps of depth k+1; the second
// one is remembered as a bit in `pending` (at most one per depth can be
// outstanding), and the walk always resumes at the deepest pending step. The
// returned timescale is the one of the last accepted step, as in the reference.
#pragma once

#include <cfloat>
#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define SSW_HD __host__ __device__ inline
#else
#define SSW_HD inline
#endif

namespace ssw {

// Unit factors and constants, src/units/mod.rs:19-108 (products as diman forms them).
constexpr double kCm = 0.01;
constexpr double kCm2 = kCm * kCm;
constexpr double kCm3 = 1e-6;
constexpr double kCm3PerS = kCm3 / 1.0;
constexpr double kErg = 1e-7 * 1.0;
constexpr double kErgCm3PerS = kErg * kCm3PerS;
constexpr double kErgPerS = kErg / 1.0;
constexpr double kEv = 1.602176634e-19 * 1.0;
constexpr double kBoltzmann = 1.380649e-23 * 1.0 / 1.0;
constexpr double kProtonMass = 1.67262192369e-27;
constexpr double kGamma = 5.0 / 3.0;
constexpr double kSigma = 2.9580524545305314e-18 * kCm2;  // NUMBER_WEIGHTED_AVERAGE_CROSS_SECTION
constexpr double kPhotonEnergy = 18.028356312818811 * kEv;
constexpr double kRydberg = 13.65693 * kEv;
constexpr double kXEps = 1e-10;             // IONIZED_HYDROGEN_FRACTION_EPSILON
constexpr double kInvEps = 1.0 / DBL_EPSILON;  // 1.0 / f64::EPSILON
constexpr int kMaxDepth = 100;

enum { kProcTemperature = 0, kProcIonization = 1, kProcPhotonRate = 2 };

// f64::powi lowered by LLVM to binary exponentiation.
SSW_HD double powi(double a, int b) {
    bool recip = b < 0;
    double r = 1.0;
    for (;;) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

// Attenuation of one cell: exp(-(rho/m_p (1-x)) sigma L), the non-absorbed
// fraction of HydrogenOnly::get_outgoing_rate (hydrogen_only/mod.rs:78-86).
SSW_HD double non_absorbed_fraction(double density, double xhii, double size) {
    double n = density / kProtonMass * (1.0 - xhii);
    return exp(-n * kSigma * size);
}

struct Cell {
    double x, T;              // ionised fraction, temperature
    double rho, vol, len, rate, a;
};

// --- rate fits (hydrogen_only/mod.rs:161-302) ---------------------------------
SSW_HD double fit_collision(double t) {
    return sqrt(t) / (1.0 + sqrt(t / 1e5)) * exp(-157809.1 / t);
}
SSW_HD double fit_collision_dt(double t) {
    const double k1 = 1.0 / 1e5, k2 = 157809.1;
    double s = sqrt(k1 * t);
    return (exp(-k2 / t) * (k1 * k2 * t + 0.5 * s * (2.0 * k2 + t))) /
           (sqrt(powi(t, 3)) * s * powi(s + 1.0, 2));
}
SSW_HD double alpha_b(double T) {
    double l = 315614.0 / T;
    return (2.753e-14 * pow(l, 1.5) / pow(1.0 + pow(l / 2.74, 0.407), 2.242)) * kCm3PerS;
}
SSW_HD double alpha_b_dt(double T) {
    double l = 315614.0 / T;
    double dl = -315614.0 / (T * T);
    const double k1 = 1.0 / 2.74, k2 = 0.407, k3 = 2.242;
    double p = pow(k1 * l, k2);
    double d = -sqrt(l) * pow(p + 1.0, -k3 - 1.0) * (k2 * k3 * p - 1.5 * p - 1.5);
    return ((2.753e-14 * d) * kCm3PerS) * dl;
}
SSW_HD double recomb_cooling(double T) {
    double l = 315614.0 / T;
    return (3.435e-30 * T * pow(l, 1.97) / pow(1.0 + pow(l / 2.25, 0.376), 3.72)) * kErgCm3PerS;
}
SSW_HD double recomb_cooling_dt(double T) {
    const double k1 = 315614.0, k2 = 1.97, k3 = 0.376, k4 = 3.72, k5 = 2.25;
    double q = pow(k1 / (k5 * T), k3);
    double v = pow(1.0 + q, -1.0 - k4) * (1.0 - 1.0 * k2 + (1.0 - 1.0 * k2 + k3 * k4) * q) * pow(k1 / T, k2);
    return ((3.435e-30 * v) * kErgCm3PerS) / 1.0;
}
SSW_HD double beta_ci(double T) { return (5.85e-11 * fit_collision(T)) * kCm3PerS; }
SSW_HD double beta_ci_dt(double T) { return ((5.85e-11 * fit_collision_dt(T)) * kCm3PerS) / 1.0; }
SSW_HD double ci_cooling(double T) { return (1.27e-21 * fit_collision(T)) * kErgCm3PerS; }
SSW_HD double ci_cooling_dt(double T) { return ((1.27e-21 * fit_collision_dt(T)) * kErgCm3PerS) / 1.0; }
SSW_HD double ce_cooling(double T) {
    return (7.5e-19 / (1.0 + sqrt(T / 1e5)) * exp(-118348.0 / T)) * kErgCm3PerS;
}
SSW_HD double ce_cooling_dt(double T) {
    const double k1 = 7.5e-19, k2 = 118348.0, k3 = 1.0 / 1e5;
    double s = sqrt(k3 * T);
    double v = (k1 * exp(-k2 / T) * (k2 * k3 * T - 0.5 * k3 * powi(T, 2) + k2 * s)) /
               (powi(T, 2) * s * powi(1.0 + s, 2));
    return (v * kErgCm3PerS) / 1.0;
}
SSW_HD double brems_cooling(double T) { return (1.42e-27 * sqrt(T)) * kErgCm3PerS; }
SSW_HD double brems_cooling_dt(double T) { return ((1.42e-27 / (2.0 * sqrt(T))) * kErgCm3PerS) / 1.0; }
SSW_HD double compton_cooling(double T, double a) {
    double x = 2.727 / a;
    return (1.017e-37 * powi(x, 4) * (T - x)) * kErgPerS;
}
SSW_HD double compton_cooling_dt(double a) {
    double x = 2.727 / a;
    return ((1.017e-37 * powi(x, 4)) * kErgPerS) / 1.0;
}

// --- derived quantities -----------------------------------------------------------
SSW_HD double n_h(const Cell& c) { return c.rho / kProtonMass; }
SSW_HD double n_hii(const Cell& c) { return n_h(c) * c.x; }
SSW_HD double n_hi(const Cell& c) { return n_h(c) * (1.0 - c.x); }
SSW_HD double mu(const Cell& c) { return 1.0 / (c.x + 1.0); }

SSW_HD double cooling(const Cell& c) {
    double ne = n_hii(c), n0 = n_hi(c), n1 = n_hii(c);
    double coll = (ce_cooling(c.T) + ci_cooling(c.T)) * ne * n0;
    double rec = recomb_cooling(c.T) * ne * n1;
    double br = brems_cooling(c.T) * ne * n1;
    double co = compton_cooling(c.T, c.a) * ne;
    return coll + rec + br + co;
}
SSW_HD double cooling_dt(const Cell& c) {
    double ne = n_hii(c), n0 = n_hi(c), n1 = n_hii(c);
    double coll = (ce_cooling_dt(c.T) + ci_cooling_dt(c.T)) * ne * n0;
    double rec = recomb_cooling_dt(c.T) * ne * n1;
    double br = brems_cooling_dt(c.T) * ne * n1;
    double co = compton_cooling_dt(c.a) * ne;
    return coll + rec + br + co;
}
SSW_HD double newly_ionized(const Cell& c, double dt) {  // :312-319
    double absorbed = 1.0 - exp(-n_hi(c) * kSigma * c.len);
    double photons = dt * c.rate;
    return photons * absorbed;
}
SSW_HD double delta_temperature(const Cell& c, double dt) {  // :304-310
    double k = (kGamma - 1.0) * kProtonMass / (c.rho * kBoltzmann);
    double heat = newly_ionized(c, dt) / c.vol * (kPhotonEnergy - kRydberg) / dt;  // photoheating_rate
    double lam = heat - cooling(c);
    double dlam = -cooling_dt(c);
    double m = mu(c);
    return k * m * lam * dt / (1.0 - k * m * dlam * dt);
}
SSW_HD double delta_ionized(const Cell& c, double dt) {  // :334-354 (Rosdahl+15 A23)
    double nh = n_h(c);
    double ne = n_hii(c);
    double al = alpha_b(c.T), dal = alpha_b_dt(c.T);
    double be = beta_ci(c.T), dbe = beta_ci_dt(c.T);
    double gam = newly_ionized(c, dt) / (n_hi(c) * c.vol) / dt;  // photoionization_rate
    double cc = be * ne + gam;
    double m = mu(c);
    double dd = al * ne;
    double x = c.x;
    double rc = ne * c.T * m * 1.0 * dbe;
    double dcdx = nh * be - rc;
    double rd = ne * c.T * m * 1.0 * dal;
    double dddx = nh * al - rd;
    double j = dcdx - (cc + dd) - x * (dcdx + dddx);
    return dt * (cc - x * (cc + dd)) / (1.0 - j * dt);
}

// --- optional outputs of the solver (src/sweep/chemistry_output.rs:25-56) -------
// evaluated with the fixed timestep of :15-19 (1 yr = 3.15576e7 s,
// src/units/mod.rs:48) on the state after the step, without a floor
constexpr double kYear = 3.15576e7;
SSW_HD double photoionization_rate(const Cell& c, double dt) {  // hydrogen_only/mod.rs:327-332
    return newly_ionized(c, dt) / (n_hi(c) * c.vol) / dt;
}
SSW_HD double photoheating_rate(const Cell& c, double dt) {  // :321-325
    return newly_ionized(c, dt) / c.vol * (kPhotonEnergy - kRydberg) / dt;
}
// which: 0 PhotoionizationRate, 1 HeatingRate (photoheating - cooling),
// 2 RecombinationRate (alpha_B n_e x), 3 CollisionalIonizationRate (beta n_e (1 - x))
SSW_HD double chemistry_output(const Cell& c, int which) {
    switch (which) {
        case 0: return photoionization_rate(c, kYear);
        case 1: return photoheating_rate(c, kYear) - cooling(c);
        case 2: return alpha_b(c.T) * n_hii(c) * c.x;
        default: return beta_ci(c.T) * n_hii(c) * (1.0 - c.x);
    }
}

struct Floor {
    bool on;
    double T, x;
};

SSW_HD void clamp_state(Cell& c, const Floor& f) {  // :356-369
    double lo = f.on ? f.x : kXEps;
    double hi = 1.0 - kXEps;
    double v = c.x;
    if (v < lo) v = lo;
    if (v > hi) v = hi;
    c.x = v;
    if (f.on && c.T < f.T) c.T = f.T;
}

// update() of :444-461; false = TimestepCriterionViolated.
SSW_HD bool accept(double& value, double change, double safety, double dt, double& rec) {
    double rel = fmin(fabs(change / value), kInvEps);
    if (rel > safety) return false;
    value += change;
    rec = dt * (safety / rel);
    return true;
}

// try_timestep_update (:371-392)
SSW_HD bool try_step(Cell& c, const Floor& f, double dt, double safety, double& ts, int& proc) {
    double dT = delta_temperature(c, dt);
    double tsT;
    if (!accept(c.T, dT, safety, dt, tsT)) return false;
    double dx = delta_ionized(c, dt);
    double tsx;
    if (!accept(c.x, dx, safety, dt, tsx)) return false;
    clamp_state(c, f);
    if (tsT < tsx) {
        ts = tsT;
        proc = kProcTemperature;
    } else {
        ts = tsx;
        proc = kProcIonization;
    }
    return true;
}

// perform_timestep (:394-441). Returns the timescale; *failed on depth overflow.
SSW_HD double perform_timestep(Cell& c, const Floor& f, double dt0, double safety, int& proc, bool& failed) {
    uint64_t pend_lo = 0, pend_hi = 0;  // pending second halves, bit = depth
    int depth = 0;
    double dt = dt0;
    double ts = 0.0;
    failed = false;
    for (;;) {
        clamp_state(c, f);
        double T0 = c.T, x0 = c.x;
        if (depth > kMaxDepth) {
            failed = true;
            proc = kProcTemperature;
            return dt0 / 10.0;
        }
        if (try_step(c, f, dt, safety, ts, proc)) {
            if (pend_lo == 0 && pend_hi == 0) return ts;
            if (pend_hi) {
                depth = 127 - __builtin_clzll(pend_hi);
                pend_hi &= ~(1ull << (depth - 64));
            } else {
                depth = 63 - __builtin_clzll(pend_lo);
                pend_lo &= ~(1ull << depth);
            }
        } else {
            c.T = T0;
            c.x = x0;
            depth += 1;
            if (depth < 64) pend_lo |= 1ull << depth;
            else pend_hi |= 1ull << (depth - 64);
        }
        dt = dt0;
        for (int k = 0; k < depth; k++) dt = dt / 2.0;
    }
}

}  // namespace ssw
