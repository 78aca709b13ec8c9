// Source rates of star particles (host C++, no device code): the BPASS
// lookup and the star-particle -> source conversion of the reference's Arepo
// post-processing input path (src/arepo_postprocess/bpass.rs:7-56,
// src/arepo_postprocess/sources.rs:57-128, src/cosmology.rs:80-107). The rates
// then go to ssw_assign_sources like any point source. Compiled with
// -ffp-contract=off: the reference's Rust arithmetic is not contracted.

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "../../include/subsweep_amd.h"
#include "bpass_table.h"

namespace {

// linear_interpolate (bpass.rs:7-9)
double lerp(double x1, double x2, double y1, double y2, double x) { return y1 + (x - x1) / (x2 - x1) * (y2 - y1); }

// `bins.binary_search(value).map(|x| x + 1).unwrap_or_else(|e| e)` (bpass.rs:12-16):
// one past an exact match, else the insertion point.
int bin_index(const double* bins, int n, double v) {
    const double* it = std::lower_bound(bins, bins + n, v);
    const int i = (int)(it - bins);
    return (i < n && *it == v) ? i + 1 : i;
}

// Rust's f64::clamp(min, max)
double rust_clamp(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

// the cosmic time of scale factor a in units of 1 / (HUBBLE h), flat
// Lambda-CDM (src/cosmology.rs:96-102)
double cosmic_time(double a, double omega_lambda, double omega_0) {
    const double factor1 = 2.0 / (3.0 * std::sqrt(omega_lambda));
    const double term1 = std::sqrt(omega_lambda / omega_0) * std::pow(a, 1.5);
    const double term2 = std::sqrt(1.0 + omega_lambda / omega_0 * (a * a * a));
    const double factor2 = std::log(term1 + term2);
    return factor1 * factor2;
}

}  // namespace

extern "C" {

double ssw_bpass_source_rate(double age_years, double metallicity, double mass_solar) {
    const double safety = 1.00001;
    const double age = rust_clamp(age_years, kBpassAgeBins[0] * safety, kBpassAgeBins[kBpassAges - 1] / safety);
    const double z = rust_clamp(metallicity, kBpassMetallicityBins[0] * safety,
                                kBpassMetallicityBins[kBpassMetallicities - 1] / safety);
    const int mi = bin_index(kBpassMetallicityBins, kBpassMetallicities, z);
    const int ai = bin_index(kBpassAgeBins, kBpassAges, age);
    const double l1 = lerp(kBpassAgeBins[ai - 1], kBpassAgeBins[ai], kBpassTable[ai - 1][mi - 1],
                           kBpassTable[ai][mi - 1], age);
    const double l2 = lerp(kBpassAgeBins[ai - 1], kBpassAgeBins[ai], kBpassTable[ai - 1][mi], kBpassTable[ai][mi], age);
    const double per_mass = lerp(kBpassMetallicityBins[mi - 1], kBpassMetallicityBins[mi], l1, l2, z);
    return per_mass * mass_solar;
}

double ssw_time_between_scale_factors(double a0, double a1, double omega_lambda, double omega_0, double h) {
    const double hubble = 3.2407789e-18;  // h / s (src/cosmology.rs:86)
    if (std::fabs((omega_lambda + omega_0) - 1.0) > 1e-2) return NAN;  // flat cosmologies only (:92-94)
    const double t0 = cosmic_time(a0, omega_lambda, omega_0);
    const double t1 = cosmic_time(a1, omega_lambda, omega_0);
    return (t1 - t0) / (hubble * h);
}

int64_t ssw_star_sources(int64_t n, const double* metallicity, const double* formation_scale_factor,
                         const double* mass_solar, double escape_fraction, double omega_lambda, double omega_0,
                         double h, double scale_factor, double* rate, uint8_t* is_star) {
    if (n < 0 || (n > 0 && (!metallicity || !formation_scale_factor || !mass_solar || !rate || !is_star)))
        return SSW_ERR_INVALID;
    if (std::fabs((omega_lambda + omega_0) - 1.0) > 1e-2) return SSW_ERR_INVALID;
    const double year = 3.15576e7;  // src/units/mod.rs:48
    int64_t stars = 0;
    for (int64_t i = 0; i < n; i++) {
        // formation scale factors that are not positive mark wind particles (sources.rs:117-118)
        const bool star = formation_scale_factor[i] > 0.0;
        is_star[i] = star ? 1 : 0;
        rate[i] = 0.0;
        if (!star) continue;
        stars++;
        const double age_s =
            ssw_time_between_scale_factors(formation_scale_factor[i], scale_factor, omega_lambda, omega_0, h);
        rate[i] = ssw_bpass_source_rate(age_s / year, metallicity[i], mass_solar[i]) * escape_fraction;
    }
    return stars;
}

}  // extern "C"
