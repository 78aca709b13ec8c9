// Oracle: a single-threaded CPU restatement of subsweep's sweep radiative
// transfer and hydrogen-only chemistry, used ONLY as test infrastructure (what
// the results are compared with in tests/ and __graft_entry__.smoke(), and the
// CPU figure of bench.py's cpu_baseline leg). Nothing in the product links or
// calls it.
//
// It follows the reference's Rust code operation by operation:
//   sweep driver      src/sweep/mod.rs:258-289, 346-398, 412-513, 549-589
//   site / rates      src/sweep/site.rs:12-57, src/chemistry/mod.rs:45-82
//   chemistry         src/chemistry/hydrogen_only/mod.rs:68-461
//   timescale min     src/chemistry/timescale.rs:32-38
//   timestep levels   src/sweep/timestep_level.rs:27-48, src/sweep/timestep_state.rs:4-90
//   active lists      src/sweep/active_list.rs:55-66, 143-153
//   directions        src/sweep/direction/mod.rs:58-109
//   constants, units  src/units/mod.rs:9-110
// Compiled with -O2 -ffp-contract=off (Rust does not contract a*b+c into FMAs);
// exp/pow/sqrt/cbrt/log2 come from the host libm, as Rust std calls them.
//
// Deliberate, documented choices where the reference is order-dependent:
//  * Task order. The reference pops tasks from a BinaryHeap ordered by direction
//    only (src/sweep/task.rs:25-35), so it sweeps direction D-1 first, then
//    D-2, ...; inside one direction the heap's tie order is unspecified. Here each
//    direction is swept in Kahn order (FIFO), directions from D-1 down to 0.
//    Only the summation order of flux corrections depends on this.
//  * Periodic lag. A periodic-source correction reaches its receiver at once in
//    the reference (src/sweep/mod.rs:505-513), so whether a receiver sees it in
//    the same sweep depends on heap order. Here the transport of a cell reads
//    the periodic source as it was when the sweep started; chemistry, which runs
//    after the sweep, reads the current value as in the reference.
//
// Parity status of each part (how it is pinned) is listed in the test files
// tests/test_oracle_*.py.

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include <omp.h>

#include "../include/subsweep_amd.h"

namespace {

// ---- constants and unit factors: src/units/mod.rs:14-108 -------------------
// Unit factors are products of the factors they are defined from, as diman
// computes them.
const double CENTIMETERS = 0.01;
const double CENTIMETERS_SQUARED = CENTIMETERS * CENTIMETERS;  // centimeters * centimeters
const double CUBIC_CENTIMETERS = 1e-6;                         // 1e-6 * cubic_meters
const double SECONDS = 1.0;
const double JOULES = 1.0;
const double ERGS = 1e-7 * JOULES;
const double ELECTRON_VOLTS = 1.602176634e-19 * JOULES;
const double CENTIMETERS_CUBED_PER_S = CUBIC_CENTIMETERS / SECONDS;
const double ERGS_CENTIMETERS_CUBED_PER_S = ERGS * CENTIMETERS_CUBED_PER_S;
const double ERGS_PER_S = ERGS / SECONDS;
const double KELVINS = 1.0;

const double BOLTZMANN_CONSTANT = 1.380649e-23 * JOULES / KELVINS;
const double PROTON_MASS = 1.67262192369e-27;
const double GAMMA = 5.0 / 3.0;
const double NUMBER_WEIGHTED_AVERAGE_CROSS_SECTION = 2.9580524545305314e-18 * CENTIMETERS_SQUARED;
const double PHOTON_AVERAGE_ENERGY = 18.028356312818811 * ELECTRON_VOLTS;
const double RYDBERG_CONSTANT = 13.65693 * ELECTRON_VOLTS;

// src/chemistry/hydrogen_only/mod.rs:32-38
const double HYDROGEN_MASS_FRACTION = 1.0;
const int MAX_DEPTH = 100;
const double IONIZED_HYDROGEN_FRACTION_EPSILON = 1e-10;

// f64::powi as LLVM lowers it (binary exponentiation, compiler-rt __powidf2).
double powi(double a, int b) {
    const bool recip = b < 0;
    double r = 1.0;
    while (true) {
        if (b & 1) r *= a;
        b /= 2;
        if (b == 0) break;
        a *= a;
    }
    return recip ? 1.0 / r : r;
}

// Rust f64::min / f64::clamp semantics for the cases used here.
inline double rust_min(double a, double b) { return std::fmin(a, b); }

// ---- Timescale: src/chemistry/timescale.rs:7-39 ------------------------------
enum Process { TEMPERATURE = 0, IONIZATION_FRACTION = 1, PHOTON_RATE = 2 };
struct Timescale {
    double time;
    int process;
    Timescale min(const Timescale& other) const { return (time < other.time) ? *this : other; }
};

// ---- Solver: src/chemistry/hydrogen_only/mod.rs:125-461 ----------------------
struct Solver {
    double ionized_hydrogen_fraction;
    double temperature;
    double density;
    double volume;
    double length;
    double rate;
    double scale_factor;
    bool has_floor;
    double floor_temperature;
    double floor_xhii;

    double hydrogen_number_density() const { return density / PROTON_MASS; }
    double ionized_hydrogen_number_density() const {
        return hydrogen_number_density() * ionized_hydrogen_fraction;
    }
    double neutral_hydrogen_number_density() const {
        return hydrogen_number_density() * (1.0 - ionized_hydrogen_fraction);
    }
    double electron_number_density() const { return ionized_hydrogen_number_density(); }
    double mu() const { return 1.0 / (ionized_hydrogen_fraction + 1.0); }

    double collision_fit_function() const {  // :161-164
        double t = temperature / KELVINS;
        return std::sqrt(t) / (1.0 + std::sqrt(t / 1e5)) * std::exp(-157809.1 / t);
    }
    double collision_fit_function_derivative() const {  // :166-173
        const double const1 = 1.0 / 1e5;
        const double const2 = 157809.1;
        double t = temperature / KELVINS;
        return (std::exp(-const2 / t) *
                (const1 * const2 * t + 0.5 * std::sqrt(const1 * t) * (2.0 * const2 + t))) /
               (std::sqrt(powi(t, 3)) * std::sqrt(const1 * t) * powi(std::sqrt(const1 * t) + 1.0, 2));
    }
    double case_b_recombination_rate() const {  // :175-180
        double lambda = (315614.0 * KELVINS) / temperature;
        return (2.753e-14 * std::pow(lambda, 1.5) / std::pow(1.0 + std::pow(lambda / 2.74, 0.407), 2.242)) *
               CENTIMETERS_CUBED_PER_S;
    }
    double case_b_recombination_rate_derivative() const {  // :182-194
        double lambda = (315614.0 * KELVINS) / temperature;
        double dlambda_dt = -(315614.0 * KELVINS) / (temperature * temperature);
        const double c1 = 1.0 / 2.74;
        const double c2 = 0.407;
        const double c3 = 2.242;
        double d = -std::sqrt(lambda) * std::pow(std::pow(c1 * lambda, c2) + 1.0, -c3 - 1.0) *
                   (c2 * c3 * std::pow(c1 * lambda, c2) - 1.5 * std::pow(c1 * lambda, c2) - 1.5);
        return ((2.753e-14 * d) * CENTIMETERS_CUBED_PER_S) * dlambda_dt;
    }
    double case_b_recombination_cooling_rate() const {  // :196-202
        double lambda = (315614.0 * KELVINS) / temperature;
        return (3.435e-30 * (temperature / KELVINS) * std::pow(lambda, 1.97) /
                std::pow(1.0 + std::pow(lambda / 2.25, 0.376), 3.72)) *
               ERGS_CENTIMETERS_CUBED_PER_S;
    }
    double case_b_recombination_cooling_rate_derivative() const {  // :204-216
        const double c1 = 315614.0, c2 = 1.97, c3 = 0.376, c4 = 3.72, c5 = 2.25;
        double t = temperature / KELVINS;
        double derivative = std::pow(1.0 + std::pow(c1 / (c5 * t), c3), -1.0 - c4) *
                            (1.0 - 1.0 * c2 + (1.0 - 1.0 * c2 + c3 * c4) * std::pow(c1 / (c5 * t), c3)) *
                            std::pow(c1 / t, c2);
        return ((3.435e-30 * derivative) * ERGS_CENTIMETERS_CUBED_PER_S) / (1.0 * KELVINS);
    }
    double collisional_ionization_rate() const {  // :218-220
        return (5.85e-11 * collision_fit_function()) * CENTIMETERS_CUBED_PER_S;
    }
    double collisional_ionization_rate_derivative() const {  // :222-225
        return ((5.85e-11 * collision_fit_function_derivative()) * CENTIMETERS_CUBED_PER_S) / (1.0 * KELVINS);
    }
    double collisional_ionization_cooling_rate() const {  // :227-229
        return (1.27e-21 * collision_fit_function()) * ERGS_CENTIMETERS_CUBED_PER_S;
    }
    double collisional_ionization_cooling_rate_derivative() const {  // :231-235
        return ((1.27e-21 * collision_fit_function_derivative()) * ERGS_CENTIMETERS_CUBED_PER_S) / (1.0 * KELVINS);
    }
    double collisional_excitation_cooling_rate() const {  // :237-242
        double t = temperature / KELVINS;
        return (7.5e-19 / (1.0 + std::sqrt(t / 1e5)) * std::exp(-118348.0 / t)) * ERGS_CENTIMETERS_CUBED_PER_S;
    }
    double collisional_excitation_cooling_rate_derivative() const {  // :244-253
        double t = temperature / KELVINS;
        const double c1 = 7.5e-19, c2 = 118348.0, c3 = 1.0 / 1e5;
        double v = (c1 * std::exp(-c2 / t) * (c2 * c3 * t - 0.5 * c3 * powi(t, 2) + c2 * std::sqrt(c3 * t))) /
                   (powi(t, 2) * std::sqrt(c3 * t) * powi(1.0 + std::sqrt(c3 * t), 2));
        return (v * ERGS_CENTIMETERS_CUBED_PER_S) / (1.0 * KELVINS);
    }
    double bremsstrahlung_cooling_rate() const {  // :255-257
        return (1.42e-27 * std::sqrt(temperature / KELVINS)) * ERGS_CENTIMETERS_CUBED_PER_S;
    }
    double bremsstrahlung_cooling_rate_derivative() const {  // :259-263
        return ((1.42e-27 / (2.0 * std::sqrt(temperature / KELVINS))) * ERGS_CENTIMETERS_CUBED_PER_S) /
               (1.0 * KELVINS);
    }
    double compton_cooling_rate() const {  // :265-268
        double x = 2.727 / scale_factor;
        return (1.017e-37 * powi(x, 4) * (temperature / KELVINS - x)) * ERGS_PER_S;
    }
    double compton_cooling_rate_derivative() const {  // :270-273
        double x = 2.727 / scale_factor;
        return ((1.017e-37 * powi(x, 4)) * ERGS_PER_S) / (1.0 * KELVINS);
    }
    double cooling_rate() const {  // :275-287
        double ne = electron_number_density();
        double nh_neutral = neutral_hydrogen_number_density();
        double nh_ionized = ionized_hydrogen_number_density();
        double collisional =
            (collisional_excitation_cooling_rate() + collisional_ionization_cooling_rate()) * ne * nh_neutral;
        double recombination = case_b_recombination_cooling_rate() * ne * nh_ionized;
        double bremsstrahlung = bremsstrahlung_cooling_rate() * ne * nh_ionized;
        double compton = compton_cooling_rate() * ne;
        return collisional + recombination + bremsstrahlung + compton;
    }
    double cooling_rate_derivative() const {  // :289-302
        double ne = electron_number_density();
        double nh_neutral = neutral_hydrogen_number_density();
        double nh_ionized = ionized_hydrogen_number_density();
        double collisional = (collisional_excitation_cooling_rate_derivative() +
                              collisional_ionization_cooling_rate_derivative()) *
                             ne * nh_neutral;
        double recombination = case_b_recombination_cooling_rate_derivative() * ne * nh_ionized;
        double bremsstrahlung = bremsstrahlung_cooling_rate_derivative() * ne * nh_ionized;
        double compton = compton_cooling_rate_derivative() * ne;
        return collisional + recombination + bremsstrahlung + compton;
    }
    double temperature_change(double timestep) const {  // :304-310
        double k = (GAMMA - 1.0) * PROTON_MASS / (density * BOLTZMANN_CONSTANT);
        double lambda = photoheating_rate(timestep) - cooling_rate();
        double dlambdadt = -cooling_rate_derivative();
        double m = mu();
        return k * m * lambda * timestep / (1.0 - k * m * dlambdadt * timestep);
    }
    double num_newly_ionized_hydrogen_atoms(double timestep) const {  // :312-319
        double n = neutral_hydrogen_number_density();
        double sigma = NUMBER_WEIGHTED_AVERAGE_CROSS_SECTION;
        double absorbed_fraction = 1.0 - std::exp(-n * sigma * length);
        double num_photons = timestep * rate;
        return num_photons * absorbed_fraction;
    }
    double photoheating_rate(double timestep) const {  // :321-325
        double num = num_newly_ionized_hydrogen_atoms(timestep);
        double ionization_density = num / volume;
        return ionization_density * (PHOTON_AVERAGE_ENERGY - RYDBERG_CONSTANT) / timestep;
    }
    double photoionization_rate(double timestep) const {  // :327-332
        double num = num_newly_ionized_hydrogen_atoms(timestep);
        double fraction = num / (neutral_hydrogen_number_density() * volume);
        return fraction / timestep;
    }
    double ionized_fraction_change(double timestep) const {  // :334-354
        double nh = hydrogen_number_density();
        double ne = electron_number_density();
        double alpha = case_b_recombination_rate();
        double dalpha = case_b_recombination_rate_derivative();
        double beta = collisional_ionization_rate();
        double dbeta = collisional_ionization_rate_derivative();
        double gamma_ph = photoionization_rate(timestep);
        double c = beta * ne + gamma_ph;
        double m = mu();
        double d = alpha * ne;
        double xhii = ionized_hydrogen_fraction;
        double rhsc = ne * temperature * m * HYDROGEN_MASS_FRACTION * dbeta;
        double dcdx = nh * beta - rhsc;
        double rhsd = ne * temperature * m * HYDROGEN_MASS_FRACTION * dalpha;
        double dddx = nh * alpha - rhsd;
        double j = dcdx - (c + d) - xhii * (dcdx + dddx);
        return timestep * (c - xhii * (c + d)) / (1.0 - j * timestep);
    }
    void clamp() {  // :356-369
        double floor_x = has_floor ? floor_xhii : IONIZED_HYDROGEN_FRACTION_EPSILON;
        double hi = 1.0 - IONIZED_HYDROGEN_FRACTION_EPSILON;
        // f64::clamp(min, max): max(min) then min(max)
        double v = ionized_hydrogen_fraction;
        if (v < floor_x) v = floor_x;
        if (v > hi) v = hi;
        ionized_hydrogen_fraction = v;
        if (has_floor && temperature < floor_temperature) temperature = floor_temperature;
    }
    // update(): :444-461. Returns false on TimestepCriterionViolated.
    static bool update(double& value, double change, double max_allowed_change, double timestep,
                       double& recommendation) {
        double relative_change = rust_min(std::fabs(change / value), 1.0 / DBL_EPSILON);
        if (relative_change > max_allowed_change) return false;
        value += change;
        recommendation = timestep * (max_allowed_change / relative_change);
        return true;
    }
    bool try_timestep_update(double timestep, double safety, Timescale& out) {  // :371-392
        double temperature_change_v = temperature_change(timestep);
        double ts_t;
        if (!update(temperature, temperature_change_v, safety, timestep, ts_t)) return false;
        double x_change = ionized_fraction_change(timestep);
        double ts_x;
        if (!update(ionized_hydrogen_fraction, x_change, safety, timestep, ts_x)) return false;
        clamp();
        out = Timescale{ts_t, TEMPERATURE}.min(Timescale{ts_x, IONIZATION_FRACTION});
        return true;
    }
    // perform_timestep_internal(): :394-424. Returns false on convergence failure.
    bool perform_timestep_internal(double timestep, double safety, int depth, Timescale& out) {
        clamp();
        double init_t = temperature, init_x = ionized_hydrogen_fraction;
        if (depth > MAX_DEPTH) return false;
        Timescale ts;
        if (try_timestep_update(timestep, safety, ts)) {
            out = ts;
            return true;
        }
        temperature = init_t;
        ionized_hydrogen_fraction = init_x;
        Timescale first;
        if (!perform_timestep_internal(timestep / 2.0, safety, depth + 1, first)) return false;
        return perform_timestep_internal(timestep / 2.0, safety, depth + 1, out);
    }
    Timescale perform_timestep(double timestep, double safety, bool* failed) {  // :426-441
        Timescale out;
        if (perform_timestep_internal(timestep, safety, 0, out)) {
            if (failed) *failed = false;
            return out;
        }
        if (failed) *failed = true;
        return Timescale{timestep / 10.0, TEMPERATURE};
    }
};

// ---- TimestepLevel / TimestepState ------------------------------------------
// src/sweep/timestep_level.rs:27-36; `as usize` saturates (NaN and negatives -> 0).
int level_from_max_and_desired(int max_num_levels, double max_timestep, double desired) {
    double ratio = max_timestep / desired;
    double v = std::ceil(std::log2(ratio));
    long long level;
    if (!(v > 0.0)) level = 0;
    else if (v >= (double)(max_num_levels - 1)) level = max_num_levels - 1;
    else level = (long long)v;
    if (level > max_num_levels - 1) level = max_num_levels - 1;
    return (int)level;
}
double level_factor(int level) { return powi(0.5, level); }  // :46-48

struct TimestepState {  // src/sweep/timestep_state.rs:4-80
    double max_timestep;
    int max_num_levels;
    int current_lowest_allowed;
    bool first_done;
    void init(double mt, int n) {
        max_timestep = mt;
        max_num_levels = n;
        current_lowest_allowed = n - 1;
        first_done = false;
    }
    int num_allowed() const { return max_num_levels - current_lowest_allowed; }
    std::vector<int> levels_in_sweep_order() const {  // :22-27, 70-75, 82-90
        int num = num_allowed();
        std::vector<int> out;
        for (uint32_t i = 0; i < (1u << (num - 1)); i++) {
            int first_bit = num - 1;
            for (int b = 0; b < 32; b++)
                if (i & (1u << b)) {
                    first_bit = b;
                    break;
                }
            out.push_back(current_lowest_allowed + (num - 1 - first_bit));
        }
        return out;
    }
    void advance() {  // :37-48
        if (first_done && current_lowest_allowed > 0) current_lowest_allowed -= 1;
        if (!first_done) first_done = true;
    }
    double timestep_at_level(int level) const { return max_timestep * level_factor(level); }
    int desired_level(double desired) const {  // :54-64
        int level = level_from_max_and_desired(max_num_levels, max_timestep, desired);
        if (level < current_lowest_allowed) level = current_lowest_allowed;
        return level;
    }
    double current_max_timestep() const { return max_timestep * level_factor(current_lowest_allowed); }
};

struct Oracle {
    ssw_sweep_parameters params;
    int D = 0;
    std::vector<double> dirs;  // D x 3
    // grid (CSR)
    int64_t N = 0;
    std::vector<int64_t> off;
    std::vector<int32_t> nbr;
    std::vector<uint8_t> kind;
    std::vector<double> area, normal, volume, size;
    bool have_grid = false, have_state = false;
    // state, Site (src/sweep/site.rs:12-23); per-direction vectors stored [c*D + d]
    std::vector<double> density, x, T, source, species_timestep, prev_rate, change_ts;
    std::vector<int> level;
    std::vector<double> I, O, P, P_snap;
    std::vector<uint32_t> missing;  // num_missing_upwind [c*D + d]
    TimestepState ts;
    int current_level = 0;
    int64_t chemistry_failures = 0;
    int64_t updates_last_step = 0;
    // TimescaleCounter (src/chemistry/timescale.rs:70-111) of the last
    // run_sweeps call: cells per limiting process (temperature, ionisation
    // fraction, photon rate) whose change timescale is below max_timestep
    int64_t limiting[3] = {0, 0, 0};
    // Threads of the sweep (one direction per thread: directions never share
    // state inside a sweep) and of the per-cell loops. Each direction's and
    // each cell's arithmetic is the same as with one thread, so every result
    // is bitwise independent of the count.
    int threads = 1;

    double dot(int64_t f, int d) const {  // glam DVec3::dot: (x*x') + (y*y') + (z*z')
        return (normal[3 * f] * dirs[3 * d] + normal[3 * f + 1] * dirs[3 * d + 1]) + normal[3 * f + 2] * dirs[3 * d + 2];
    }
    int face_kind(int64_t f) const {
        // map_ptype (src/voronoi/constructor/mod.rs:125-135): periodic faces are
        // boundaries unless the sweep is periodic.
        int k = kind[f];
        if (k == SSW_FACE_PERIODIC && !params.periodic) return SSW_FACE_BOUNDARY;
        return k;
    }
    bool is_active(int64_t c, int lvl) const { return level[c] >= lvl; }  // timestep_level.rs:38-40

    // ActiveList::enumerate_active: bins from current level upwards, each bin in
    // index order (src/sweep/active_list.rs:55-66, 143-153).
    std::vector<int64_t> enumerate_active(int lvl) const {
        std::vector<int64_t> out;
        for (int b = lvl; b < params.num_timestep_levels; b++)
            for (int64_t c = 0; c < N; c++)
                if (level[c] == b) out.push_back(c);
        return out;
    }

    // Sweep::init_counts (src/sweep/mod.rs:346-386), single rank.
    void init_counts(int lvl, const std::vector<int64_t>& active) {
#pragma omp parallel for num_threads(threads) schedule(static)
        for (size_t i = 0; i < active.size(); i++) {
            const int64_t c = active[i];
            for (int d = 0; d < D; d++) {
                uint32_t n = 0;
                for (int64_t f = off[c]; f < off[c + 1]; f++) {
                    int k = face_kind(f);
                    if (!(dot(f, d) < 0.0) || k == SSW_FACE_BOUNDARY) continue;
                    if (!is_active(nbr[f], lvl)) continue;
                    if (k == SSW_FACE_LOCAL) n++;
                }
                missing[c * D + d] = n;
            }
        }
    }

    double source_per_dir(int64_t c) const { return source[c] / (double)D; }  // site.rs:49-51

    // HydrogenOnly::get_outgoing_rate (src/chemistry/hydrogen_only/mod.rs:72-88)
    double outgoing_rate(int64_t c, double incoming) const {
        double n = density[c] / PROTON_MASS * (1.0 - x[c]);
        double sigma = NUMBER_WEIGHTED_AVERAGE_CROSS_SECTION;
        if (incoming < params.significant_rate_threshold) return 0.0;
        double non_absorbed = std::exp(-n * sigma * size[c]);
        return incoming * non_absorbed;
    }

    // Sweep::solve_task (src/sweep/mod.rs:412-485) + handle_local[_periodic]_neighbour (:487-513)
    void solve_task(int64_t c, int d, int lvl, std::deque<int64_t>& ready, std::vector<uint32_t>* depth) {
        double& Icd = I[c * D + d];
        if (Icd < 0.0) Icd = 0.0;  // Photons::make_positive (src/chemistry/mod.rs:61-65)
        double incoming = (Icd + source_per_dir(c)) + P_snap[c * D + d];  // Site::get_rate
        double out = outgoing_rate(c, incoming);
        double correction = out - O[c * D + d];
        O[c * D + d] = out;
        double total = 0.0;
        for (int64_t f = off[c]; f < off[c + 1]; f++) {
            double dd = dot(f, d);
            if (dd > 0.0) total += area[f] * dd;
        }
        for (int64_t f = off[c]; f < off[c + 1]; f++) {
            double dd = dot(f, d);
            if (!(dd > 0.0)) continue;
            double effective = area[f] * dd;
            double delta = correction * (effective / total);
            int k = face_kind(f);
            int64_t nb = nbr[f];
            if (k == SSW_FACE_LOCAL) {
                I[nb * D + d] += delta;
                if (is_active(nb, lvl)) {
                    uint32_t& m = missing[nb * D + d];
                    m -= 1;
                    if (depth) {
                        uint32_t cand = (*depth)[d * N + c] + 1;
                        if (cand > (*depth)[d * N + nb] || (*depth)[d * N + nb] == 0xFFFFFFFFu)
                            (*depth)[d * N + nb] = cand;
                    }
                    if (m == 0) ready.push_back(nb);
                }
            } else if (k == SSW_FACE_PERIODIC) {
                P[nb * D + d] += delta;
            }
        }
    }

    // Kahn layering of one direction restricted to the active set (no transport):
    // the level-set order. depth[d*N + c].
    void levelsets(int lvl, std::vector<uint32_t>& depth) {
        std::vector<int64_t> active = enumerate_active(lvl);
        init_counts(lvl, active);
        depth.assign((size_t)D * N, 0xFFFFFFFFu);
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
        for (int d = 0; d < D; d++) {
            std::deque<int64_t> q;
            for (int64_t c : active)
                if (missing[c * D + d] == 0) {
                    q.push_back(c);
                    depth[d * N + c] = 0;
                }
            while (!q.empty()) {
                int64_t c = q.front();
                q.pop_front();
                for (int64_t f = off[c]; f < off[c + 1]; f++) {
                    if (!(dot(f, d) > 0.0) || face_kind(f) != SSW_FACE_LOCAL) continue;
                    int64_t nb = nbr[f];
                    if (!is_active(nb, lvl)) continue;
                    uint32_t cand = depth[d * N + c] + 1;
                    if (depth[d * N + nb] == 0xFFFFFFFFu || cand > depth[d * N + nb]) depth[d * N + nb] = cand;
                    if (--missing[nb * D + d] == 0) q.push_back(nb);
                }
            }
        }
    }

    // Sweep::single_sweep (src/sweep/mod.rs:274-289) for the current level.
    bool single_sweep(int lvl) {
        std::vector<int64_t> active = enumerate_active(lvl);
        init_counts(lvl, active);
        P_snap = P;  // periodic source as it is when the sweep starts
        int64_t solved = 0;
        // BinaryHeap pops the largest direction first; directions share no
        // state inside a sweep, so with several threads each runs on its own
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1) reduction(+ : solved)
        for (int d = D - 1; d >= 0; d--) {
            std::deque<int64_t> ready;
            for (int64_t c : active)  // get_initial_tasks (src/sweep/mod.rs:388-398)
                if (missing[c * D + d] == 0) ready.push_back(c);
            while (!ready.empty()) {
                int64_t c = ready.front();
                ready.pop_front();
                solve_task(c, d, lvl, ready, nullptr);
                solved++;
            }
        }
        updates_last_step += solved;
        if (solved != (int64_t)active.size() * D) return false;  // dependency cycle
        update_chemistry(lvl, active);
        return true;
    }

    // Sweep::update_chemistry (src/sweep/mod.rs:549-574)
    void update_chemistry(int lvl, const std::vector<int64_t>& active) {
        (void)lvl;
        int64_t failures = 0, lim0 = 0, lim1 = 0, lim2 = 0;
#pragma omp parallel for num_threads(threads) schedule(static) reduction(+ : failures, lim0, lim1, lim2)
        for (size_t i = 0; i < active.size(); i++) {
            const int64_t c = active[i];
            double timestep = ts.timestep_at_level(level[c]);
            double rate = 0.0;
            for (int d = 0; d < D; d++) rate += (I[c * D + d] + source_per_dir(c)) + P[c * D + d];
            double relative_change;
            if (std::fabs(rate) < std::fabs(params.significant_rate_threshold)) {  // below_threshold
                relative_change = 0.0;
            } else {  // relative_change_to(..).abs()
                relative_change =
                    std::fabs(rust_min(std::fabs(std::fabs(rate - prev_rate[c]) / rate), 1.0 / DBL_EPSILON));
            }
            prev_rate[c] = rate;
            Timescale rate_ts{timestep / relative_change, PHOTON_RATE};
            // HydrogenOnly::update_abundances (src/chemistry/hydrogen_only/mod.rs:90-120)
            Solver s;
            s.ionized_hydrogen_fraction = x[c];
            s.temperature = T[c];
            s.density = density[c];
            s.volume = volume[c];
            s.length = size[c];
            s.rate = rate;
            s.scale_factor = params.scale_factor;
            s.has_floor = params.prevent_cooling != 0;
            s.floor_temperature = T[c];
            s.floor_xhii = x[c];
            bool failed = false;
            Timescale used = s.perform_timestep(timestep, params.chemistry_timestep_safety_factor, &failed);
            if (failed) failures++;
            T[c] = s.temperature;
            x[c] = s.ionized_hydrogen_fraction;
            species_timestep[c] = used.time;
            Timescale change = rate_ts.min(used);
            change_ts[c] = change.time;
            if (change.time < params.max_timestep) {  // TimescaleCounter::count (timescale.rs:80-87)
                lim0 += change.process == TEMPERATURE;
                lim1 += change.process == IONIZATION_FRACTION;
                lim2 += change.process == PHOTON_RATE;
            }
        }
        chemistry_failures += failures;
        limiting[0] += lim0;
        limiting[1] += lim1;
        limiting[2] += lim2;
    }

    // Sweep::update_timestep_levels (src/sweep/mod.rs:576-589)
    void update_timestep_levels() {
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int64_t c = 0; c < N; c++) {
            double desired = params.timestep_safety_factor * change_ts[c];
            level[c] = ts.desired_level(desired);
        }
    }

    // Sweep::run_sweeps (src/sweep/mod.rs:258-272)
    int run_sweeps(double* elapsed) {
        updates_last_step = 0;
        limiting[0] = limiting[1] = limiting[2] = 0;  // show_timestep_limiting_processes resets it (:267)
        std::vector<int> order = ts.levels_in_sweep_order();
        std::vector<int64_t> counts(params.num_timestep_levels, 0);
        for (int l = 0; l < params.num_timestep_levels; l++)
            for (int64_t c = 0; c < N; c++)
                if (level[c] >= l) counts[l]++;
        for (int l : order) {
            if (counts[l] > 0) {
                current_level = l;
                if (!single_sweep(l)) return SSW_ERR_CYCLE;
            }
        }
        double time_elapsed = ts.current_max_timestep();
        ts.advance();
        update_timestep_levels();
        if (elapsed) *elapsed = time_elapsed;
        return SSW_OK;
    }
};

thread_local std::string g_err;
int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

}  // namespace

extern "C" {

const char* orc_last_error(void) { return g_err.c_str(); }

// `tabulated`: for DirectionsSpecification::Num (no explicit directions), the
// num_directions x 3 bins of Directions::from_num (src/sweep/direction/mod.rs:58-75),
// used as tabulated; the caller reads them from the reference's tables.
void* orc_create(const ssw_sweep_parameters* p, const double* tabulated) {
    if (!p || p->num_timestep_levels < 1 || p->num_timestep_levels >= 32 || p->num_directions < 1) return nullptr;
    Oracle* o = new Oracle();
    o->params = *p;
    o->D = p->num_directions;
    if (p->explicit_directions) {
        // Direction(dir.normalize()): glam DVec3::normalize = v * (1 / sqrt(dot(v, v)))
        for (int d = 0; d < o->D; d++) {
            const double* v = p->explicit_directions + 3 * d;
            double len = std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
            double r = 1.0 / len;
            o->dirs.push_back(v[0] * r);
            o->dirs.push_back(v[1] * r);
            o->dirs.push_back(v[2] * r);
        }
    } else {
        if (!tabulated) {
            delete o;
            return nullptr;
        }
        o->dirs.assign(tabulated, tabulated + 3 * o->D);
    }
    o->ts.init(p->max_timestep, p->num_timestep_levels);
    return o;
}

void orc_destroy(void* h) { delete static_cast<Oracle*>(h); }

void orc_set_threads(void* h, int n) { static_cast<Oracle*>(h)->threads = n < 1 ? 1 : n; }

// Replaces the per-cell source rates of a running problem, keeping every
// other state. The reference sets sources once, at startup
// (src/source_systems.rs:48-80, 96-100); the oracle's tests use this to
// change the summed rate between steps and pin update_chemistry's
// rate timescale (src/sweep/mod.rs:554-569).
int orc_set_source(void* h, const double* source) {
    Oracle* o = static_cast<Oracle*>(h);
    if (!o || !o->have_state) return fail(SSW_ERR_INVALID, "state first");
    o->source.assign(source, source + o->N);
    return SSW_OK;
}

// Cells per limiting process of the last step's TimescaleCounter:
// out = {temperature, ionization fraction, photon rate}.
void orc_limiting_processes(void* h, int64_t* out) {
    Oracle* o = static_cast<Oracle*>(h);
    for (int k = 0; k < 3; k++) out[k] = o->limiting[k];
}

int orc_upload_grid(void* h, int64_t n, const int64_t* off, const int32_t* nbr, const uint8_t* kind,
                    const double* area, const double* normal, const double* volume, const double* size) {
    Oracle* o = static_cast<Oracle*>(h);
    if (!o || n <= 0) return fail(SSW_ERR_INVALID, "bad grid");
    int64_t F = off[n];
    o->N = n;
    o->off.assign(off, off + n + 1);
    o->nbr.assign(nbr, nbr + F);
    o->kind.assign(kind, kind + F);
    o->area.assign(area, area + F);
    o->normal.assign(normal, normal + 3 * F);
    o->volume.assign(volume, volume + n);
    o->size.assign(size, size + n);
    o->have_grid = true;
    return SSW_OK;
}

int orc_upload_state(void* h, const double* density, const double* xhii, const double* temp, const double* source) {
    Oracle* o = static_cast<Oracle*>(h);
    if (!o || !o->have_grid) return fail(SSW_ERR_INVALID, "grid first");
    int64_t N = o->N;
    int D = o->D;
    o->density.assign(density, density + N);
    o->x.assign(xhii, xhii + N);
    o->T.assign(temp, temp + N);
    o->source.assign(source, source + N);
    o->species_timestep.assign(N, 0.0);
    o->prev_rate.assign(N, 0.0);
    o->change_ts.assign(N, 0.0);
    o->level.assign(N, o->params.num_timestep_levels - 1);  // src/sweep/mod.rs:206
    o->I.assign((size_t)N * D, 0.0);
    o->O.assign((size_t)N * D, 0.0);
    o->P.assign((size_t)N * D, 0.0);
    o->P_snap.assign((size_t)N * D, 0.0);
    o->missing.assign((size_t)N * D, 0);
    o->ts.init(o->params.max_timestep, o->params.num_timestep_levels);
    o->chemistry_failures = 0;
    o->have_state = true;
    return SSW_OK;
}

int orc_step(void* h, double* elapsed) {
    Oracle* o = static_cast<Oracle*>(h);
    if (!o || !o->have_state) return fail(SSW_ERR_INVALID, "state first");
    int rc = o->run_sweeps(elapsed);
    if (rc != SSW_OK) return fail(rc, "dependency cycle");
    return SSW_OK;
}

int orc_download(void* h, double* xhii, double* temp, double* species_ts, double* photon_rate, uint8_t* level,
                 double* change_ts) {
    Oracle* o = static_cast<Oracle*>(h);
    if (!o || !o->have_state) return fail(SSW_ERR_INVALID, "state first");
    for (int64_t c = 0; c < o->N; c++) {
        if (xhii) xhii[c] = o->x[c];
        if (temp) temp[c] = o->T[c];
        if (species_ts) species_ts[c] = o->species_timestep[c];
        if (photon_rate) {
            double s = 0.0;
            for (int d = 0; d < o->D; d++) s += o->I[c * o->D + d];
            photon_rate[c] = s;
        }
        if (level) level[c] = (uint8_t)o->level[c];
        if (change_ts) change_ts[c] = o->change_ts[c];
    }
    return SSW_OK;
}

int orc_download_counts(void* h, int32_t lvl, uint32_t* counts) {
    Oracle* o = static_cast<Oracle*>(h);
    if (!o || !o->have_state) return fail(SSW_ERR_INVALID, "state first");
    std::vector<int64_t> active = o->enumerate_active(lvl);
    o->init_counts(lvl, active);
    for (size_t i = 0; i < (size_t)o->N * o->D; i++) counts[i] = 0xFFFFFFFFu;
    for (int64_t c : active)
        for (int d = 0; d < o->D; d++) counts[d * o->N + c] = o->missing[c * o->D + d];
    return SSW_OK;
}

int orc_download_levelsets(void* h, int32_t lvl, uint32_t* depth) {
    Oracle* o = static_cast<Oracle*>(h);
    if (!o || !o->have_state) return fail(SSW_ERR_INVALID, "state first");
    std::vector<uint32_t> dep;
    o->levelsets(lvl, dep);
    std::memcpy(depth, dep.data(), dep.size() * sizeof(uint32_t));
    return SSW_OK;
}

int orc_download_fluxes(void* h, double* incoming, double* outgoing, double* periodic) {
    Oracle* o = static_cast<Oracle*>(h);
    if (!o || !o->have_state) return fail(SSW_ERR_INVALID, "state first");
    int D = o->D;
    int64_t N = o->N;
    for (int64_t c = 0; c < N; c++)
        for (int d = 0; d < D; d++) {
            if (incoming) incoming[d * N + c] = o->I[c * D + d];
            if (outgoing) outgoing[d * N + c] = o->O[c * D + d];
            if (periodic) periodic[d * N + c] = o->P[c * D + d];
        }
    return SSW_OK;
}

// The optional chemistry outputs (src/sweep/chemistry_output.rs:25-56) of one
// cell from the Solver of Sweep::get_solver (src/sweep/mod.rs:612-632: no
// floor), timestep 1 yr (chemistry_output.rs:15-19; src/units/mod.rs:48).
// out[0..3]: photoionization, heating (photoheating - cooling), recombination
// (alpha_B n_e x), collisional ionization (beta n_e (1 - x)).
void orc_chemistry_outputs_cell(double xhii, double temperature, double density, double volume, double length,
                                double rate, double scale_factor, double* out) {
    Solver s{};
    s.ionized_hydrogen_fraction = xhii;
    s.temperature = temperature;
    s.density = density;
    s.volume = volume;
    s.length = length;
    s.rate = rate;
    s.scale_factor = scale_factor;
    s.has_floor = false;
    const double timestep = 3.15576e7;
    out[0] = s.photoionization_rate(timestep);
    out[1] = s.photoheating_rate(timestep) - s.cooling_rate();
    out[2] = s.case_b_recombination_rate() * s.electron_number_density() * s.ionized_hydrogen_fraction;
    out[3] = s.collisional_ionization_rate() * s.electron_number_density() * (1.0 - s.ionized_hydrogen_fraction);
}

// The same for every cell of the oracle's current state, rate = sum over
// directions of Site::get_rate (src/sweep/site.rs:53-56) in direction order.
int orc_download_chemistry_outputs(void* h, double* photoionization, double* heating, double* recombination,
                                   double* collisional_ionization) {
    Oracle* o = static_cast<Oracle*>(h);
    if (!o || !o->have_state) return fail(SSW_ERR_INVALID, "state first");
    double* outs[4] = {photoionization, heating, recombination, collisional_ionization};
    for (int64_t c = 0; c < o->N; c++) {
        double rate = 0.0;
        for (int d = 0; d < o->D; d++) rate += (o->I[c * o->D + d] + o->source_per_dir(c)) + o->P[c * o->D + d];
        double v[4];
        orc_chemistry_outputs_cell(o->x[c], o->T[c], o->density[c], o->volume[c], o->size[c], rate,
                                   o->params.scale_factor, v);
        for (int k = 0; k < 4; k++)
            if (outs[k]) outs[k][c] = v[k];
    }
    return SSW_OK;
}

// rotate_directions_system (src/sweep/direction/mod.rs:158-175) with the
// rotation matrix given (row-major 3 x 3; the reference draws it from its
// StdRng, get_random_rotation_matrix :137-148): every direction becomes
// matrix . direction (multiply_by_matrix :150-155), then incoming, outgoing
// and periodic rates of every site are remapped (remap :190-205). remap's
// inner loop assigns values[i] = old[j] * kernel[i][j] for every j, so only
// j = D-1 survives: values[i] = old[D-1] * kernel[i][D-1], where kernel[i][j]
// is 1 if new direction j is the last of the new directions most aligned
// with old direction i (max_by_key keeps the last maximum of 1 + old_i . new_j,
// kernel_f :177-188) and 0 otherwise.
int orc_rotate_directions(void* h, const double* m) {
    Oracle* o = static_cast<Oracle*>(h);
    if (!o || !m) return fail(SSW_ERR_INVALID, "null argument");
    const int D = o->D;
    std::vector<double> old_dirs = o->dirs;
    for (int d = 0; d < D; d++) {
        const double x = o->dirs[3 * d], y = o->dirs[3 * d + 1], z = o->dirs[3 * d + 2];
        o->dirs[3 * d] = x * m[0] + y * m[1] + z * m[2];
        o->dirs[3 * d + 1] = x * m[3] + y * m[4] + z * m[5];
        o->dirs[3 * d + 2] = x * m[6] + y * m[7] + z * m[8];
    }
    std::vector<double> kept(D);  // kernel[i][D-1]
    for (int i = 0; i < D; i++) {
        int best = 0;
        double best_v = 0.0;
        for (int j = 0; j < D; j++) {
            const double v = 1.0 + ((old_dirs[3 * i] * o->dirs[3 * j] + old_dirs[3 * i + 1] * o->dirs[3 * j + 1]) +
                                    old_dirs[3 * i + 2] * o->dirs[3 * j + 2]);
            if (j == 0 || v >= best_v) {  // Iterator::max_by_key: the last maximum
                best = j;
                best_v = v;
            }
        }
        kept[i] = best == D - 1 ? 1.0 : 0.0;
    }
    if (o->have_state) {
        for (std::vector<double>* arr : {&o->I, &o->O, &o->P}) {
            std::vector<double>& a = *arr;
            for (int64_t c = 0; c < o->N; c++) {
                const double last = a[c * D + D - 1];
                for (int i = 0; i < D; i++) a[c * D + i] = last * kept[i];
            }
        }
    }
    return SSW_OK;
}

int64_t orc_updates_last_step(void* h) { return static_cast<Oracle*>(h)->updates_last_step; }
int64_t orc_chemistry_failures(void* h) { return static_cast<Oracle*>(h)->chemistry_failures; }
int orc_num_directions(void* h) { return static_cast<Oracle*>(h)->D; }
void orc_directions(void* h, double* out) {
    Oracle* o = static_cast<Oracle*>(h);
    std::memcpy(out, o->dirs.data(), o->dirs.size() * sizeof(double));
}

// ---- unit-level entry points used by the known-answer tests -------------------
// Evaluates every rate fit and derivative of Solver at (T, x, density, ...).
// out[0..15]: alpha_B, dalpha_B, beta, dbeta, recomb_cool, drecomb_cool,
// coll_ion_cool, dcoll_ion_cool, coll_exc_cool, dcoll_exc_cool, brems, dbrems,
// compton, dcompton, cooling_rate, cooling_rate_derivative.
void orc_solver_rates(double temperature, double xhii, double density, double scale_factor, double* out) {
    Solver s{};
    s.temperature = temperature;
    s.ionized_hydrogen_fraction = xhii;
    s.density = density;
    s.scale_factor = scale_factor;
    s.has_floor = false;
    out[0] = s.case_b_recombination_rate();
    out[1] = s.case_b_recombination_rate_derivative();
    out[2] = s.collisional_ionization_rate();
    out[3] = s.collisional_ionization_rate_derivative();
    out[4] = s.case_b_recombination_cooling_rate();
    out[5] = s.case_b_recombination_cooling_rate_derivative();
    out[6] = s.collisional_ionization_cooling_rate();
    out[7] = s.collisional_ionization_cooling_rate_derivative();
    out[8] = s.collisional_excitation_cooling_rate();
    out[9] = s.collisional_excitation_cooling_rate_derivative();
    out[10] = s.bremsstrahlung_cooling_rate();
    out[11] = s.bremsstrahlung_cooling_rate_derivative();
    out[12] = s.compton_cooling_rate();
    out[13] = s.compton_cooling_rate_derivative();
    out[14] = s.cooling_rate();
    out[15] = s.cooling_rate_derivative();
}

// One Solver::perform_timestep on a single cell. state[0..1] = (x, T) in/out.
// Returns the timescale; *process and *failed describe it.
double orc_solver_perform_timestep(double* state, double density, double volume, double length, double rate,
                                   double scale_factor, int has_floor, double timestep, double safety,
                                   int* process, int* failed) {
    Solver s{};
    s.ionized_hydrogen_fraction = state[0];
    s.temperature = state[1];
    s.density = density;
    s.volume = volume;
    s.length = length;
    s.rate = rate;
    s.scale_factor = scale_factor;
    s.has_floor = has_floor != 0;
    s.floor_xhii = state[0];
    s.floor_temperature = state[1];
    bool f = false;
    Timescale t = s.perform_timestep(timestep, safety, &f);
    state[0] = s.ionized_hydrogen_fraction;
    state[1] = s.temperature;
    if (process) *process = t.process;
    if (failed) *failed = f ? 1 : 0;
    return t.time;
}

// The terms of one implicit Solver update at state (x, T) without floor,
// for the tests that pin temperature_change and ionized_fraction_change:
// out[0] temperature_change(dt) (mod.rs:304-310), out[1] photoheating_rate(dt)
// (:321-325), out[2] cooling_rate (:275-287), out[3] cooling_rate_derivative
// (:289-302), out[4] mu (1 / (1 + x)), out[5] ionized_fraction_change(dt)
// (:334-354), out[6] photoionization_rate(dt) (:327-332).
void orc_solver_terms(double xhii, double temperature, double density, double volume, double length, double rate,
                      double scale_factor, double timestep, double* out) {
    Solver s{};
    s.ionized_hydrogen_fraction = xhii;
    s.temperature = temperature;
    s.density = density;
    s.volume = volume;
    s.length = length;
    s.rate = rate;
    s.scale_factor = scale_factor;
    s.has_floor = false;
    out[0] = s.temperature_change(timestep);
    out[1] = s.photoheating_rate(timestep);
    out[2] = s.cooling_rate();
    out[3] = s.cooling_rate_derivative();
    out[4] = s.mu();
    out[5] = s.ionized_fraction_change(timestep);
    out[6] = s.photoionization_rate(timestep);
}

// One attempt of perform_timestep_internal (mod.rs:394-424) at depth 0: the
// clamp, then try_timestep_update (:371-392). Accepted: state[0..1] = (x, T)
// after the update and the returned timescale is the attempt's (process in
// *process). Rejected (*accepted = 0): state holds the clamped state the
// sub-steps start from, and the return value is 0. The floor, when has_floor,
// is (floor_x, floor_T).
double orc_solver_try_timestep(double* state, double density, double volume, double length, double rate,
                               double scale_factor, int has_floor, double floor_x, double floor_T, double timestep,
                               double safety, int* accepted, int* process) {
    Solver s{};
    s.ionized_hydrogen_fraction = state[0];
    s.temperature = state[1];
    s.density = density;
    s.volume = volume;
    s.length = length;
    s.rate = rate;
    s.scale_factor = scale_factor;
    s.has_floor = has_floor != 0;
    s.floor_xhii = floor_x;
    s.floor_temperature = floor_T;
    s.clamp();
    const double init_t = s.temperature, init_x = s.ionized_hydrogen_fraction;
    Timescale ts{0.0, TEMPERATURE};
    const bool ok = s.try_timestep_update(timestep, safety, ts);
    if (!ok) {
        s.temperature = init_t;
        s.ionized_hydrogen_fraction = init_x;
    }
    state[0] = s.ionized_hydrogen_fraction;
    state[1] = s.temperature;
    if (accepted) *accepted = ok ? 1 : 0;
    if (process) *process = ok ? ts.process : 0;
    return ok ? ts.time : 0.0;
}

int orc_timestep_level(int max_num_levels, double max_timestep, double desired) {
    return level_from_max_and_desired(max_num_levels, max_timestep, desired);
}

// Sweep-order levels of successive run_sweeps calls (TimestepState cadence).
// Writes up to `cap` ints per call: returns the number written for this call
// and advances the state it keeps in *lowest/*first_done.
int orc_levels_in_sweep_order(int num_levels, int* lowest, int* first_done, int* out, int cap) {
    TimestepState t;
    t.init(1.0, num_levels);
    t.current_lowest_allowed = *lowest;
    t.first_done = *first_done != 0;
    std::vector<int> v = t.levels_in_sweep_order();
    int n = (int)std::min<size_t>(v.size(), (size_t)cap);
    for (int i = 0; i < n; i++) out[i] = v[i];
    t.advance();
    *lowest = t.current_lowest_allowed;
    *first_done = t.first_done ? 1 : 0;
    return n;
}

}  // extern "C"
