// Timestep-level selection shared by host and device code.
//
// TimestepLevel::from_max_timestep_and_desired_timestep
// (src/sweep/timestep_level.rs:27-36) computes clamp(ceil(log2(max/desired)) as
// usize, 0, L-1). The integer result must not depend on which log2
// implementation runs, so ceil(log2(r)) is evaluated exactly from the binary
// exponent of r, treating log2 as correctly rounded: r = m 2^e with m in
// [0.5, 1); log2(r) = (e-1) + log2(2m). If 2m == 1 the result is e-1. Otherwise
// log2(r) lies strictly between e-1 and e, and a correctly rounded log2 returns
// exactly e-1 only when log2(2m) is below half an ulp of e-1; then the ceiling
// is e-1, else e. Rust's `as usize` saturates: NaN and negatives give 0.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#define SSW_LM_HD __host__ __device__ inline
#else
#define SSW_LM_HD inline
#endif

namespace ssw {

SSW_LM_HD int level_from_ratio(double ratio, int num_levels) {
    const int top = num_levels - 1;
    if (ratio != ratio) return 0;      // NaN -> 0
    if (!(ratio > 0.0)) return 0;      // log2 of 0 or negative: -inf / NaN -> 0
    if (ratio == INFINITY) return top;
    int e = 0;
    double m = frexp(ratio, &e);       // exact
    double t = 2.0 * m - 1.0;          // exact, in [0, 1)
    int k = e - 1;
    int c;                             // ceil(log2(ratio)) under correct rounding
    if (t == 0.0) {
        c = k;
    } else if (k >= 1) {
        // half ulp of k: 2^(floor(log2 k) - 53)
        int fl = 0;
        for (int v = k; v > 1; v >>= 1) fl++;
        double half_ulp = ldexp(1.0, fl - 53);
        double delta = t * 1.4426950408889634 * (1.0 - 0.5 * t);  // log2(1+t) for small t
        c = (t < 1e-6 && delta < half_ulp) ? k : e;
    } else {
        c = e;                          // k <= 0: the level is 0 either way
    }
    if (c <= 0) return 0;
    return c >= top ? top : c;
}

}  // namespace ssw
