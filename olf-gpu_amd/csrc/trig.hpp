// sin and cos of one double for the branch-flow terms, small enough in registers to keep the Newton-step kernel's
// fill and J * v passes from spilling: Cody-Waite reduction by pi/2 (W. J. Cody and W. Waite, Software Manual for the
// Elementary Functions, 1980), k = rint(x * 2/pi) and r = x - k pi/2 with pi/2 split into three doubles and each
// product subtracted by one fused multiply-add (exact products, so the reduction stays accurate while k pi/2 is within
// a rounding of x: |x| up to about 1e15), then the Taylor series of sin (degree 17) and cos (degree 16) on
// |r| <= pi/4, whose truncation stays below 1e-19. Absolute error about 3e-16. Arguments beyond 1e15 rad, which only
// a diverged Newton iterate reaches, are first brought below it modulo the double nearest 2 pi (so sin and cos stay
// finite and bounded, with no claim of accuracy); NaN and infinities give NaN. Also compiles as host code.
#pragma once

#include <cmath>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define OLFHIP_HD __host__ __device__
#else
#define OLFHIP_HD
#endif

namespace olfhip {

// pi/2 = PIO2_A + PIO2_B + PIO2_C to about 160 bits (each the double nearest to what the previous ones leave)
constexpr double PIO2_A = 0x1.921fb54442d18p+0;
constexpr double PIO2_B = 0x1.1a62633145c07p-54;
constexpr double PIO2_C = -0x1.f1976b7ed8fbcp-110;
constexpr double TWO_OVER_PI = 0.6366197723675814;
constexpr double TWO_PI = 0x1.921fb54442d18p+2;
constexpr double SINCOS_REDUCED_MAX = 1e15;

OLFHIP_HD inline void sincosReduced(double x, double *s, double *c) {
    // each pass leaves at most pi + |x| 2^-52, so a finite x comes below the bound within about twenty passes
    while (std::fabs(x) > SINCOS_REDUCED_MAX) x = std::fma(-std::rint(x * (1.0 / TWO_PI)), TWO_PI, x);
    const double kd = std::rint(x * TWO_OVER_PI);
    double r = std::fma(-kd, PIO2_A, x);
    r = std::fma(-kd, PIO2_B, r);
    r = std::fma(-kd, PIO2_C, r);
    const double z = r * r;
    // sin r = r - r z (1/3! - z/5! + ...), cos r = 1 - z/2 + z^2 (1/4! - z/6! + ...), both by Horner in z = r^2
    double ps = -1.0 / 355687428096000.0;          // -1/17!
    ps = std::fma(ps, z, 1.0 / 1307674368000.0);   //  1/15!
    ps = std::fma(ps, z, -1.0 / 6227020800.0);     // -1/13!
    ps = std::fma(ps, z, 1.0 / 39916800.0);        //  1/11!
    ps = std::fma(ps, z, -1.0 / 362880.0);         // -1/9!
    ps = std::fma(ps, z, 1.0 / 5040.0);            //  1/7!
    ps = std::fma(ps, z, -1.0 / 120.0);
    ps = std::fma(ps, z, 1.0 / 6.0);
    // ps = 1/3! - z/5! + z^2/7! - ...: sin r = r - r z ps
    const double sr = std::fma(-r * z, ps, r);
    double pc = 1.0 / 20922789888000.0;            //  1/16!; pc = 1/4! - z/6! + ...
    pc = std::fma(pc, z, -1.0 / 87178291200.0);    // -1/14!
    pc = std::fma(pc, z, 1.0 / 479001600.0);       //  1/12!
    pc = std::fma(pc, z, -1.0 / 3628800.0);        // -1/10!
    pc = std::fma(pc, z, 1.0 / 40320.0);           //  1/8!
    pc = std::fma(pc, z, -1.0 / 720.0);            // -1/6!
    pc = std::fma(pc, z, 1.0 / 24.0);              //  1/4!
    const double hz = 0.5 * z;
    const double w = 1.0 - hz;
    const double cr = w + (((1.0 - w) - hz) + z * z * pc);
    const int q = (int)(kd - 4.0 * std::floor(kd * 0.25));   // kd mod 4, exact: kd is an integer below 2^51
    const double sv = (q & 1) ? cr : sr, cv = (q & 1) ? sr : cr;
    *s = (q & 2) ? -sv : sv;
    *c = ((q + 1) & 2) ? -cv : cv;
}

}  // namespace olfhip
