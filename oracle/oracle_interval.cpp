// ORACLE — test infrastructure only (see oracle_interval.hpp).
#include "oracle_interval.hpp"

#include <cfloat>

namespace oracle {

namespace {

// Restores the FPU rounding mode on scope exit, like Boost's
// save_state<Rounding> (the protected-interval rounding object).
struct SaveState {
    int mode;
    SaveState() : mode(std::fegetround()) {}
    ~SaveState() { std::fesetround(mode); }
};

// Cody-Waite split constants (fdlibm): pi/2 and pi as hi + lo.
const double PIO2_HI = 1.57079632679489655800e+00;  // 0x3FF921FB54442D18
const double PIO2_LO = 6.12323399573676603587e-17;  // 0x3C91A62633145C07
const double PI_HI = 3.14159265358979311600e+00;    // 0x400921FB54442D18
const double PI_LO = 1.22464679914735317720e-16;    // 0x3CA1A62633145C07
const double PIO4 = 0.78539816339744830962;
const double THREE_PIO4 = 2.35619449019234492885;

// Taylor polynomial of cos on |z| <= 0.87 (truncation < 1e-17).
double cos_poly(double z) {
    const double z2 = z * z;
    double p = 1.0 / 20922789888000.0;  // 1/16!
    p = p * z2 - 1.0 / 87178291200.0;   // 1/14!
    p = p * z2 + 1.0 / 479001600.0;     // 1/12!
    p = p * z2 - 1.0 / 3628800.0;       // 1/10!
    p = p * z2 + 1.0 / 40320.0;         // 1/8!
    p = p * z2 - 1.0 / 720.0;           // 1/6!
    p = p * z2 + 1.0 / 24.0;            // 1/4!
    p = p * z2 - 0.5;                   // 1/2!
    return p * z2 + 1.0;
}

// Taylor polynomial of sin on |z| <= 0.8 (truncation < 2e-19).
double sin_poly(double z) {
    const double z2 = z * z;
    double q = 1.0 / 355687428096000.0;  // 1/17!
    q = q * z2 - 1.0 / 1307674368000.0;  // 1/15!
    q = q * z2 + 1.0 / 6227020800.0;     // 1/13!
    q = q * z2 - 1.0 / 39916800.0;       // 1/11!
    q = q * z2 + 1.0 / 362880.0;         // 1/9!
    q = q * z2 - 1.0 / 5040.0;           // 1/7!
    q = q * z2 + 1.0 / 120.0;            // 1/5!
    q = q * z2 - 1.0 / 6.0;              // 1/3!
    return z + (z * z2) * q;
}

// FILibRounding::cos_down / cos_up (filib_rounding.cpp:102-136) with the
// filib kernel replaced by cos_rn() +- 1e-15, clamped to [-1, 1].
const double COS_ABS_ERR = 1e-15;

double cos_down(double x) {
    double c;
    {
        SaveState s;
        std::fesetround(FE_TONEAREST);
        c = pred(cos_rn(x) - COS_ABS_ERR);
        ORACLE_FLOPS(1);
    }
    if (c < -1.0) c = -1.0;
    if (c > 1.0) c = 1.0;
    return c;
}

double cos_up(double x) {
    double c;
    {
        SaveState s;
        std::fesetround(FE_TONEAREST);
        c = succ(cos_rn(x) + COS_ABS_ERR);
        ORACLE_FLOPS(1);
    }
    if (c < -1.0) c = -1.0;
    if (c > 1.0) c = 1.0;
    return c;
}

// Boost fmod(interval, interval) on the unprotected type: the quotient's
// floor is taken through FILibRounding::int_down(), which leaves the FPU in
// round-downward mode for everything that follows.
Interval fmod_unprotected(const Interval& x, const Interval& y) {
    if (x.is_empty() || y.is_empty()) return Interval::empty();
    const double yb = (x.lo < 0) ? y.lo : y.hi;
    const double q = div_down(x.lo, yb);
    std::fesetround(FE_DOWNWARD);
    const double n = std::rint(q);
    return x - n * y;
}

// Boost cos() on the unprotected type (no rounding-mode save/restore).
Interval cos_unprotected(const Interval& x) {
    if (x.is_empty()) return Interval::empty();
    const Interval pi2 = Interval::raw(pi_lower() * 2, pi_upper() * 2);
    const Interval tmp = fmod_unprotected(x, pi2);
    if (width(tmp) >= pi2.lo) return Interval::raw(-1.0, 1.0);
    if (tmp.lo >= pi_upper())
        return -cos_unprotected(tmp - Interval::raw(pi_lower(), pi_upper()));
    const double l = tmp.lo, u = tmp.hi;
    if (u <= pi_lower()) return Interval::raw(cos_down(u), cos_up(l));
    if (u <= pi2.lo) return Interval::raw(-1.0, cos_up(std::min(sub_down(pi2.lo, u), l)));
    return Interval::raw(-1.0, 1.0);
}

const double taylor_0_bound = DBL_EPSILON;               // sinc.cpp:6
const double taylor_2_bound = 1.4901161193847656e-08;    // sqrt(eps) = 2^-26
const double taylor_n_bound = 1.220703125e-04;           // sqrt(sqrt(eps)) = 2^-13

// sinc.cpp:36-47 (_sinc_interval_taylor)
Interval sinc_interval_taylor(double x_double) {
    Interval x(x_double);
    if (std::fabs(x.lo) >= taylor_n_bound) return sin(x) / x;
    Interval x2 = x * x;
    return 1.0 + x2 / 6.0 * (x2 / 20.0 - 1.0);
}

}  // namespace

double cos_rn(double x) {
    // cos_poly: 1 mul + 9 (mul, add); sin_poly: 2 + 8 (mul, add) + 2; reduction 2
    ORACLE_FLOPS(x <= PIO4 ? 19 : 22);
    x = std::fabs(x);
    if (x <= PIO4) return cos_poly(x);
    if (x <= THREE_PIO4) return -sin_poly((x - PIO2_HI) - PIO2_LO);
    return -cos_poly((PI_HI - x) + PI_LO);
}

double sin_rn_full(double x) {
    // fdlibm three-part pi/2 (pio2_1, pio2_2, pio2_2t); k*pio2_1 and
    // k*pio2_2 are exact for |k| < 2^20.
    const double INV_PIO2 = 6.36619772367581382433e-01;
    const double P1 = 1.57079632673412561417e+00;  // 0x3FF921FB54400000
    const double P2 = 6.07710050630396597660e-11;  // 0x3DD0B4611A600000
    const double P3 = 2.02226624879595063154e-21;  // 0x3BA3198A2E037073
    if (std::fabs(x) <= PIO4) return sin_poly(x);
    const double k = std::nearbyint(x * INV_PIO2);
    const double r = ((x - k * P1) - k * P2) - k * P3;
    long q = static_cast<long>(k) & 3;
    switch (q) {
    case 0: return sin_poly(r);
    case 1: return cos_poly(r);
    case 2: return -sin_poly(r);
    default: return -cos_poly(r);
    }
}

Interval cos(const Interval& x) {
    SaveState s;
    return cos_unprotected(x);
}

Interval sin(const Interval& x) {
    SaveState s;
    if (x.is_empty()) return Interval::empty();
    const Interval pi_half = Interval::raw(pi_lower() / 2, pi_upper() / 2);
    return cos_unprotected(x - pi_half);
}

// sinc.cpp:49-101
Interval sinc(const Interval& x) {
    static const double monotonic_bound = 4.4934094579;
    Interval y = Interval::empty(), x_pos = x;
    if (x.lo < 0) {
        if (x.hi <= 0) return sinc(-x);
        y = sinc(Interval(0, -x.lo));
        x_pos = Interval(0, x.hi);
    }
    Interval x_gt_monotonic = x_pos;
    if (x_pos.lo <= monotonic_bound) {
        Interval x_monotonic = x_pos;
        if (x_monotonic.hi > monotonic_bound) {
            x_monotonic = Interval(x_monotonic.lo, monotonic_bound);
            x_gt_monotonic = Interval(monotonic_bound, x_pos.hi);
        } else {
            x_gt_monotonic = Interval::empty();
        }
        y = hull(y, Interval(std::max(sinc_interval_taylor(x_monotonic.hi).lo, -0.271724),
                             std::min(sinc_interval_taylor(x_monotonic.lo).hi, 1.0)));
    }
    if (!empty(x_gt_monotonic)) y = hull(y, sin(x_gt_monotonic) / x_gt_monotonic);
    return y;
}

// sinc.cpp:10-34, with std::sin replaced by sin_rn_full().
double sinc(double x) {
    if (std::fabs(x) >= taylor_n_bound) return sin_rn_full(x) / x;
    double result = 1;
    if (std::fabs(x) >= taylor_0_bound) {
        double x2 = x * x;
        result -= x2 / 6.0;
        if (std::fabs(x) >= taylor_2_bound) result += (x2 * x2) / 120.0;
    }
    return result;
}

}  // namespace oracle
