// ORACLE — test infrastructure only. Never linked into, called by or shipped
// with the product library. Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py may load it, and only as the checker.
//
// CPU restatement of the interval arithmetic used by Rigid IPC's rigid CCD.
//
//   Interval type  : boost::numeric::interval<double,
//                      policies<save_state<FILibRounding>, checking_catch_nan>>
//                    (src/interval/interval.hpp:17-30)
//   Endpoint rule  : FILibRounding (src/interval/filib_rounding.cpp:12-60):
//                    the round-to-nearest result stepped one ulp outward,
//                    with exact-zero special cases.
//   Operator cases : Boost.Interval 1.86 (third-party, not in the reference
//                    tree): arith.hpp (+,-,*,/ sign-case tables),
//                    transc.hpp/trigo.hpp (sqrt, cos, sin, fmod),
//                    utility.hpp (width, median, hull, zero_in, overlap).
//                    Restated from the published library; parity of these
//                    tables is pinned by the enclosure tests in
//                    tests/test_oracle_interval.py, not by running Boost.
//
// Hardware rounding mode: Boost's interval fmod() calls
// FILibRounding::int_down(), which switches the FPU to round-downward and,
// because cos() operates on the unprotected interval type, the mode stays
// downward until the outermost protected cos()/sin() returns
// (filib_rounding.cpp:258-268 + Boost save_state). This oracle reproduces
// that literally with fesetround(); it must be compiled with -frounding-math.
//
// cos_down/cos_up: filib's q_cos and its error factors q_cosp/q_cosm are not
// in the reference tree, so their bits cannot be reproduced. They are
// replaced by a restated Cody-Waite + Taylor cosine evaluated in
// round-to-nearest and widened by an absolute 1e-15 (a proven bound on the
// polynomial's error of < 5e-16 over [0, pi]). PARITY OF cos/sin BITS WITH
// THE REFERENCE IS UNPINNED; enclosure of the true cosine is pinned by tests.
#pragma once

#include <cfenv>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <algorithm>

namespace oracle {

inline double pred(double x) { return std::nextafter(x, -INFINITY); }
inline double succ(double x) { return std::nextafter(x, INFINITY); }

// Op-counting build (-DORACLE_COUNT_FLOPS, oracle/_build/liboracle_flops.so):
// counts every FP64 add, sub, mul, div and sqrt of the interval arithmetic
// and of the cosine kernel, to price the CCD inclusion function.
#ifdef ORACLE_COUNT_FLOPS
inline long& flop_counter() { static thread_local long c = 0; return c; }
#define ORACLE_FLOPS(n) (oracle::flop_counter() += (n))
#else
#define ORACLE_FLOPS(n) ((void)0)
#endif

// --- FILibRounding (filib_rounding.cpp:12-60); arithmetic follows the
// current hardware rounding mode, exactly as in the reference.
inline double add_down(double x, double y) { ORACLE_FLOPS(1); return x == -y ? 0.0 : pred(x + y); }
inline double add_up(double x, double y) { ORACLE_FLOPS(1); return x == -y ? 0.0 : succ(x + y); }
inline double sub_down(double x, double y) { ORACLE_FLOPS(1); return x == y ? 0.0 : pred(x - y); }
inline double sub_up(double x, double y) { ORACLE_FLOPS(1); return x == y ? 0.0 : succ(x - y); }
inline double mul_down(double x, double y) { ORACLE_FLOPS(1); return (x == 0 || y == 0) ? 0.0 : pred(x * y); }
inline double mul_up(double x, double y) { ORACLE_FLOPS(1); return (x == 0 || y == 0) ? 0.0 : succ(x * y); }
inline double div_down(double x, double y) { ORACLE_FLOPS(1); return x == 0 ? 0.0 : pred(x / y); }
inline double div_up(double x, double y) { ORACLE_FLOPS(1); return x == 0 ? 0.0 : succ(x / y); }
inline double sqrt_down(double x) { ORACLE_FLOPS(1); return x == 0 ? 0.0 : pred(std::sqrt(x)); }
inline double sqrt_up(double x) { ORACLE_FLOPS(1); return x == 0 ? 0.0 : succ(std::sqrt(x)); }

// Boost.Interval pi constants (boost/numeric/interval/constants.hpp): the two
// doubles bracketing pi, and their exact halves / doubles.
inline double pi_lower() { return (3373259426.0 + 273688.0 / (1 << 21)) / (1 << 30); }
inline double pi_upper() { return (3373259426.0 + 273689.0 / (1 << 21)) / (1 << 30); }

// Restated cosine, accurate to < 5e-16 absolute on |x| <= 4 (evaluated in
// round-to-nearest). Cody-Waite reduction to |z| <= pi/4, Taylor polynomials
// (truncation < 2e-18).
double cos_rn(double x);
// Restated sin/cos for the double-precision (non-interval) rotation path,
// used wherever the reference calls std::sin (src/utils/sinc.cpp:13).
double sin_rn_full(double x);

struct Interval {
    double lo, hi;
    Interval() : lo(0), hi(0) {}
    explicit Interval(double v) : lo(v), hi(v) {}
    Interval(double l, double u) : lo(l), hi(u) {
        if (!(l <= u)) { lo = hi = std::numeric_limits<double>::quiet_NaN(); }
    }
    static Interval raw(double l, double u) { Interval r; r.lo = l; r.hi = u; return r; }
    static Interval empty() {
        return raw(std::numeric_limits<double>::quiet_NaN(), std::numeric_limits<double>::quiet_NaN());
    }
    bool is_empty() const { return !(lo <= hi); }
    double lower() const { return lo; }
    double upper() const { return hi; }
};

inline bool empty(const Interval& x) { return x.is_empty(); }

// ---- utility.hpp
inline double width(const Interval& x) { return x.is_empty() ? 0.0 : sub_up(x.hi, x.lo); }
inline bool zero_in(const Interval& x) { return !x.is_empty() && !(x.lo > 0) && !(x.hi < 0); }
inline bool overlap(const Interval& x, const Interval& y) {
    if (x.is_empty() || y.is_empty()) return false;
    return x.lo <= y.hi && y.lo <= x.hi;
}
inline Interval hull(const Interval& x, const Interval& y) {
    bool bx = x.is_empty(), by = y.is_empty();
    if (bx) return by ? Interval::empty() : y;
    if (by) return x;
    return Interval::raw(std::min(x.lo, y.lo), std::max(x.hi, y.hi));
}
// FILibRounding::median -> filib q_mid; every box bisected by the CCD root
// finder has dyadic endpoints in [0,1], where the midpoint is exact.
inline double median(const Interval& x) { return 0.5 * (x.lo + x.hi); }

// ---- arith.hpp
inline Interval operator-(const Interval& x) {
    if (x.is_empty()) return Interval::empty();
    return Interval::raw(-x.hi, -x.lo);
}
inline Interval operator+(const Interval& x, const Interval& y) {
    if (x.is_empty() || y.is_empty()) return Interval::empty();
    return Interval::raw(add_down(x.lo, y.lo), add_up(x.hi, y.hi));
}
inline Interval operator+(const Interval& x, double y) {
    if (x.is_empty()) return Interval::empty();
    return Interval::raw(add_down(x.lo, y), add_up(x.hi, y));
}
inline Interval operator+(double x, const Interval& y) {
    if (y.is_empty()) return Interval::empty();
    return Interval::raw(add_down(x, y.lo), add_up(x, y.hi));
}
inline Interval operator-(const Interval& x, const Interval& y) {
    if (x.is_empty() || y.is_empty()) return Interval::empty();
    return Interval::raw(sub_down(x.lo, y.hi), sub_up(x.hi, y.lo));
}
inline Interval operator-(const Interval& x, double y) {
    if (x.is_empty()) return Interval::empty();
    return Interval::raw(sub_down(x.lo, y), sub_up(x.hi, y));
}
inline Interval operator-(double x, const Interval& y) {
    if (y.is_empty()) return Interval::empty();
    return Interval::raw(sub_down(x, y.hi), sub_up(x, y.lo));
}
inline Interval operator*(const Interval& x, const Interval& y) {
    if (x.is_empty() || y.is_empty()) return Interval::empty();
    const double xl = x.lo, xu = x.hi, yl = y.lo, yu = y.hi;
    if (xl < 0) {
        if (xu > 0) {
            if (yl < 0) {
                if (yu > 0)  // M * M
                    return Interval::raw(std::min(mul_down(xl, yu), mul_down(xu, yl)),
                                         std::max(mul_up(xl, yl), mul_up(xu, yu)));
                return Interval::raw(mul_down(xu, yl), mul_up(xl, yl));  // M * N
            }
            if (yu > 0) return Interval::raw(mul_down(xl, yu), mul_up(xu, yu));  // M * P
            return Interval::raw(0.0, 0.0);                                      // M * Z
        }
        if (yl < 0) {
            if (yu > 0) return Interval::raw(mul_down(xl, yu), mul_up(xl, yl));  // N * M
            return Interval::raw(mul_down(xu, yu), mul_up(xl, yl));              // N * N
        }
        if (yu > 0) return Interval::raw(mul_down(xl, yu), mul_up(xu, yl));      // N * P
        return Interval::raw(0.0, 0.0);                                          // N * Z
    }
    if (xu > 0) {
        if (yl < 0) {
            if (yu > 0) return Interval::raw(mul_down(xu, yl), mul_up(xu, yu));  // P * M
            return Interval::raw(mul_down(xu, yl), mul_up(xl, yu));              // P * N
        }
        if (yu > 0) return Interval::raw(mul_down(xl, yl), mul_up(xu, yu));      // P * P
        return Interval::raw(0.0, 0.0);                                          // P * Z
    }
    return Interval::raw(0.0, 0.0);  // Z * ?
}
inline Interval operator*(const Interval& x, double y) {
    if (x.is_empty()) return Interval::empty();
    if (y < 0) return Interval::raw(mul_down(x.hi, y), mul_up(x.lo, y));
    if (y == 0) return Interval::raw(0.0, 0.0);
    return Interval::raw(mul_down(x.lo, y), mul_up(x.hi, y));
}
inline Interval operator*(double x, const Interval& y) { return y * x; }

inline Interval div_non_zero(const Interval& x, const Interval& y) {
    const double xl = x.lo, xu = x.hi, yl = y.lo, yu = y.hi;
    if (xu < 0) {
        if (yu < 0) return Interval::raw(div_down(xu, yl), div_up(xl, yu));
        return Interval::raw(div_down(xl, yl), div_up(xu, yu));
    }
    if (xl < 0) {
        if (yu < 0) return Interval::raw(div_down(xu, yu), div_up(xl, yu));
        return Interval::raw(div_down(xl, yl), div_up(xu, yl));
    }
    if (yu < 0) return Interval::raw(div_down(xu, yu), div_up(xl, yl));
    return Interval::raw(div_down(xl, yu), div_up(xu, yl));
}
inline Interval operator/(const Interval& x, const Interval& y) {
    if (x.is_empty() || y.is_empty()) return Interval::empty();
    if (zero_in(y)) {
        // Division by an interval containing zero: the CCD path never takes
        // this branch (sinc divides by x >= 4.49); return the whole line.
        if (y.lo == 0 && y.hi == 0) return Interval::empty();
        return Interval::raw(-INFINITY, INFINITY);
    }
    return div_non_zero(x, y);
}
inline Interval operator/(const Interval& x, double y) {
    if (x.is_empty() || y == 0) return Interval::empty();
    if (y < 0) return Interval::raw(div_down(x.hi, y), div_up(x.lo, y));
    return Interval::raw(div_down(x.lo, y), div_up(x.hi, y));
}
inline Interval operator/(double x, const Interval& y) {
    if (y.is_empty()) return Interval::empty();
    if (zero_in(y)) return Interval::raw(-INFINITY, INFINITY);
    if (x < 0) return Interval::raw(div_down(x, y.lo), div_up(x, y.hi));
    return Interval::raw(div_down(x, y.hi), div_up(x, y.lo));
}

inline Interval abs(const Interval& x) {
    if (x.is_empty()) return Interval::empty();
    if (!(x.lo < 0)) return x;
    if (!(x.hi > 0)) return -x;
    return Interval::raw(0.0, std::max(-x.lo, x.hi));
}
inline Interval sqrt(const Interval& x) {
    if (x.is_empty() || x.hi < 0) return Interval::empty();
    double l = !(x.lo > 0) ? 0.0 : sqrt_down(x.lo);
    return Interval::raw(l, sqrt_up(x.hi));
}

// ---- trigo.hpp (cos, sin) with the FPU-mode behaviour described above.
Interval cos(const Interval& x);
Interval sin(const Interval& x);

// src/utils/sinc.cpp:49-101 and the double version :10-34
Interval sinc(const Interval& x);
double sinc(double x);

}  // namespace oracle
