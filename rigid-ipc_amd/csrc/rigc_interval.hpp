// Device interval arithmetic for the rigid-trajectory CCD (gfx950).
//
// Semantics follow the reference's interval type, boost::interval<double,
// save_state<FILibRounding>> (src/interval/interval.hpp:17-30,
// src/interval/filib_rounding.cpp:12-60): every endpoint is the
// round-to-nearest result stepped one ulp outward, except that an exactly
// zero result stays zero. Because that outward step is monotone, the
// endpoint of a product is the stepped min/max of the four endpoint
// products, which is what this file computes (branch-free, no sign-case
// table). The stretch of Boost's cos() after fmod() runs in the FPU's
// round-downward mode in the reference (FILibRounding::int_down, never
// restored inside cos()); here those operations are rounded downward
// exactly from round-to-nearest results and their error-free residuals
// (TwoSum for sums, an FMA residual for products and quotients). The OCML
// directed-rounding entry points behind HIP's _rd intrinsics are not present
// in this ROCm device library, and the residual form needs no mode switch.
//
// filib's cos kernel (q_cos with q_cosp/q_cosm) is not available, so
// cos_down/cos_up evaluate a Cody-Waite + Taylor cosine (error < 5e-16 on
// [0, 4]) and widen it by 1e-15 absolute before clamping to [-1, 1].
#pragma once

#include <hip/hip_runtime.h>

namespace rigc {

struct Iv {
    double lo, hi;
};

#define RIGC_DI __device__ __forceinline__

// Outward one-ulp steps that keep an exact zero at +0.
RIGC_DI double step_dn(double p) {
    long long b = __double_as_longlong(p);
    b += (p > 0.0) ? -1 : 1;
    return p == 0.0 ? 0.0 : __longlong_as_double(b);
}
RIGC_DI double step_up(double p) {
    long long b = __double_as_longlong(p);
    b += (p > 0.0) ? 1 : -1;
    return p == 0.0 ? 0.0 : __longlong_as_double(b);
}

// IEEE next-down / next-up (nextafter toward -inf / +inf) of a finite value.
RIGC_DI double next_dn(double v) {
    if (v == 0.0) return -4.9406564584124654e-324;
    return __longlong_as_double(__double_as_longlong(v) + (v > 0.0 ? -1 : 1));
}
RIGC_DI double next_up(double v) {
    if (v == 0.0) return 4.9406564584124654e-324;
    return __longlong_as_double(__double_as_longlong(v) + (v > 0.0 ? 1 : -1));
}

// Round-downward sum, difference, product and quotient of finite operands.
RIGC_DI double add_rd(double a, double b) {
    const double s = a + b, bb = s - a;
    const double err = (a - (s - bb)) + (b - bb);
    return err < 0.0 ? next_dn(s) : s;
}
RIGC_DI double sub_rd(double a, double b) { return add_rd(a, -b); }
RIGC_DI double mul_rd(double a, double b) {
    const double p = a * b;
    return fma(a, b, -p) < 0.0 ? next_dn(p) : p;
}
RIGC_DI double div_rd(double a, double b) {
    const double q = a / b;
    const double r = fma(-q, b, a);
    return (r != 0.0 && ((r < 0.0) != (b < 0.0))) ? next_dn(q) : q;
}

RIGC_DI Iv iv(double a) { return Iv{ a, a }; }
RIGC_DI Iv iv(double a, double b) { return Iv{ a, b }; }

RIGC_DI Iv operator+(Iv x, Iv y) { return Iv{ step_dn(x.lo + y.lo), step_up(x.hi + y.hi) }; }
RIGC_DI Iv operator-(Iv x, Iv y) { return Iv{ step_dn(x.lo - y.hi), step_up(x.hi - y.lo) }; }
RIGC_DI Iv operator-(Iv x) { return Iv{ -x.hi, -x.lo }; }
RIGC_DI Iv operator+(Iv x, double y) { return Iv{ step_dn(x.lo + y), step_up(x.hi + y) }; }
RIGC_DI Iv operator+(double x, Iv y) { return y + x; }
RIGC_DI Iv operator-(Iv x, double y) { return Iv{ step_dn(x.lo - y), step_up(x.hi - y) }; }

RIGC_DI Iv operator*(Iv x, Iv y) {
    const double a = x.lo * y.lo, b = x.lo * y.hi, c = x.hi * y.lo, d = x.hi * y.hi;
    return Iv{ step_dn(fmin(fmin(a, b), fmin(c, d))), step_up(fmax(fmax(a, b), fmax(c, d))) };
}
RIGC_DI Iv operator*(Iv x, double y) {
    const double a = x.lo * y, b = x.hi * y;
    return Iv{ step_dn(fmin(a, b)), step_up(fmax(a, b)) };
}
RIGC_DI Iv operator*(double x, Iv y) { return y * x; }
RIGC_DI Iv operator/(Iv x, double y) {
    const double a = x.lo / y, b = x.hi / y;
    return Iv{ step_dn(fmin(a, b)), step_up(fmax(a, b)) };
}
// Division by an interval that does not contain zero.
RIGC_DI Iv div_nonzero(Iv x, Iv y) {
    const double a = x.lo / y.lo, b = x.lo / y.hi, c = x.hi / y.lo, d = x.hi / y.hi;
    return Iv{ step_dn(fmin(fmin(a, b), fmin(c, d))), step_up(fmax(fmax(a, b), fmax(c, d))) };
}

RIGC_DI bool iv_empty(Iv x) { return !(x.lo <= x.hi); }
RIGC_DI Iv iv_emptyval() { return Iv{ __longlong_as_double(0x7ff8000000000000LL), __longlong_as_double(0x7ff8000000000000LL) }; }
RIGC_DI Iv hull(Iv x, Iv y) {
    if (iv_empty(x)) return y;
    if (iv_empty(y)) return x;
    return Iv{ fmin(x.lo, y.lo), fmax(x.hi, y.hi) };
}
RIGC_DI bool zero_in(Iv x) { return x.lo <= 0.0 && x.hi >= 0.0; }
RIGC_DI bool overlap(Iv x, Iv y) { return x.lo <= y.hi && y.lo <= x.hi; }
// Boost width(): hi - lo rounded up with the outward step.
RIGC_DI double width(Iv x) { return step_up(x.hi - x.lo); }

RIGC_DI Iv iabs(Iv x) {
    if (!(x.lo < 0.0)) return x;
    if (!(x.hi > 0.0)) return Iv{ -x.hi, -x.lo };
    return Iv{ 0.0, fmax(-x.lo, x.hi) };
}
RIGC_DI Iv isqrt(Iv x) {
    const double l = (x.lo > 0.0) ? step_dn(sqrt(x.lo)) : 0.0;
    return Iv{ l, step_up(sqrt(x.hi)) };
}

// ---- round-downward variants (the FPU state inside Boost's cos after fmod)
RIGC_DI Iv sub_rd(Iv x, Iv y) {
    return Iv{ step_dn(sub_rd(x.lo, y.hi)), step_up(sub_rd(x.hi, y.lo)) };
}
RIGC_DI Iv mul_rd(Iv x, double y) {
    const double a = mul_rd(x.lo, y), b = mul_rd(x.hi, y);
    return Iv{ step_dn(fmin(a, b)), step_up(fmax(a, b)) };
}

// ---- restated cosine (Cody-Waite + Taylor), round-to-nearest
RIGC_DI double cos_poly(double z) {
    const double z2 = z * z;
    double p = 1.0 / 20922789888000.0;
    p = p * z2 - 1.0 / 87178291200.0;
    p = p * z2 + 1.0 / 479001600.0;
    p = p * z2 - 1.0 / 3628800.0;
    p = p * z2 + 1.0 / 40320.0;
    p = p * z2 - 1.0 / 720.0;
    p = p * z2 + 1.0 / 24.0;
    p = p * z2 - 0.5;
    return p * z2 + 1.0;
}
RIGC_DI double sin_poly(double z) {
    const double z2 = z * z;
    double q = 1.0 / 355687428096000.0;
    q = q * z2 - 1.0 / 1307674368000.0;
    q = q * z2 + 1.0 / 6227020800.0;
    q = q * z2 - 1.0 / 39916800.0;
    q = q * z2 + 1.0 / 362880.0;
    q = q * z2 - 1.0 / 5040.0;
    q = q * z2 + 1.0 / 120.0;
    q = q * z2 - 1.0 / 6.0;
    return z + (z * z2) * q;
}
RIGC_DI double cos_rn(double x) {
    x = fabs(x);
    if (x <= 0.78539816339744830962) return cos_poly(x);
    if (x <= 2.35619449019234492885)
        return -sin_poly((x - 1.57079632679489655800e+00) - 6.12323399573676603587e-17);
    return -cos_poly((3.14159265358979311600e+00 - x) + 1.22464679914735317720e-16);
}
RIGC_DI double clamp1(double c) { return fmin(fmax(c, -1.0), 1.0); }
RIGC_DI double cos_down(double x) { return clamp1(next_dn(cos_rn(x) - 1e-15)); }
RIGC_DI double cos_up(double x) { return clamp1(next_up(cos_rn(x) + 1e-15)); }

// Restated sine for the double-precision rotation path (fdlibm three-part
// pi/2 reduction, valid for |x| < 2^20 * pi/2).
__host__ __device__ inline double rigc_sin_poly_hd(double z) {
    const double z2 = z * z;
    double q = 1.0 / 355687428096000.0;
    q = q * z2 - 1.0 / 1307674368000.0;
    q = q * z2 + 1.0 / 6227020800.0;
    q = q * z2 - 1.0 / 39916800.0;
    q = q * z2 + 1.0 / 362880.0;
    q = q * z2 - 1.0 / 5040.0;
    q = q * z2 + 1.0 / 120.0;
    q = q * z2 - 1.0 / 6.0;
    return z + (z * z2) * q;
}
__host__ __device__ inline double rigc_cos_poly_hd(double z) {
    const double z2 = z * z;
    double p = 1.0 / 20922789888000.0;
    p = p * z2 - 1.0 / 87178291200.0;
    p = p * z2 + 1.0 / 479001600.0;
    p = p * z2 - 1.0 / 3628800.0;
    p = p * z2 + 1.0 / 40320.0;
    p = p * z2 - 1.0 / 720.0;
    p = p * z2 + 1.0 / 24.0;
    p = p * z2 - 0.5;
    return p * z2 + 1.0;
}

// Boost interval constants: the doubles bracketing pi, 2*pi, pi/2.
#define RIGC_PI_LO ((3373259426.0 + 273688.0 / 2097152.0) / 1073741824.0)
#define RIGC_PI_HI ((3373259426.0 + 273689.0 / 2097152.0) / 1073741824.0)

// Boost cos() entered in round-to-nearest mode: fmod by [2pi] (quotient
// rounded with div_down, floored), after which all arithmetic is in the
// round-downward state; at most one reflection through pi.
RIGC_DI Iv cos_iv(Iv x) {
    const Iv pi2 = Iv{ RIGC_PI_LO * 2, RIGC_PI_HI * 2 };
    const Iv pi = Iv{ RIGC_PI_LO, RIGC_PI_HI };
    bool negate = false;
    bool downward = false;  // FPU state at the current fmod
    Iv r;
#pragma unroll 1
    for (int level = 0; level < 4; level++) {
        const double yb = (x.lo < 0.0) ? pi2.lo : pi2.hi;
        const double q = downward ? step_dn(div_rd(x.lo, yb)) : step_dn(x.lo / yb);
        const double n = floor(q);
        downward = true;
        const Iv tmp = sub_rd(x, mul_rd(pi2, n));
        const double w = step_up(sub_rd(tmp.hi, tmp.lo));
        if (w >= pi2.lo) {
            r = Iv{ -1.0, 1.0 };
            break;
        }
        if (tmp.lo >= RIGC_PI_HI) {
            x = sub_rd(tmp, pi);
            negate = !negate;
            continue;
        }
        const double l = tmp.lo, u = tmp.hi;
        if (u <= RIGC_PI_LO) {
            r = Iv{ cos_down(u), cos_up(l) };
        } else if (u <= pi2.lo) {
            r = Iv{ -1.0, cos_up(fmin(step_dn(sub_rd(pi2.lo, u)), l)) };
        } else {
            r = Iv{ -1.0, 1.0 };
        }
        break;
    }
    return negate ? Iv{ -r.hi, -r.lo } : r;
}

// Boost sin(x) = cos(x - [pi/2]) with the subtraction in round-to-nearest.
RIGC_DI Iv sin_iv(Iv x) { return cos_iv(x - Iv{ RIGC_PI_LO / 2, RIGC_PI_HI / 2 }); }

// One end of cos_iv(x) (the upper one if `upper`), with the same operations
// as cos_iv for that end: the reduction is shared, and only the cosine the
// end needs is evaluated.
RIGC_DI double cos_iv_end(Iv x, bool upper) {
    const Iv pi2 = Iv{ RIGC_PI_LO * 2, RIGC_PI_HI * 2 };
    const Iv pi = Iv{ RIGC_PI_LO, RIGC_PI_HI };
    bool negate = false;
    bool downward = false;
    // r = [-1, 1] (kind 0), [cos_down(u), cos_up(l)] (kind 1) or
    // [-1, cos_up(m)] (kind 2)
    int kind = 0;
    double l = 0.0, u = 0.0;
#pragma unroll 1
    for (int level = 0; level < 4; level++) {
        const double yb = (x.lo < 0.0) ? pi2.lo : pi2.hi;
        const double q = downward ? step_dn(div_rd(x.lo, yb)) : step_dn(x.lo / yb);
        const double n = floor(q);
        downward = true;
        const Iv tmp = sub_rd(x, mul_rd(pi2, n));
        const double w = step_up(sub_rd(tmp.hi, tmp.lo));
        if (w >= pi2.lo) break;
        if (tmp.lo >= RIGC_PI_HI) {
            x = sub_rd(tmp, pi);
            negate = !negate;
            continue;
        }
        l = tmp.lo;
        u = tmp.hi;
        if (u <= RIGC_PI_LO) {
            kind = 1;
        } else if (u <= pi2.lo) {
            kind = 2;
            l = fmin(step_dn(sub_rd(pi2.lo, u)), l);
        }
        break;
    }
    // the end of r that gives the requested end after the negation
    const bool r_hi = upper != negate;
    double e;
    if (kind == 0 || (kind == 2 && !r_hi)) {
        e = r_hi ? 1.0 : -1.0;
    } else {
        const double c = cos_rn(r_hi ? l : u);
        e = r_hi ? clamp1(next_up(c + 1e-15)) : clamp1(next_dn(c - 1e-15));
    }
    return negate ? -e : e;
}

// _sinc_interval_taylor (src/utils/sinc.cpp:36-47)
RIGC_DI Iv sinc_point(double xd) {
    const Iv x = iv(xd);
    if (fabs(xd) >= 1.220703125e-04) return div_nonzero(sin_iv(x), x);
    const Iv x2 = x * x;
    return 1.0 + (x2 / 6.0) * ((x2 / 20.0) - 1.0);
}

// One end of sinc_point(xd) for xd >= 0, the same bits as sinc_point's.
// Off the Taylor branch xd > 0, so the quotient by the point [xd, xd] is
// monotone: its ends are those of sin over xd, and each needs one end of
// the sine.
RIGC_DI double sinc_point_end(double xd, bool upper) {
    if (!(fabs(xd) >= 1.220703125e-04)) {
        const Iv x = iv(xd);
        const Iv x2 = x * x;
        const Iv y = 1.0 + (x2 / 6.0) * ((x2 / 20.0) - 1.0);
        return upper ? y.hi : y.lo;
    }
    const Iv a = iv(xd) - Iv{ RIGC_PI_LO / 2, RIGC_PI_HI / 2 };
    const double s = cos_iv_end(a, upper);
    return upper ? step_up(s / xd) : step_dn(s / xd);
}

// sinc(Interval) (src/utils/sinc.cpp:49-101) for x.lo >= 0 (x is always a
// norm here), in three parts that may run on different lanes: the upper
// end of the monotone piece (sinc at x.lo), its lower end (sinc at
// min(x.hi, mb)) together with the enclosure beyond mb, and their
// combination.
#define RIGC_SINC_MB 4.4934094579
RIGC_DI bool sinc_has_gt(Iv x) { return !(x.lo <= RIGC_SINC_MB) || x.hi > RIGC_SINC_MB; }
RIGC_DI double sinc_part_lo(Iv x) { return x.lo <= RIGC_SINC_MB ? sinc_point_end(x.lo, true) : 0.0; }
RIGC_DI double sinc_part_hi(Iv x, Iv& g) {
    double v = 0.0;
    if (x.lo <= RIGC_SINC_MB) v = sinc_point_end(x.hi > RIGC_SINC_MB ? RIGC_SINC_MB : x.hi, false);
    g = iv(0.0);
    if (sinc_has_gt(x)) {
        const Iv xg = (x.lo <= RIGC_SINC_MB) ? Iv{ RIGC_SINC_MB, x.hi } : x;
        g = div_nonzero(sin_iv(xg), xg);
    }
    return v;
}
RIGC_DI Iv sinc_combine(Iv x, double lo_part, double hi_part, Iv g) {
    Iv y = iv_emptyval();
    if (x.lo <= RIGC_SINC_MB) y = Iv{ fmax(hi_part, -0.271724), fmin(lo_part, 1.0) };
    if (sinc_has_gt(x)) y = hull(y, g);
    return y;
}
RIGC_DI Iv sinc_iv(Iv x) {
    Iv g;
    const double lo = sinc_part_lo(x);
    const double hi = sinc_part_hi(x, g);
    return sinc_combine(x, lo, hi, g);
}

}  // namespace rigc
