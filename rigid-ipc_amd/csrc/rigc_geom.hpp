// Double-precision rigid-body geometry shared by host and device code:
// Rodrigues rotation from a rotation vector and world vertices.
//
//   construct_rotation_matrix<double>  src/physics/pose.tpp:202-218
//   sinc(double)                       src/utils/sinc.cpp:10-34
//   RigidBody::world_vertex            src/physics/rigid_body.tpp:18-24
//
// std::sin is replaced by a Cody-Waite (fdlibm three-part pi/2) + Taylor
// sine so host and device round identically; no FMA contraction.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace rigc {

#define RIGC_HD __host__ __device__ inline

RIGC_HD double geom_sin_poly(double z) {
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
RIGC_HD double geom_cos_poly(double z) {
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
RIGC_HD double geom_sin(double x) {
    if (fabs(x) <= 0.78539816339744830962) return geom_sin_poly(x);
    const double k = rint(x * 6.36619772367581382433e-01);
    const double r = ((x - k * 1.57079632673412561417e+00) - k * 6.07710050630396597660e-11)
        - k * 2.02226624879595063154e-21;
    switch (static_cast<long long>(k) & 3) {
    case 0: return geom_sin_poly(r);
    case 1: return geom_cos_poly(r);
    case 2: return -geom_sin_poly(r);
    default: return -geom_cos_poly(r);
    }
}

RIGC_HD double geom_sinc(double x) {
    if (fabs(x) >= 1.220703125e-04) return geom_sin(x) / x;
    double result = 1;
    if (fabs(x) >= 2.220446049250313e-16) {
        const double x2 = x * x;
        result -= x2 / 6.0;
        if (fabs(x) >= 1.4901161193847656e-08) result += (x2 * x2) / 120.0;
    }
    return result;
}

RIGC_HD double geom_norm3(double a, double b, double c) {
    a = fabs(a); b = fabs(b); c = fabs(c);
    return sqrt((a * a + b * b) + c * c);
}

// R = I + sinc(|r|)[r]x + 1/2 sinc(|r|/2)^2 [r]x^2, row-major R[3*i+j].
RIGC_HD void rotation_d(const double* r, double* R) {
    const double sa = geom_sinc(geom_norm3(r[0], r[1], r[2]));
    const double sh = geom_sinc(geom_norm3(r[0] / 2.0, r[1] / 2.0, r[2] / 2.0));
    const double K[3][3] = { { 0.0, -r[2], r[1] }, { r[2], 0.0, -r[0] }, { -r[1], r[0], 0.0 } };
    const double c2 = (0.5 * sh) * sh;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const double k2 = K[i][0] * K[0][j] + (K[i][1] * K[1][j] + K[i][2] * K[2][j]);
            R[3 * i + j] = sa * K[i][j] + c2 * k2;
        }
    R[0] += 1.0;
    R[4] += 1.0;
    R[8] += 1.0;
}

RIGC_HD void world_vertex_d(const double* R, const double* p, const double* v, double* w) {
    for (int j = 0; j < 3; j++) w[j] = ((v[0] * R[3 * j] + v[1] * R[3 * j + 1]) + v[2] * R[3 * j + 2]) + p[j];
}

}  // namespace rigc
