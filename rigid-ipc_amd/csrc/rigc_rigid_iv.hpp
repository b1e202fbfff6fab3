// Interval rigid-body kinematics on the device: pose interpolation over an
// interval of time, interval Rodrigues rotation and interval world vertices.
#pragma once

#include "rigc_interval.hpp"

namespace rigc {

// construct_rotation_matrix<Interval> (src/physics/pose.tpp:202-218), in
// parts: the norms of r and r / 2 (the sinc arguments), and the matrix
// from the two sinc values.
RIGC_DI void rotation_norms(const Iv r[3], Iv& x_angle, Iv& x_half) {
    const Iv a0 = iabs(r[0]), a1 = iabs(r[1]), a2 = iabs(r[2]);
    x_angle = isqrt((a0 * a0 + a1 * a1) + a2 * a2);
    Iv h[3];
    for (int i = 0; i < 3; i++) h[i] = r[i] / 2.0;
    const Iv h0 = iabs(h[0]), h1 = iabs(h[1]), h2 = iabs(h[2]);
    x_half = isqrt((h0 * h0 + h1 * h1) + h2 * h2);
}

RIGC_DI void rotation_from_sincs(const Iv r[3], Iv sinc_angle, Iv sinc_half, Iv R[3][3]) {
    const Iv z = iv(0.0);
    const Iv K[3][3] = { { z, -r[2], r[1] }, { r[2], z, -r[0] }, { -r[1], r[0], z } };
    const Iv c2 = (sinc_half * 0.5) * sinc_half;
    // K2 = K * K and sinc * K term by term in Eigen's order; a product with
    // a diagonal (exactly zero) entry of K is the exact zero interval, so it
    // is not computed, while its (outward-rounded) addition is kept.
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const Iv t0 = (i == 0 || j == 0) ? z : K[i][0] * K[0][j];
            const Iv t1 = (i == 1 || j == 1) ? z : K[i][1] * K[1][j];
            const Iv t2 = (i == 2 || j == 2) ? z : K[i][2] * K[2][j];
            const Iv k2 = t0 + (t1 + t2);
            R[i][j] = (i == j ? z : sinc_angle * K[i][j]) + c2 * k2;
        }
#pragma unroll
    for (int i = 0; i < 3; i++) R[i][i] = R[i][i] + iv(1.0);
}

RIGC_DI void rotation_iv(const Iv r[3], Iv R[3][3]) {
    Iv xa, xh;
    rotation_norms(r, xa, xh);
    rotation_from_sincs(r, sinc_iv(xa), sinc_iv(xh), R);
}

// Pose<Interval>::interpolate (pose.tpp:141-148) from double endpoints:
// position and rotation vector over the time interval t.
RIGC_DI void pose_pos_iv(const double* p0, const double* p1, Iv t, Iv p[3]) {
#pragma unroll
    for (int i = 0; i < 3; i++) p[i] = (iv(p1[i]) - iv(p0[i])) * t + iv(p0[i]);
}
RIGC_DI void pose_rot_iv(const double* p0, const double* p1, Iv t, Iv r[3]) {
#pragma unroll
    for (int i = 0; i < 3; i++) r[i] = (iv(p1[3 + i]) - iv(p0[3 + i])) * t + iv(p0[3 + i]);
}

RIGC_DI void pose_iv(const double* p0, const double* p1, Iv t, Iv p[3], Iv R[3][3]) {
    Iv r[3];
    pose_pos_iv(p0, p1, t, p);
    pose_rot_iv(p0, p1, t, r);
    rotation_iv(r, R);
}

// RigidBody::world_vertex (src/physics/rigid_body.tpp:18-24)
RIGC_DI void world_vertex_iv(const Iv R[3][3], const Iv p[3], const double* v, Iv w[3]) {
    const double v0 = v[0], v1 = v[1], v2 = v[2];
#pragma unroll
    for (int j = 0; j < 3; j++) w[j] = ((R[j][0] * v0 + R[j][1] * v1) + R[j][2] * v2) + p[j];
}

}  // namespace rigc
