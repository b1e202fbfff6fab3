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

// c2 = sinc(|r| / 2)^2 / 2 as the matrix's second-order coefficient.
RIGC_DI Iv rotation_c2(Iv sinc_half) { return (sinc_half * 0.5) * sinc_half; }

// Entry (i, j) of the matrix: K2 = K * K and sinc * K term by term in
// Eigen's order; a product with a diagonal (exactly zero) entry of K is the
// exact zero interval, so it is not computed, while its (outward-rounded)
// addition is kept.
RIGC_DI Iv hat_entry(const Iv r[3], int i, int j) {
    // K = [[0, -r2, r1], [r2, 0, -r0], [-r1, r0, 0]]
    if (i == j) return iv(0.0);
    const int k = 3 - i - j;  // the component of r at (i, j)
    const bool neg = (j == (i + 1) % 3);
    return neg ? -r[k] : r[k];
}
RIGC_DI Iv rotation_entry(const Iv r[3], Iv sinc_angle, Iv c2, int i, int j) {
    const Iv z = iv(0.0);
    const Iv t0 = (i == 0 || j == 0) ? z : hat_entry(r, i, 0) * hat_entry(r, 0, j);
    const Iv t1 = (i == 1 || j == 1) ? z : hat_entry(r, i, 1) * hat_entry(r, 1, j);
    const Iv t2 = (i == 2 || j == 2) ? z : hat_entry(r, i, 2) * hat_entry(r, 2, j);
    const Iv k2 = t0 + (t1 + t2);
    const Iv e = (i == j ? z : sinc_angle * hat_entry(r, i, j)) + c2 * k2;
    return i == j ? e + iv(1.0) : e;
}

RIGC_DI void rotation_from_sincs(const Iv r[3], Iv sinc_angle, Iv sinc_half, Iv R[3][3]) {
    const Iv c2 = rotation_c2(sinc_half);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) R[i][j] = rotation_entry(r, sinc_angle, c2, i, j);
}

RIGC_DI void rotation_iv(const Iv r[3], Iv R[3][3]) {
    Iv xa, xh;
    rotation_norms(r, xa, xh);
    rotation_from_sincs(r, sinc_iv(xa), sinc_iv(xh), R);
}

// Pose<Interval>::interpolate (pose.tpp:141-148) from double endpoints:
// position and rotation vector over the time interval t.
RIGC_DI Iv pose_pos_component(const double* p0, const double* p1, Iv t, int i) {
    return (iv(p1[i]) - iv(p0[i])) * t + iv(p0[i]);
}
RIGC_DI void pose_pos_iv(const double* p0, const double* p1, Iv t, Iv p[3]) {
#pragma unroll
    for (int i = 0; i < 3; i++) p[i] = pose_pos_component(p0, p1, t, i);
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
RIGC_DI Iv world_coordinate(Iv R0, Iv R1, Iv R2, Iv pj, const double* v) {
    return ((R0 * v[0] + R1 * v[1]) + R2 * v[2]) + pj;
}
RIGC_DI void world_vertex_iv(const Iv R[3][3], const Iv p[3], const double* v, Iv w[3]) {
#pragma unroll
    for (int j = 0; j < 3; j++) w[j] = world_coordinate(R[j][0], R[j][1], R[j][2], p[j], v);
}

}  // namespace rigc
