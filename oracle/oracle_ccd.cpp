// ORACLE — test infrastructure only (see oracle_interval.hpp).
#include "oracle_ccd.hpp"

namespace oracle {

namespace {

// norm() of sinc.hpp:16-23: explicit abs, then dot as a dynamic-size sum.
Interval norm3(const Vec3I& r) {
    Interval a0 = abs(r[0]), a1 = abs(r[1]), a2 = abs(r[2]);
    return sqrt((a0 * a0 + a1 * a1) + a2 * a2);
}
double norm3(const double r[3]) {
    double a0 = std::fabs(r[0]), a1 = std::fabs(r[1]), a2 = std::fabs(r[2]);
    return std::sqrt((a0 * a0 + a1 * a1) + a2 * a2);
}

template <typename T> struct Hat3 { T k[3][3]; };

}  // namespace

// pose.tpp:202-218 for T = Interval.
Mat3I rotation_matrix(const Vec3I& r) {
    const Interval sinc_angle = sinc(norm3(r));
    Vec3I rh;
    const Interval two(2.0);
    for (int i = 0; i < 3; i++) rh[i] = r[i] / two;
    const Interval sinc_half = sinc(norm3(rh));
    const Interval zero(0.0);
    Interval K[3][3] = {
        { zero, -r[2], r[1] },
        { r[2], zero, -r[0] },
        { -r[1], r[0], zero },
    };
    Interval K2[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            K2[i][j] = K[i][0] * K[0][j] + (K[i][1] * K[1][j] + K[i][2] * K[2][j]);
    const Interval c2 = (0.5 * sinc_half) * sinc_half;
    Mat3I R;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) R.m[i][j] = sinc_angle * K[i][j] + c2 * K2[i][j];
    for (int i = 0; i < 3; i++) R.m[i][i] = R.m[i][i] + Interval(1.0);
    return R;
}

// pose.tpp:202-218 for T = double.
Mat3d rotation_matrix(const double r[3]) {
    const double sinc_angle = sinc(norm3(r));
    const double rh[3] = { r[0] / 2.0, r[1] / 2.0, r[2] / 2.0 };
    const double sinc_half = sinc(norm3(rh));
    const double K[3][3] = {
        { 0.0, -r[2], r[1] },
        { r[2], 0.0, -r[0] },
        { -r[1], r[0], 0.0 },
    };
    const double c2 = (0.5 * sinc_half) * sinc_half;
    Mat3d R;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const double k2 = K[i][0] * K[0][j] + (K[i][1] * K[1][j] + K[i][2] * K[2][j]);
            R.m[i][j] = sinc_angle * K[i][j] + c2 * k2;
        }
    for (int i = 0; i < 3; i++) R.m[i][i] += 1.0;
    return R;
}

// pose.tpp:141-148 with Pose<double>::cast<Interval>() endpoints.
PoseI interpolate(const double pose_t0[6], const double pose_t1[6], const Interval& t) {
    PoseI p;
    for (int i = 0; i < 3; i++) {
        p.p[i] = (Interval(pose_t1[i]) - Interval(pose_t0[i])) * t + Interval(pose_t0[i]);
        p.r[i] = (Interval(pose_t1[3 + i]) - Interval(pose_t0[3 + i])) * t + Interval(pose_t0[3 + i]);
    }
    return p;
}

// rigid_body.tpp:18-24: (v^T R^T) + p^T, row-vector product summed left to right.
Vec3I world_vertex(const Mat3I& R, const Vec3I& p, const double v[3]) {
    Vec3I w;
    for (int j = 0; j < 3; j++)
        w[j] = ((v[0] * R.m[j][0] + v[1] * R.m[j][1]) + v[2] * R.m[j][2]) + p[j];
    return w;
}

void world_vertex(const Mat3d& R, const double p[3], const double v[3], double out[3]) {
    for (int j = 0; j < 3; j++)
        out[j] = ((v[0] * R.m[j][0] + v[1] * R.m[j][1]) + v[2] * R.m[j][2]) + p[j];
}

namespace {

// edge_trajectory_aabb (rigid_trajectory_aabb.cpp:21-36)
Vec3I edge_point(const double pose_t0[6], const double pose_t1[6], const double e0[3],
                 const double e1[3], const Interval& t, const Interval& alpha) {
    PoseI pose = interpolate(pose_t0, pose_t1, t);
    Mat3I R = rotation_matrix(pose.r);
    Vec3I w0 = world_vertex(R, pose.p, e0);
    Vec3I w1 = world_vertex(R, pose.p, e1);
    Vec3I out;
    for (int j = 0; j < 3; j++) out[j] = (w1[j] - w0[j]) * alpha + w0[j];
    return out;
}

double edge_length(const double a[3], const double b[3]) {
    const double d0 = b[0] - a[0], d1 = b[1] - a[1], d2 = b[2] - a[2];
    return std::sqrt((d0 * d0 + d1 * d1) + d2 * d2);
}

}  // namespace

// edge_edge_aabb (rigid_trajectory_aabb.cpp:70-85)
Vec3I edge_edge_aabb(const EEQuery& q, const Interval& t, const Interval& alpha, const Interval& beta) {
    Vec3I ea = edge_point(q.poseA_t0, q.poseA_t1, q.a0, q.a1, t, alpha);
    Vec3I eb = edge_point(q.poseB_t0, q.poseB_t1, q.b0, q.b1, t, beta);
    Vec3I d;
    for (int j = 0; j < 3; j++) d[j] = ea[j] - eb[j];
    return d;
}

// face_vertex_aabb (rigid_trajectory_aabb.cpp:87-102)
Vec3I face_vertex_aabb(const FVQuery& q, const Interval& t, const Interval& u, const Interval& v) {
    PoseI pa = interpolate(q.poseA_t0, q.poseA_t1, t);
    Mat3I Ra = rotation_matrix(pa.r);
    Vec3I wv = world_vertex(Ra, pa.p, q.v);

    PoseI pb = interpolate(q.poseB_t0, q.poseB_t1, t);
    Mat3I Rb = rotation_matrix(pb.r);
    Vec3I f0 = world_vertex(Rb, pb.p, q.f0);
    Vec3I f1 = world_vertex(Rb, pb.p, q.f1);
    Vec3I f2 = world_vertex(Rb, pb.p, q.f2);
    Vec3I d;
    for (int j = 0; j < 3; j++) {
        Interval fp = ((f1[j] - f0[j]) * u + (f2[j] - f0[j]) * v) + f0[j];
        d[j] = wv[j] - fp;
    }
    return d;
}

// time_of_impact.cpp:138-217
bool edge_edge_toi(const EEQuery& q, double& toi, double earliest_toi, double toi_tol, long* n_evals,
                   const std::atomic<double>* shared_min) {
    double tol[3] = { toi_tol, 1e-4 / edge_length(q.a0, q.a1), 1e-4 / edge_length(q.b0, q.b1) };
    Interval x0[3] = { Interval(0, earliest_toi), Interval(0, 1), Interval(0, 1) };
    Interval x[3];
    bool hit = interval_root_finder_nd<3>(
        [&](const Interval* p, Interval* y) {
            Vec3I d = edge_edge_aabb(q, p[0], p[1], p[2]);
            for (int i = 0; i < 3; i++) y[i] = d[i];
        },
        [](const Interval*) { return true; }, [](const Interval*) { return true; }, x0, tol, x,
        n_evals, shared_min);
    toi = hit ? x[0].lo : INFINITY;
    return hit;
}

// time_of_impact.cpp:222-310. The face's u runs along face edge 0 (f0,f1)
// and v along face edge 1 (f1,f2), whose lengths set the tolerances.
bool face_vertex_toi(const FVQuery& q, double& toi, double earliest_toi, double toi_tol, long* n_evals,
                     const std::atomic<double>* shared_min) {
    double tol[3] = { toi_tol, 1e-4 / edge_length(q.f0, q.f1), 1e-4 / edge_length(q.f1, q.f2) };
    Interval x0[3] = { Interval(0, earliest_toi), Interval(0, 1), Interval(0, 1) };
    Interval x[3];
    bool hit = interval_root_finder_nd<3>(
        [&](const Interval* p, Interval* y) {
            Vec3I d = face_vertex_aabb(q, p[0], p[1], p[2]);
            for (int i = 0; i < 3; i++) y[i] = d[i];
        },
        [](const Interval* p) { return overlap(p[1] + p[2], Interval(0, 1)); },
        [](const Interval*) { return true; }, x0, tol, x, n_evals, shared_min);
    toi = hit ? x[0].lo : INFINITY;
    return hit;
}

}  // namespace oracle
