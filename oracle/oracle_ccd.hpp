// ORACLE — test infrastructure only (see oracle_interval.hpp).
//
// Restatement of the rigid-trajectory CCD of Rigid IPC:
//   Pose<T>::interpolate            src/physics/pose.tpp:141-148
//   construct_rotation_matrix       src/physics/pose.tpp:202-218
//   norm / sinc_normx               src/utils/sinc.hpp:16-32
//   Hat                             src/utils/eigen_ext.tpp:21-29
//   RigidBody::world_vertex         src/physics/rigid_body.tpp:18-24
//   edge/face/vertex trajectories   src/ccd/rigid/rigid_trajectory_aabb.cpp:7-102
//   interval_root_finder            src/interval/interval_root_finder.cpp:97-203
//   compute_*_time_of_impact        src/ccd/rigid/time_of_impact.cpp:138-310
//
// Eigen evaluation orders restated (Eigen 3.3.7 is not in the reference
// tree): fixed-size 3-term sums (K*K) as a0 + (a1 + a2) (unrolled redux);
// dynamic-size sums (dot in norm(), row-vector times matrix in world_vertex)
// left to right ((a0 + a1) + a2). No FMA contraction.
#pragma once

#include <atomic>

#include "oracle_interval.hpp"

#include <vector>

namespace oracle {

struct Vec3I { Interval v[3]; Interval& operator[](int i) { return v[i]; } const Interval& operator[](int i) const { return v[i]; } };
struct Mat3I { Interval m[3][3]; };
struct Mat3d { double m[3][3]; };

// Rodrigues rotation from a rotation vector (pose.tpp:202-218).
Mat3I rotation_matrix(const Vec3I& r);
Mat3d rotation_matrix(const double r[3]);

// Interval world vertex R(t) v + p(t) for one body-local vertex at interval t.
struct PoseI { Vec3I p, r; };
PoseI interpolate(const double pose_t0[6], const double pose_t1[6], const Interval& t);
Vec3I world_vertex(const Mat3I& R, const Vec3I& p, const double v[3]);
void world_vertex(const Mat3d& R, const double p[3], const double v[3], double out[3]);

// Per-query rigid CCD inputs: body-local vertices + poses (dof layout
// [px,py,pz,rx,ry,rz]) at t0 and t1, as stored in the reference's CCD query
// fixtures (src/ccd/save_queries.cpp:40-185).
struct EEQuery {
    double a0[3], a1[3], b0[3], b1[3];
    double poseA_t0[6], poseA_t1[6], poseB_t0[6], poseB_t1[6];
};
struct FVQuery {
    double v[3], f0[3], f1[3], f2[3];                          // v in body A, face in body B
    double poseA_t0[6], poseA_t1[6], poseB_t0[6], poseB_t1[6];
};

// compute_edge_edge_time_of_impact / compute_face_vertex_time_of_impact
// with the t-domain [0, earliest_toi]. Returns is_impacting; toi = +inf when
// not impacting. n_evals (optional) counts inclusion-function evaluations;
// shared_min (optional) is the running minimum of an earliest-TOI call
// (interval_root_finder_nd).
bool edge_edge_toi(const EEQuery& q, double& toi, double earliest_toi = 1.0,
                   double toi_tol = 1e-4, long* n_evals = nullptr,
                   const std::atomic<double>* shared_min = nullptr);
bool face_vertex_toi(const FVQuery& q, double& toi, double earliest_toi = 1.0,
                     double toi_tol = 1e-4, long* n_evals = nullptr,
                     const std::atomic<double>* shared_min = nullptr);
// Piecewise-linear trajectories (oracle_plccd.cpp).
bool pl_edge_edge_toi(const EEQuery& q, double& toi, double earliest_toi, double msd, long* n_pieces,
                      long* n_boxes);
bool pl_face_vertex_toi(const FVQuery& q, double& toi, double earliest_toi, double msd, long* n_pieces,
                        long* n_boxes);

// Inclusion functions (rigid_trajectory_aabb.cpp:70-102), exposed for tests.
Vec3I edge_edge_aabb(const EEQuery& q, const Interval& t, const Interval& alpha, const Interval& beta);
Vec3I face_vertex_aabb(const FVQuery& q, const Interval& t, const Interval& u, const Interval& v);

// Generic 1-D root finder used by the reference's own root-finder tests
// (interval_root_finder.cpp:20-47 -> :97-203).
template <typename F, typename P>
bool interval_root_finder_1d(F f, P constraint_predicate, const Interval& x0, double tol, Interval& x);

}  // namespace oracle

#include "oracle_root_finder.tpp"
