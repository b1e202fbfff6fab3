// ORACLE — test infrastructure only (see oracle_interval.hpp).
//
// Restatement of the collision-constraint set and of the barrier / friction
// assembly of DistanceBarrierRBProblem:
//   construct_constraint_set      src/opt/distance_barrier_constraint.cpp:267-306
//                                 -> ipc::construct_constraint_set (ipc-toolkit @d41d6ca)
//   compute_barrier_term          src/problems/distance_barrier_rb_problem.cpp:1299-1399
//   apply_chain_rule              :1182-1248 (+ project_to_psd)
//   world_vertices_diff           src/physics/rigid_body.tpp:26-78 (autodiff on Rodrigues)
//   sinc_normx derivatives        src/utils/sinc.cpp:103-142, sinc.hpp:64-93
//   update_friction_constraints   :168-188 -> ipc::construct_friction_constraint_set
//   compute_friction_term         :1410-1572
//   compute_minimum_distance      src/opt/distance_barrier_constraint.cpp:308-323
// The ipc-toolkit pieces (distance-type classification, distances, IPC log
// barrier, edge-edge mollifier, tangent bases, smoothed friction f0/f1,
// project_to_psd) are not in the reference tree; they are restated from the
// IPC formulation and differentiated here with forward-mode second-order
// dual numbers, as ipc-toolkit's generated derivatives differentiate the
// same formulas. PARITY WITH THE TRUE REFERENCE IS UNPINNED at that boundary;
// it is pinned by finite differences and closed forms in the tests.
#pragma once

#include <array>
#include <vector>

#include "oracle_scene.hpp"

namespace oracle {

enum ConstraintType { VV = 0, EV = 1, EE = 2, FV = 3 };

// One collision constraint, with vertex_indices in ipc-toolkit order:
// VV (v0, v1); EV (v, e0, e1); EE (ea0, ea1, eb0, eb1); FV (v, f0, f1, f2).
struct Constraint {
    int type = 0;
    int a = 0, b = 0;  // VV (v0, v1); EV (edge, vertex); EE (edge0, edge1); FV (face, vertex)
    int nv = 0;
    int v[4] = { 0, 0, 0, 0 };
    int local_body[4] = { 0, 0, 0, 0 };
    int body[2] = { 0, 0 };
    long multiplicity = 1;
    double eps_x = -1;
};

struct FrictionConstraint {
    Constraint c;
    double coords[2] = { 0, 0 };  // closest-point parameters
    double basis[3][2];           // tangent basis (columns)
    double normal_force = 0;
    double mu = 0;
};

// Static candidates (inflation (dhat + dmin) / 2) -> constraints with
// squared distance < (dhat + dmin)^2.
void construct_constraint_set(const Scene& s, const double* poses, double dhat, double dmin,
                              std::vector<Constraint>& out);

// Sum of multiplicity * [mollifier *] b(d - dmin^2, 2 dmin dhat + dhat^2) and its
// body-space gradient (6 nb, accumulated) and per-constraint projected 12x12
// Hessians (local_hess[c] row-major, bodies c.body[0], c.body[1]).
double barrier_term(const Scene& s, const double* poses, const std::vector<Constraint>& C, double dhat, double dmin,
                    double* grad, std::vector<std::array<double, 144>>* local_hess);

double minimum_distance(const Scene& s, const double* poses, const std::vector<Constraint>& C);
// Squared distance of an EE (ea0, ea1, eb0, eb1) or FV (v, f0, f1, f2)
// vertex quadruple X[4][3], distance type classified at X.
double element_distance_sq(int type, const double* X);

// One constraint's potential and its unprojected vertex-space derivatives
// (X: 4 slots x 3; g: 12; H: 12x12 row-major).
double constraint_local(int type, const double* X, double eps_x, double dhat, double dmin, long multiplicity,
                        double* g, double* H);

void construct_friction_constraint_set(const Scene& s, const double* poses, const std::vector<Constraint>& C,
                                       double dhat, double dmin, double barrier_stiffness, double mu,
                                       std::vector<FrictionConstraint>& out);

double friction_term(const Scene& s, const double* poses, const double* poses_t0,
                     const std::vector<FrictionConstraint>& FC, double epsv_times_h, double* grad,
                     std::vector<std::array<double, 144>>* local_hess);

// (d grad_x D(x; x0) / d x0)^T lambda over the friction set (lambda, out:
// [nb][6]); the set must have been built at poses_t0 with kappa.
void friction_adjoint(const Scene& s, const double* poses, const double* poses_t0,
                      const std::vector<FrictionConstraint>& FC, double epsv_times_h, double dhat, double dmin,
                      double kappa, const double* lambda, double* out);

// Vertex position and derivatives w.r.t. its body's 6 dofs: w (3),
// jac (3x6 row-major), hess (3 x 3x3 rotational block, [j][k][l]).
void vertex_derivatives(const Scene& s, const double* pose, int vertex, double w[3], double jac[18],
                        double hess[27]);
// The same for a body-local point v.
void point_derivatives(const double* pose, const double* v, double w[3], double jac[18], double hess[27]);

// project_to_psd on a symmetric n x n matrix (row-major, in place).
void project_to_psd(double* A, int n);
// Tests switch the projection off to finite-difference the raw Hessian.
bool& psd_projection_enabled();

}  // namespace oracle
