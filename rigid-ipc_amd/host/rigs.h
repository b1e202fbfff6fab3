/*
 * rigs — host time stepper of Rigid IPC over a contact backend.
 *
 * The Newton solve stays on the host, as in the reference: this layer
 * restates DistanceBarrierRBProblem's simulation step
 * (src/problems/distance_barrier_rb_problem.cpp:118-626: update_constraints,
 * solve_constraints with lagged friction, take_step with the implicit-Euler
 * velocity update, compute_objective = E / m + kappa / m B + D / m), the
 * IPC solver's adaptive barrier stiffness (src/solvers/ipc_solver.cpp) and
 * the Newton solver with its CCD-filtered backtracking line search and
 * Tikhonov regularisation (src/solvers/newton_solver.cpp:160-520). The
 * contact terms (CCD, constraint set, barrier, friction, minimum distance)
 * go through a rigs_backend: by default the device library (include/rigc.h);
 * any other implementation of the same table can be passed in.
 *
 * Layout: poses / velocities [num_bodies][6] (position, rotation vector;
 * linear velocity in world coordinates, angular velocity in the body's
 * principal frame), inertia [num_bodies][3] principal moments.
 */
#ifndef RIGS_H
#define RIGS_H

#include <stdint.h>

#include "../../include/rigc.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Contact backend: the rigc entry points with the context as `self`. */
typedef struct rigs_backend {
    void* self;
    rigc_status (*set_scene)(void* self, const rigc_scene* scene);
    rigc_status (*compute_earliest_toi)(void* self, const double* poses_t0, const double* poses_t1,
                                        double minimum_separation_distance, double* earliest_toi);
    rigc_status (*construct_constraint_set)(void* self, const double* poses, double dhat, double dmin,
                                            int64_t* counts);
    rigc_status (*compute_barrier_term)(void* self, const double* poses, double dhat, double dmin, int32_t flags,
                                        double* energy, double* grad, rigc_block_hessian* hess);
    rigc_status (*compute_minimum_distance)(void* self, const double* poses, double dhat, double dmin,
                                            int32_t flags, double* min_distance);
    rigc_status (*update_friction_constraints)(void* self, const double* poses, double dhat, double dmin,
                                               double barrier_stiffness, double mu, int32_t flags, int64_t* count);
    rigc_status (*compute_friction_term)(void* self, const double* poses, const double* poses_t0,
                                         double epsv_times_h, int32_t flags, double* energy, double* grad,
                                         rigc_block_hessian* hess);
    const char* (*last_error)(void* self);
    /* Optional (may be NULL): counters of the last call (rigc_get_stats). */
    rigc_status (*get_stats)(void* self, rigc_stats* stats);
    /* Optional (may be NULL): number of intersecting edge-face pairs at
     * poses (rigc_detect_intersections without the list). */
    rigc_status (*detect_intersections)(void* self, const double* poses, int64_t* count);
    /* Optional (may be NULL; then piecewise-linear scenes are refused):
     * compute_earliest_toi over piecewise-linear trajectories
     * (rigc_compute_earliest_toi_piecewise_linear). */
    rigc_status (*compute_earliest_toi_piecewise_linear)(void* self, const double* poses_t0, const double* poses_t1,
                                                         double minimum_separation_distance, double* earliest_toi);
    /* Optional (may be NULL; then rollouts cannot be differentiated):
     * rigc_friction_adjoint. */
    rigc_status (*friction_adjoint)(void* self, const double* poses, const double* poses_t0, double epsv_times_h,
                                    double dhat, double dmin, double barrier_stiffness, const double* lambda,
                                    double* out);
    /* Optional (may be NULL; then only RIGC_DETECTION_HASH_GRID): the
     * broad phase of compute_earliest_toi (rigc_set_detection_method). */
    rigc_status (*set_detection_method)(void* self, int32_t method);
} rigs_backend;

/* Trajectory types of the CCD (TrajectoryType, src/ccd/ccd.hpp:30-38; the
 * linearised and Redon types are not built). */
enum { RIGS_TRAJECTORY_RIGID = 0, RIGS_TRAJECTORY_PIECEWISE_LINEAR = 1 };

/* Simulation settings; rigs_default_params fills the reference defaults
 * (src/SimState.cpp:130-208, src/constants.hpp). */
typedef struct rigs_params {
    double timestep;
    double gravity[3];
    double dhat;                       /* initial_barrier_activation_distance */
    double minimum_separation_distance;
    double coefficient_friction;
    int32_t friction_iterations;
    double static_friction_speed_bound;
    double velocity_conv_tol;          /* relative to the world bbox diagonal */
    double line_search_lower_bound;
    int32_t max_iterations;
    double dhat_epsilon;
    double min_barrier_stiffness_scale;
    int32_t do_intersection_check;     /* check the poses of every step for
                                        * intersections (RigidBodyProblem,
                                        * rigid_body_problem.cpp:261-267) */
    int32_t trajectory_type;           /* RIGS_TRAJECTORY_*: the line search's
                                        * CCD (distance_barrier_constraint.trajectory_type) */
    int32_t newton_polish;             /* full Newton steps taken after the solve has
                                        * converged (0 in the reference): each is
                                        * taken when the CCD allows the whole step,
                                        * with the barrier stiffness held, so that the
                                        * step ends at the minimiser to rounding and
                                        * a rollout is a smooth function of its inputs
                                        * (differentiable rollouts) */
    int32_t detection_method;          /* RIGC_DETECTION_*: the broad phase of the
                                        * line search's CCD (distance_barrier_constraint.
                                        * detection_method; the reference's default is
                                        * BVH, rigs_default_params sets HASH_GRID) */
} rigs_params;

typedef struct rigs_bodies {
    int32_t num_bodies;
    const double* poses;
    const double* velocities;
    const double* mass;
    const double* inertia;
    /* Optional (NULL: none): external force and torque [num_bodies][6],
     * constant, in world coordinates (RigidBody::force; compute_body_energy,
     * distance_barrier_rb_problem.cpp:847-944). */
    const double* force;
    /* Optional (NULL: none): kinematic bodies (RigidBodyType::KINEMATIC).
     * kinematic[b] = 1 marks body b (not static in the scene): it has no body
     * energy, and an augmented Lagrangian (distance_barrier_rb_problem.cpp:
     * 190-235, 339-435, 946-1109) moves it each step to its next scripted
     * pose (kinematic_poses[kinematic_pose_start[b] ..], consumed one per
     * step; degrees of freedom as `poses`) or, without one, by its velocity
     * times the timestep. Once kinematic_max_time[b] (NULL or < 0: never)
     * runs out it becomes static (step_kinematic_bodies, :237-251). */
    const int32_t* kinematic;
    const double* kinematic_max_time;
    const int32_t* kinematic_pose_start; /* [num_bodies + 1], NULL: no scripted poses */
    const double* kinematic_poses;       /* [kinematic_pose_start[num_bodies]][6] */
} rigs_bodies;

typedef struct rigs_stats {
    int64_t steps;
    int64_t newton_iterations;
    int64_t line_search_iterations;
    int64_t ccd_calls;
    int64_t constraint_builds;
    int64_t objective_evaluations;
    int64_t regularizations;
    int64_t failed_solves;
    double contact_ms;   /* time inside backend calls */
    double host_ms;      /* body energy, assembly of the Newton system, LDL^T; the body
                          * energy computed beside backend calls counts only for its
                          * wait */
    double total_ms;
    double kappa;        /* barrier stiffness of the last step */
    double min_distance; /* after the last step (-1: no constraints) */
    int64_t num_constraints; /* of the last constraint-set build */
    /* CCD calls (line-search filter): candidates, host wall time around the
     * call, and, when the backend reports them, device broad / narrow phase
     * time and interval box evaluations. */
    int64_t ccd_candidates;
    double ccd_ms;
    double ccd_broad_phase_ms;
    double ccd_narrow_phase_ms;
    int64_t ccd_box_evaluations;
    double ccd_kernel_ms;        /* CCD traversal kernel launches alone */
    int64_t ccd_launches;
    /* host wall time of the other backend calls, by kind */
    double constraint_ms;        /* construct_constraint_set */
    double barrier_ms;           /* compute_barrier_term */
    double friction_ms;          /* compute_friction_term, update_friction_constraints */
    double distance_ms;          /* compute_minimum_distance */
    /* Intersection checks (backend's detect_intersections): intersecting
     * edge-face pairs of the initial state (checked at creation, as
     * rigid_body_problem.cpp:100 does), of the last step's end poses when
     * do_intersection_check is set, and the number of steps that ended
     * with intersections; -1 where no check ran. */
    int64_t initial_intersections;
    int64_t intersections;
    int64_t intersecting_steps;
} rigs_stats;

typedef struct rigs_sim rigs_sim;

void rigs_default_params(rigs_params* params);

/* backend NULL: the device library on HIP device `device`. */
rigc_status rigs_create(rigs_sim** sim, const rigs_backend* backend, int32_t device, const rigc_scene* scene,
                        const rigs_bodies* bodies, const rigs_params* params);
rigc_status rigs_destroy(rigs_sim* sim);
const char* rigs_last_error(const rigs_sim* sim);
/* Advance num_steps time steps (DistanceBarrierRBProblem::simulation_step). */
rigc_status rigs_step(rigs_sim* sim, int32_t num_steps);
rigc_status rigs_get_state(const rigs_sim* sim, double* poses, double* velocities);
/* Back to poses / velocities (layout of rigs_bodies) with the solver state
 * of a fresh rigs_create (barrier stiffness, caches); clears the record. */
rigc_status rigs_set_state(rigs_sim* sim, const double* poses, const double* velocities);

/* ---- Differentiable rollouts (no counterpart in the reference, which has
 * no gradients: python/src/bindings.cpp:30-193).
 * rigs_rollout advances num_steps steps as rigs_step does and keeps a
 * checkpoint of each step's start state (poses, velocities, Qdot) and of
 * the barrier stiffnesses its solve used.
 * rigs_rollout_backward differentiates a loss L of the rollout's end poses
 * and (optionally, may be NULL) end velocities: given dL_dposes [nb][6]
 * and dL_dvelocities [nb][6], it writes dL/d(initial velocities) [nb][6]
 * (linear, then angular in the body frame, as rigs_bodies holds them) and,
 * when dL_dposes0 is not NULL, dL/d(initial poses). Each step is rebuilt
 * from its checkpoint (constraint set at its end poses, lagged friction set
 * at its start poses) and differentiated through its converged Newton
 * solve by the implicit function theorem with the unprojected Hessian.
 * The barrier stiffness is held at the values the forward solve used.
 * Needs converged steps, one lagged friction set per step, and a backend
 * with friction_adjoint when mu > 0; the simulation's end state is kept. */
rigc_status rigs_rollout(rigs_sim* sim, int32_t num_steps);
rigc_status rigs_rollout_backward(rigs_sim* sim, const double* dL_dposes, const double* dL_dvelocities,
                                  double* dL_dvelocities0, double* dL_dposes0);
rigc_status rigs_get_stats(const rigs_sim* sim, rigs_stats* stats);

/* ---- Batches of independent scenes (no counterpart in the reference,
 * which steps one scene per process).
 * RIGS_BATCH_FUSED: each scene's stepper runs on a thread of its own; their
 *   contact calls are gathered into rounds and served by fused calls on one
 *   multi-scene device context (rigc_set_scenes / rigc_batch_*). Results
 *   are bit-identical to stepping each scene alone.
 * RIGS_BATCH_SEQUENTIAL: one scene after another, each with its own
 *   single-scene context (the plain path). */
enum { RIGS_BATCH_SEQUENTIAL = 0, RIGS_BATCH_FUSED = 1 };

typedef struct rigs_batch_stats {
    int64_t rounds;          /* executor rounds (requests taken together) */
    int64_t fused_calls;     /* device calls made for those rounds */
    int64_t requests;        /* scene requests they served */
    double device_ms;        /* wall time of the rounds, summed over executors */
    int64_t ccd_calls;       /* fused earliest-TOI calls */
    int64_t ccd_candidates;
    double ccd_kernel_ms;
    int64_t ccd_launches;
    int64_t ccd_box_evaluations;
    int32_t executors;       /* device contexts (streams) serving the rounds */
    int32_t pad;
    rigc_tuning tuning;      /* CCD tunables the autotuner chose (zeros: not run) */
    double autotune_ms;      /* their device time over the autotuner's sample */
} rigs_batch_stats;

typedef struct rigs_batch rigs_batch;

/* scenes, bodies, params: arrays of num_scenes. */
rigc_status rigs_batch_create(rigs_batch** batch, int32_t device, int32_t num_scenes, const rigc_scene* scenes,
                              const rigs_bodies* bodies, const rigs_params* params, int32_t mode);
rigc_status rigs_batch_destroy(rigs_batch* batch);
const char* rigs_batch_last_error(const rigs_batch* batch);
/* Every scene advances num_steps time steps. */
rigc_status rigs_batch_step(rigs_batch* batch, int32_t num_steps);
/* Fused mode's dispatch rule. A round runs once every scene still stepping
 * has a request waiting, or once the oldest waiting request has waited
 * gather_us microseconds (gather_us < 0: only the first rule). The default
 * is the environment's RIGS_BATCH_GATHER_US, else 200. Results do not
 * depend on it, only the grouping of requests into rounds. */
rigc_status rigs_batch_set_gather(rigs_batch* batch, double gather_us);
rigc_status rigs_batch_get_state(const rigs_batch* batch, int32_t scene, double* poses, double* velocities);
rigc_status rigs_batch_get_stats(const rigs_batch* batch, int32_t scene, rigs_stats* stats);
rigc_status rigs_batch_get_batch_stats(const rigs_batch* batch, rigs_batch_stats* stats);

/* Sparse block LDL^T solve used for the Newton system (exposed for tests):
 * n bodies, diag [n][36], m pairs (i != j) with blocks (i,j), rhs b [6n]. */
rigc_status rigs_block_ldlt_solve(int32_t n, const double* diag, int64_t m, const int32_t* pairs,
                                  const double* off, const double* b, double* x);

#ifdef __cplusplus
}
#endif

#endif /* RIGS_H */
