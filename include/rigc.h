/*
 * rigc — MI355X (gfx950) contact hot path for rigid-body IPC.
 *
 * A C ABI over device kernels for the broad phase, the rigid-trajectory
 * interval CCD and the barrier / friction assembly of Rigid IPC. Every
 * entry point takes plain host pointers and sizes, returns a rigc_status
 * and never throws. One context owns one device, one HIP stream and the
 * device copies of one scene; calls on a context are serialised by the
 * caller (the reference calls these from its single Newton thread).
 *
 * Layout conventions (those of the reference, src/physics/pose.tpp:32-43,
 * src/physics/rigid_body_assembler.cpp:19-60):
 *   poses      [num_bodies][6] doubles: position x,y,z then rotation vector
 *   vertices   body-local (principal-axis frame) coordinates, concatenated
 *              body after body; edge / face ids are global (concatenated)
 *   candidates edge-edge pairs (e0 < e1); face-vertex pairs (face, vertex)
 */
#ifndef RIGC_H
#define RIGC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t rigc_status;
#define RIGC_OK 0
#define RIGC_ERR_INVALID_ARGUMENT 1
#define RIGC_ERR_DEVICE 2
#define RIGC_ERR_NO_SCENE 3
#define RIGC_ERR_CAPACITY 4   /* output too small; the count written is the size needed */
#define RIGC_ERR_UNRESOLVED 5 /* a CCD query exceeded its evaluation budget */

typedef struct rigc_ctx rigc_ctx;

/* Scene geometry: replaces the RigidBodyAssembler state the reference's
 * broad phase, CCD and barrier problem read
 * (src/physics/rigid_body_assembler.hpp, src/physics/rigid_body.hpp). */
typedef struct rigc_scene {
    int32_t num_bodies;
    int32_t num_vertices;
    int32_t num_edges;
    int32_t num_faces;
    const double* vertices;            /* [num_vertices][3] body-local */
    const int32_t* edges;              /* [num_edges][2] global vertex ids */
    const int32_t* faces;              /* [num_faces][3] global vertex ids */
    const int32_t* face_edges;         /* [num_faces][3] global edge of (f0,f1),(f1,f2),(f2,f0) */
    const int32_t* body_vertex_start;  /* [num_bodies+1] */
    const int32_t* body_edge_start;    /* [num_bodies+1] */
    const int32_t* body_face_start;    /* [num_bodies+1] */
    const int32_t* group_ids;          /* [num_bodies] */
    const uint8_t* is_static;          /* [num_bodies] RigidBodyType::STATIC */
} rigc_scene;

/* Device-side timings and counters of the most recent call. */
typedef struct rigc_stats {
    double broad_phase_ms;     /* all broad-phase kernels of the last call */
    double narrow_phase_ms;    /* the CCD kernel of the last call */
    double assembly_ms;        /* the assembly kernels of the last call */
    int64_t box_evaluations;   /* inclusion-function evaluations (CCD) */
    int64_t num_candidates;    /* candidates passed to the narrow phase */
    int64_t num_constraints;   /* records of the constraint set the last building call built */
    double ccd_kernel_ms;      /* the CCD traversal kernel launches alone (CCD) */
    int64_t ccd_launches;      /* CCD traversal kernel launches (passes) of the last call */
} rigc_stats;

/* Create a context on HIP device `device` (ordinal). */
rigc_status rigc_create(rigc_ctx** ctx, int32_t device);
rigc_status rigc_destroy(rigc_ctx* ctx);
/* Message of the last failing call on this context ("" if none). */
const char* rigc_last_error(const rigc_ctx* ctx);
rigc_status rigc_get_stats(const rigc_ctx* ctx, rigc_stats* stats);

/* Upload a scene. Replaces RigidBodyAssembler::init
 * (src/physics/rigid_body_assembler.cpp:19-118): derives per-edge rest
 * lengths, per-body r_max and the mean edge length on the host. */
rigc_status rigc_set_scene(rigc_ctx* ctx, const rigc_scene* scene);

/* Upload several independent scenes into one context (batched simulation;
 * no counterpart in the reference, which holds one scene per
 * RigidBodyAssembler). Scene k keeps its own body / element numbering in
 * the rigc_batch_* calls below; no body of one scene is ever paired with a
 * body of another. The single-scene entry points above and below act on
 * all scenes at once (energies summed, earliest TOI and minimum distance
 * taken over all scenes). */
rigc_status rigc_set_scenes(rigc_ctx* ctx, int32_t num_scenes, const rigc_scene* scenes);
rigc_status rigc_get_scene_count(const rigc_ctx* ctx, int32_t* num_scenes);

/* Narrow phase over given candidates. Replaces
 * DistanceBarrierConstraint::compute_earliest_toi_narrow_phase
 * (src/opt/distance_barrier_constraint.cpp:151-265) with edge_edge_ccd /
 * face_vertex_ccd (src/ccd/ccd.cpp:221-357) for TrajectoryType::RIGID,
 * i.e. compute_{edge_edge,face_vertex}_time_of_impact
 * (src/ccd/rigid/time_of_impact.cpp:158-310). Every query searches
 * t in [0,1]. earliest_toi receives the minimum TOI (+inf if none).
 * toi_out / hit_out (optional, length n_ee + n_fv, EE first) receive
 * per-query results: hit 1 = impact, 0 = none, -1 = budget exceeded.
 * eval_budget > 0 caps box visits per query (0 = unbounded, as the
 * reference). per_query != 0 disables cross-query pruning so that every
 * query's own TOI is computed. */
rigc_status rigc_narrow_phase(rigc_ctx* ctx, const double* poses_t0, const double* poses_t1,
                              const int32_t* ee, int64_t n_ee, const int32_t* fv, int64_t n_fv,
                              double toi_tolerance, int64_t eval_budget, int32_t per_query,
                              double* earliest_toi, double* toi_out, int8_t* hit_out);

/* rigc_narrow_phase that allocates nothing: no counterpart in the reference;
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid it). Same arguments
 * and results as rigc_narrow_phase with eval_budget = 0. Every per-call
 * buffer (poses, candidate lists, per-query state, results) lives in one
 * static device array of RIGC_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends; a call needs about 96 bytes per body of the scene plus 41
 * per query (each buffer rounded up to 256 bytes) and returns
 * RIGC_ERR_CAPACITY, allocating nothing, when that does not fit. The narrow
 * phase runs in one pass, which gives the same results. Calls that overlap
 * on one GPU (two contexts or threads at once) share the array and race. */
#define RIGC_NOALLOC_SCRATCH_BYTES (32u << 20)
rigc_status rigc_narrow_phase_noalloc(rigc_ctx* ctx, const double* poses_t0, const double* poses_t1,
                                      const int32_t* ee, int64_t n_ee, const int32_t* fv, int64_t n_fv,
                                      double toi_tolerance, int32_t per_query, double* earliest_toi,
                                      double* toi_out, int8_t* hit_out);

/* rigc_narrow_phase's earliest TOI with device-resident inputs and output,
 * queued on the caller's stream as kernels only (no host copy, no
 * synchronisation, no memset), so that a caller can capture it in a HIP
 * graph and replay it on new contents of the same buffers. No counterpart
 * in the reference (its CCD runs on the host); for callers that batch
 * device work in graphs. poses_t0 / poses_t1 [num_bodies][6], ee [n_ee][2],
 * fv [n_fv][2] and earliest_toi (one double, +inf when nothing collides)
 * are device pointers (the element ids are not range-checked: they stay
 * on the device); the search is the one-pass narrow phase of
 * rigc_narrow_phase with eval_budget 0 (same result). The per-call state
 * lives in a buffer of the context used by no other call: a call outside
 * capture grows it (waiting for the stream first), a call under capture
 * that would need more returns RIGC_ERR_CAPACITY, and a graph captured
 * before a call that grew it must be captured again. Calls that overlap on
 * one context share the state and race. */
rigc_status rigc_narrow_phase_enqueue(rigc_ctx* ctx, void* stream, const double* poses_t0, const double* poses_t1,
                                      const int32_t* ee, int64_t n_ee, const int32_t* fv, int64_t n_fv,
                                      double toi_tolerance, double* earliest_toi);

/* Broad phase. Replaces detect_collision_candidates_rigid with
 * DetectionMethod::HASH_GRID, continuous (poses_t1 != NULL;
 * src/ccd/rigid/broad_phase.cpp:127-203, RigidBodyHashGrid::{resize,addBodies}
 * src/ccd/rigid/rigid_body_hash_grid.cpp:90-344) or discrete at poses_t0
 * (poses_t1 == NULL; broad_phase.cpp:19-90, :36-62, :122-162).
 * Writes sorted, duplicate-free edge-edge (e0 < e1) and face-vertex
 * (face, vertex) pairs. If a capacity is too small, nothing is copied,
 * *n_ee / *n_fv hold the sizes needed and RIGC_ERR_CAPACITY is returned. */
rigc_status rigc_broad_phase(rigc_ctx* ctx, const double* poses_t0, const double* poses_t1,
                             double inflation_radius, int32_t* ee, int64_t ee_capacity, int64_t* n_ee,
                             int32_t* fv, int64_t fv_capacity, int64_t* n_fv);

/* Earliest time of impact of the rigid trajectories poses_t0 -> poses_t1:
 * continuous broad phase with inflation minimum_separation_distance / 2,
 * then the narrow phase with cross-query pruning, all on the device.
 * Replaces DistanceBarrierConstraint::compute_earliest_toi
 * (src/opt/distance_barrier_constraint.cpp:130-149) for RIGID trajectories.
 * *earliest_toi = +inf when nothing collides. */
rigc_status rigc_compute_earliest_toi(rigc_ctx* ctx, const double* poses_t0, const double* poses_t1,
                                      double minimum_separation_distance, double* earliest_toi);

/* Narrow phase for TrajectoryType::PIECEWISE_LINEAR. Replaces
 * edge_edge_ccd / face_vertex_ccd (src/ccd/ccd.cpp:271-274, 339-342) with
 * compute_piecewise_linear_{edge_edge,face_vertex}_time_of_impact
 * (src/ccd/piecewise_linear/time_of_impact.cpp:296-649): the rigid
 * trajectory is cut into pieces until each piece's deviation from a straight
 * line is bounded, and each piece goes through Tight-Inclusion's linear CCD
 * with that bound added to the separation distance. Every query searches
 * t in [0, earliest_toi_in] (the reference's earliest_toi argument).
 * toi_out / hit_out (optional, length n_ee + n_fv, EE first): hit 1 =
 * impact (toi_out = TOI), 0 = none (+inf). *earliest_toi = the minimum over
 * impacting queries (+inf if none). */
rigc_status rigc_narrow_phase_piecewise_linear(rigc_ctx* ctx, const double* poses_t0, const double* poses_t1,
                                               const int32_t* ee, int64_t n_ee, const int32_t* fv, int64_t n_fv,
                                               double earliest_toi_in, double minimum_separation_distance,
                                               double* earliest_toi, double* toi_out, int8_t* hit_out);

/* rigc_compute_earliest_toi for TrajectoryType::PIECEWISE_LINEAR: the same
 * continuous hash-grid broad phase (inflation minimum_separation_distance / 2,
 * src/ccd/ccd.cpp:75-81), then the piecewise-linear narrow phase over
 * t in [0,1] (DistanceBarrierConstraint::compute_earliest_toi,
 * src/opt/distance_barrier_constraint.cpp:130-149). */
rigc_status rigc_compute_earliest_toi_piecewise_linear(rigc_ctx* ctx, const double* poses_t0, const double* poses_t1,
                                                       double minimum_separation_distance, double* earliest_toi);

/* ---------------------------------------------------------------- assembly
 * Derivative flags of the potential entry points. */
#define RIGC_GRADIENT 1
#define RIGC_HESSIAN 2            /* implies RIGC_GRADIENT */
#define RIGC_BUILD_CONSTRAINTS 4  /* rebuild the constraint set at `poses` first */
#define RIGC_UNPROJECTED 8        /* Hessian blocks without project_to_psd: the exact
                                   * Hessian that an implicit-function adjoint needs
                                   * (the reference always projects,
                                   * distance_barrier_rb_problem.cpp:1241) */

/* Block-sparse symmetric Hessian in the 6-dof body layout. diag holds
 * every body's (i,i) block; pairs / off hold the coupled body pairs
 * (i < j) with block (i,j); block (j,i) is its transpose. Row-major 6x6.
 * count is written with the number of off-diagonal blocks; when it
 * exceeds capacity, pairs / off are left untouched and the call returns
 * RIGC_ERR_CAPACITY (energy, grad and diag are still written). */
typedef struct rigc_block_hessian {
    double* diag;     /* [num_bodies][36] */
    int32_t* pairs;   /* [capacity][2] */
    double* off;      /* [capacity][36] */
    int64_t capacity;
    int64_t count;
} rigc_block_hessian;

/* Collision constraint set at `poses`: static hash-grid candidates with
 * inflation (dhat + dmin) / 2, kept if the squared distance is below
 * (dhat + dmin)^2, classified by closest-feature pair, vertex-vertex and
 * edge-vertex duplicates merged into a multiplicity, nearly parallel
 * edge-edge pairs mollified (eps_x from the current positions).
 * Replaces DistanceBarrierConstraint::construct_constraint_set
 * (src/opt/distance_barrier_constraint.cpp:267-306). The set stays on the
 * device for the calls below; counts[4] receives VV, EV, EE, FV sizes. */
rigc_status rigc_construct_constraint_set(rigc_ctx* ctx, const double* poses, double dhat, double dmin,
                                          int64_t* counts);

/* Copy of the current constraint set, VV then EV then EE then FV, each
 * sorted by its key: records [n][6] = type (0 VV, 1 EV, 2 EE, 3 FV), a, b
 * (VV: v0 < v1; EV: edge, vertex; EE: edge0, edge1; FV: face, vertex),
 * multiplicity, body0, body1; eps_x [n] (-1 unless EE). */
rigc_status rigc_get_constraint_set(rigc_ctx* ctx, int64_t capacity, int32_t* records, double* eps_x, int64_t* n);

/* Barrier potential sum_c m_c [mollifier] b(d_c - dmin^2, 2 dmin dhat + dhat^2)
 * over the current constraint set, its gradient (grad [num_bodies][6]) and
 * its Hessian with each constraint's 12x12 block projected to PSD.
 * Replaces DistanceBarrierRBProblem::compute_barrier_term
 * (src/problems/distance_barrier_rb_problem.cpp:1299-1399, with
 * apply_chain_rule :1182-1248). grad / hess may be NULL when the matching
 * flag is off. */
rigc_status rigc_compute_barrier_term(rigc_ctx* ctx, const double* poses, double dhat, double dmin, int32_t flags,
                                      double* energy, double* grad, rigc_block_hessian* hess);

/* sqrt of the minimum squared distance over the constraint set (+inf when
 * the set is empty), the set being rebuilt at `poses` first when flags
 * holds RIGC_BUILD_CONSTRAINTS. Replaces
 * DistanceBarrierConstraint::compute_minimum_distance
 * (src/opt/distance_barrier_constraint.cpp:308-323). */
rigc_status rigc_compute_minimum_distance(rigc_ctx* ctx, const double* poses, double dhat, double dmin,
                                          int32_t flags, double* min_distance);

/* Edge-triangle intersections at `poses`: every (edge, face) pair, the
 * edge and the face on two close bodies of different groups, whose closed
 * segment and closed triangle meet (a segment lying in the triangle's
 * plane is not counted), decided with exact orientation predicates. *count
 * receives the number of such pairs; pairs (optional, [capacity][2] =
 * edge, face, sorted) receives them, RIGC_ERR_CAPACITY when more than
 * capacity were found. The reference's answer is *count > 0. Replaces
 * RigidBodyProblem::detect_intersections (3D branch,
 * src/physics/rigid_body_problem.cpp:289-360, with
 * detect_intersection_candidates_rigid_bvh, src/ccd/rigid/broad_phase.cpp:265-305,
 * and ipc-toolkit's is_edge_intersecting_triangle). stats.broad_phase_ms
 * receives its device time. */
rigc_status rigc_detect_intersections(rigc_ctx* ctx, const double* poses, int32_t* pairs, int64_t capacity,
                                      int64_t* count);

/* Lagged friction set from the constraint set (rebuilt at `poses` first
 * when flags holds RIGC_BUILD_CONSTRAINTS): each constraint's closest
 * point, tangent basis and normal force -kappa b'(d) 2 sqrt(d) at `poses`
 * are frozen (ipc::construct_friction_constraint_set as called by
 * DistanceBarrierRBProblem::update_friction_constraints,
 * src/problems/distance_barrier_rb_problem.cpp:168-188). The friction set
 * survives later rebuilds of the constraint set. */
rigc_status rigc_update_friction_constraints(rigc_ctx* ctx, const double* poses, double dhat, double dmin,
                                             double barrier_stiffness, double mu, int32_t flags, int64_t* count);

/* Copy of the friction set: records [n][9] = normal force, closest-point
 * coordinates (2), tangent basis (3x2 row-major), in constraint-set order. */
rigc_status rigc_get_friction_constraints(rigc_ctx* ctx, int64_t capacity, double* records, int64_t* n);

/* Friction potential sum_c m_c mu lambda_c f0(|T_c^T (V(poses) - V(poses_t0))|;
 * epsv_times_h) with derivatives as rigc_compute_barrier_term. Replaces
 * DistanceBarrierRBProblem::compute_friction_term
 * (src/problems/distance_barrier_rb_problem.cpp:1410-1572). */
rigc_status rigc_compute_friction_term(rigc_ctx* ctx, const double* poses, const double* poses_t0,
                                       double epsv_times_h, int32_t flags, double* energy, double* grad,
                                       rigc_block_hessian* hess);

/* Adjoint of the friction gradient with respect to the start-of-step poses:
 * out = (d grad_x D(x; x0) / d x0)^T lambda, with x = poses, x0 = poses_t0
 * and D the friction potential of rigc_compute_friction_term over the
 * current lagged friction set, which must have been built at poses_t0 with
 * barrier_stiffness, dhat and dmin. D depends on x0 through the lagged
 * displacement V(x) - V(x0) and through each normal force
 * -kappa b'(d(x0)) 2 sqrt(d(x0)); the frozen tangent bases and closest-point
 * coordinates are held fixed. lambda, out: [num_bodies][6]. No counterpart
 * in the reference, which has no gradients (python/src/bindings.cpp:30-193):
 * it serves the implicit-function adjoint of a time step. */
rigc_status rigc_friction_adjoint(rigc_ctx* ctx, const double* poses, const double* poses_t0, double epsv_times_h,
                                  double dhat, double dmin, double barrier_stiffness, const double* lambda,
                                  double* out);

/* ---- Batched scenes (rigc_set_scenes). One call serves the n scenes
 * scene_ids[0..n) (strictly ascending) with one launch sequence. Pose
 * arrays hold those scenes' bodies concatenated in list order ([sum nb][6]),
 * and so do grad ([sum nb][6]) and hess->diag ([sum nb][36]); hess->pairs
 * use the same concatenated body numbering. Per-scene results are equal,
 * bit for bit, to those of a context holding that scene alone. */

/* Per-scene rigc_compute_earliest_toi: toi[i] for scene scene_ids[i]. */
rigc_status rigc_batch_earliest_toi(rigc_ctx* ctx, int32_t n, const int32_t* scene_ids, const double* poses_t0,
                                    const double* poses_t1, double minimum_separation_distance, double* toi);
/* DetectionMethod (src/ccd/ccd.hpp:20-24) of rigc_broad_phase and of the
 * earliest-TOI calls' candidates: RIGC_DETECTION_HASH_GRID (default) or
 * RIGC_DETECTION_BVH (the reference's default detection_method,
 * src/SimState.cpp:178; detect_collision_candidates_rigid_bvh,
 * src/ccd/rigid/broad_phase.cpp:92-121, 205-230). The BVH path needs every
 * vertex on an edge and every edge on a face, and at most 64 KiB of
 * per-body staging (48 B per vertex + 8 B per face of the largest body);
 * RIGC_ERR_INVALID_ARGUMENT otherwise. The constraint set is built from
 * the hash grid's candidates either way: every pair within the activation
 * distance is a candidate of both, so the set is the same. A scene set
 * afterwards keeps the method (set it again to re-check the scene). */
#define RIGC_DETECTION_HASH_GRID 0
#define RIGC_DETECTION_BVH 1
rigc_status rigc_set_detection_method(rigc_ctx* ctx, int32_t method);

/* Tunables of the rigid CCD narrow phase. No counterpart in the reference
 * (its CCD is sequential per query); none of them changes a result, only
 * the schedule of the device's search. */
typedef struct rigc_tuning {
    int32_t ccd_walk_budget;   /* boxes a lane walks before its open subtrees go
                                * to the next pass (0: never); default 4096 */
    int32_t ccd_spec_levels;   /* levels below a box that idle lanes evaluate
                                * ahead of its owner (0..5); default 5 */
    double ccd_window_factor;  /* first pass of an earliest-TOI call bounded by
                                * this times the scene's previous answer (0: no
                                * window); default 4 */
    int64_t ccd_wide_queries;  /* calls with at least this many queries launch
                                * two blocks per CU (throughput), fewer one
                                * (latency); default 4096 */
} rigc_tuning;
rigc_status rigc_get_tuning(const rigc_ctx* ctx, rigc_tuning* tuning);
rigc_status rigc_set_tuning(rigc_ctx* ctx, const rigc_tuning* tuning);
/* Autotuner: coordinate descent over the tunings selected by knobs (an OR
 * of RIGC_TUNE_*; 0 means all: speculative levels, walk budget, window
 * factor, wide-launch query count; three values each) on num_samples
 * earliest-TOI calls (poses_t0 / poses_t1: [num_samples][num_bodies][6],
 * num_bodies of the whole context; each call as rigc_compute_earliest_toi,
 * in order, so a line-search-like sequence exercises the window), median of
 * three timed runs per tuning; a knob changes only for a value at least 5%
 * faster. Knobs left out keep the context's setting. The window factor acts
 * through each scene's previous TOI, so a sample tunes it only when it
 * follows the workload's own call sequence. Checks that every tuning
 * returns the same TOIs bit for bit (RIGC_ERR_DEVICE otherwise) and sets
 * the choice on the context. best / best_ms (optional): the tuning chosen
 * and its median device time over the sample. */
#define RIGC_TUNE_SPEC_LEVELS 1
#define RIGC_TUNE_WALK_BUDGET 2
#define RIGC_TUNE_WINDOW_FACTOR 4
#define RIGC_TUNE_WIDE_QUERIES 8
rigc_status rigc_autotune_earliest_toi(rigc_ctx* ctx, int32_t num_samples, const double* poses_t0,
                                       const double* poses_t1, double minimum_separation_distance, int32_t knobs,
                                       rigc_tuning* best, double* best_ms);

/* The previous earliest TOI of each listed scene (>= 0, +inf for none): the
 * bound of the first pass of its next earliest-TOI call on this context.
 * A context keeps it by itself from its own calls; a caller that spreads a
 * scene's calls over several contexts hands it over. Results do not depend
 * on it. No counterpart in the reference (its search has no first pass). */
rigc_status rigc_batch_set_previous_toi(rigc_ctx* ctx, int32_t n, const int32_t* scene_ids, const double* toi);
/* Box evaluations of the rigid CCD per scene in the last earliest-TOI call
 * of a multi-scene context (0 for a scene the call did not serve). No
 * counterpart in the reference: it tells a batch caller which scenes'
 * interval searches made a fused call's work. */
rigc_status rigc_batch_ccd_evaluations(rigc_ctx* ctx, int32_t n, const int32_t* scene_ids, int64_t* evals);
/* Per-scene rigc_compute_barrier_term with the constraint set built at
 * build_poses (DistanceBarrierConstraint::construct_constraint_set) and the
 * potential evaluated at poses; energy[i] per scene; counts (optional,
 * [n][4]) receive each scene's VV / EV / EE / FV constraint counts;
 * min_distance (optional, [n]) receives what rigc_batch_minimum_distance
 * returns for the same arguments, from the same constraint set (a stepper
 * that asks for both at one point saves a call). */
rigc_status rigc_batch_barrier_term(rigc_ctx* ctx, int32_t n, const int32_t* scene_ids, const double* build_poses,
                                    const double* poses, double dhat, double dmin, int32_t flags, double* energy,
                                    double* grad, rigc_block_hessian* hess, int64_t* counts, double* min_distance);
/* Per-scene rigc_detect_intersections at poses: counts[i] intersecting
 * edge-face pairs of scene scene_ids[i]. */
rigc_status rigc_batch_detect_intersections(rigc_ctx* ctx, int32_t n, const int32_t* scene_ids, const double* poses,
                                            int64_t* counts);
/* Per-scene rigc_compute_minimum_distance over the set built at build_poses. */
rigc_status rigc_batch_minimum_distance(rigc_ctx* ctx, int32_t n, const int32_t* scene_ids,
                                        const double* build_poses, const double* poses, double dhat, double dmin,
                                        double* min_distance);
/* Per-scene rigc_update_friction_constraints (constraint set built at
 * poses) with per-scene barrier_stiffness[i] and mu[i]; the other scenes'
 * friction sets are kept. counts (optional, [n]) receive the set sizes. */
rigc_status rigc_batch_update_friction_constraints(rigc_ctx* ctx, int32_t n, const int32_t* scene_ids,
                                                   const double* poses, double dhat, double dmin,
                                                   const double* barrier_stiffness, const double* mu,
                                                   int64_t* counts);
/* Per-scene rigc_compute_friction_term with per-scene epsv_times_h[i]. */
rigc_status rigc_batch_friction_term(rigc_ctx* ctx, int32_t n, const int32_t* scene_ids, const double* poses,
                                     const double* poses_t0, const double* epsv_times_h, int32_t flags,
                                     double* energy, double* grad, rigc_block_hessian* hess);

#ifdef __cplusplus
}
#endif

#endif /* RIGC_H */
