// Host time stepper (rigs.h): DistanceBarrierRBProblem's simulation step,
// the IPC solver's barrier-stiffness logic and the Newton solver, over a
// contact backend.
//
//   simulation_step / update_constraints / solve_constraints / take_step
//                            src/problems/distance_barrier_rb_problem.cpp:118-626
//   compute_objective        :633-703
//   compute_energy_term / compute_body_energy (IMPLICIT_EULER)  :706-944
//   IPCSolver::init_solve / post_step_update  src/solvers/ipc_solver.cpp:35-88
//   NewtonSolver::solve / line_search / compute_regularized_direction /
//   converged                src/solvers/newton_solver.cpp:95-520
//   free_dof                 distance_barrier_rb_problem.cpp:100-110
//   constraint-set cache     src/opt/distance_barrier_constraint.cpp:272-281
// Adaptive stiffness restates ipc-toolkit @d41d6ca's
// initial_barrier_stiffness / update_barrier_stiffness (not in the
// reference tree).
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../csrc/rigc_barrier.hpp"
#include "block_ldlt.hpp"
#include "helper.hpp"
#include "rigs.h"

using rigs::Block;

namespace {

constexpr int OVERLAP_MIN_BODIES = 200;

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ------------------------------------------------------------ device backend
rigc_status gb_set_scene(void* s, const rigc_scene* sc) { return rigc_set_scene((rigc_ctx*)s, sc); }
rigc_status gb_toi(void* s, const double* p0, const double* p1, double msd, double* toi) {
    return rigc_compute_earliest_toi((rigc_ctx*)s, p0, p1, msd, toi);
}
rigc_status gb_build(void* s, const double* p, double dhat, double dmin, int64_t* counts) {
    return rigc_construct_constraint_set((rigc_ctx*)s, p, dhat, dmin, counts);
}
rigc_status gb_barrier(void* s, const double* p, double dhat, double dmin, int32_t flags, double* e, double* g,
                       rigc_block_hessian* h) {
    return rigc_compute_barrier_term((rigc_ctx*)s, p, dhat, dmin, flags, e, g, h);
}
rigc_status gb_mind(void* s, const double* p, double dhat, double dmin, int32_t flags, double* out) {
    return rigc_compute_minimum_distance((rigc_ctx*)s, p, dhat, dmin, flags, out);
}
rigc_status gb_fupdate(void* s, const double* p, double dhat, double dmin, double kappa, double mu, int32_t flags,
                       int64_t* n) {
    return rigc_update_friction_constraints((rigc_ctx*)s, p, dhat, dmin, kappa, mu, flags, n);
}
rigc_status gb_friction(void* s, const double* p, const double* p0, double eh, int32_t flags, double* e, double* g,
                        rigc_block_hessian* h) {
    return rigc_compute_friction_term((rigc_ctx*)s, p, p0, eh, flags, e, g, h);
}
const char* gb_error(void* s) { return rigc_last_error((rigc_ctx*)s); }
rigc_status gb_stats(void* s, rigc_stats* st) { return rigc_get_stats((rigc_ctx*)s, st); }
rigc_status gb_isect(void* s, const double* p, int64_t* n) {
    return rigc_detect_intersections((rigc_ctx*)s, p, nullptr, 0, n);
}
rigc_status gb_fadjoint(void* s, const double* p, const double* p0, double eh, double dhat, double dmin, double kappa,
                        const double* lambda, double* out) {
    return rigc_friction_adjoint((rigc_ctx*)s, p, p0, eh, dhat, dmin, kappa, lambda, out);
}
rigc_status gb_detection(void* s, int32_t method) { return rigc_set_detection_method((rigc_ctx*)s, method); }
rigc_status gb_toi_pl(void* s, const double* p0, const double* p1, double msd, double* toi) {
    return rigc_compute_earliest_toi_piecewise_linear((rigc_ctx*)s, p0, p1, msd, toi);
}

struct BodyState {
    double mass = 0;
    double inertia[3] = { 0, 0, 0 };
    double J[3] = { 0, 0, 0 };      // compute_J: 1/2 (-I0 + I1 + I2), ...
    double Jsqrt[3] = { 0, 0, 0 };  // compute_Jsqrt: sqrt(max(J, 0))
    bool dynamic = false;
    double pose[6], pose_prev[6], velocity[6];
    double Qdot[9];
    double force[6] = { 0, 0, 0, 0, 0, 0 };  // external force and torque (world)
    // inertial targets of the current step: position and rotation matrix,
    // and h^2 Q_t0^T [torque]_x
    double target[3], Pm[9], Tau[9];
    // RigidBodyType::KINEMATIC: scripted poses (front first) and the time
    // after which the body becomes static
    bool kinematic = false;
    double kin_max_time = HUGE_VAL;
    std::deque<std::array<double, 6>> kin_poses;
};


// Symmetric block Hessian in body layout.
struct BlockHess {
    std::vector<Block> diag;
    std::map<std::pair<int, int>, Block> off;  // i < j
    void reset(int nb) {
        diag.assign(nb, Block{});
        off.clear();
    }
};

}  // namespace

struct rigs_sim {
    rigs_backend be{};
    rigc_ctx* own_ctx = nullptr;
    rigs_params P{};
    int nb = 0, nv = 0;
    std::vector<double> V;             // body-local vertices
    std::vector<int> vstart;           // body vertex ranges
    std::vector<BodyState> bodies;
    std::vector<int> dyn_index;        // body -> free block index (-1 fixed)
    int n_free = 0;
    // augmented Lagrangian of the kinematic bodies (distance_barrier_rb_problem.cpp:
    // 190-235 init, 339-435 update, 946-1109 potential); per body, only the
    // kinematic bodies' entries are used
    int n_kin = 0;
    std::vector<double> x_pred;           // [nb][6] the step's target poses
    std::vector<unsigned char> dof_sat;   // [nb][6] is_dof_satisfied: out of the free set
    double al_kq = 1e3, al_kQ = 1e3;      // linear and angular penalties
    std::vector<double> al_lam, al_Lam;   // [nb][3] and [nb][9] multipliers
    std::vector<BodyState> bodies_init;   // kinematic state restored by rigs_set_state
    double bbox_diagonal = 0, average_mass = 0;
    double kappa = 1, max_kappa = 0, prev_min_distance = -1;
    std::vector<double> poses_t0;      // start of the current step
    std::vector<double> cached_poses;  // poses of the backend's constraint set
    bool cache_valid = false;
    // backend output staging (one per term: the friction term may run on a
    // helper thread beside the barrier term)
    struct Staging {
        std::vector<double> hoff;
        std::vector<int32_t> hpairs;
        int64_t hcap = 0;
    };
    Staging stg_barrier, stg_friction;
    // The friction term's backend: with a context of the stepper's own, a
    // second context (its own stream, its own copy of the lagged friction
    // set) so that friction runs beside the barrier term; else the backend.
    rigc_ctx* fric_ctx = nullptr;
    rigs_backend fbe{};
    rigs_stats stats{};
    std::string err;
    int verbose = 0;  // RIGS_VERBOSE=1: one line per Newton iteration on stderr; 2: line-search trials too
    double last_terms[3] = { 0, 0, 0 };  // E / m, kappa / m B, D / m of the last objective

    rigc_status fail(rigc_status s, const std::string& m) {
        err = m;
        return s;
    }
    rigc_status backend_fail(rigc_status s, const char* where) {
        err = std::string(where) + ": " + (be.last_error ? be.last_error(be.self) : "backend error");
        return s;
    }

    // ------------------------------------------------------------ contact calls
    rigc_status ensure_constraints(const double* poses) {
        if (cache_valid && std::memcmp(cached_poses.data(), poses, sizeof(double) * 6 * nb) == 0) return RIGC_OK;
        int64_t counts[4];
        const double t0 = now_ms();
        rigc_status s = be.construct_constraint_set(be.self, poses, P.dhat, P.minimum_separation_distance, counts);
        const double dt = now_ms() - t0;
        stats.contact_ms += dt;
        stats.constraint_ms += dt;
        if (s != RIGC_OK) return backend_fail(s, "construct_constraint_set");
        stats.constraint_builds++;
        stats.num_constraints = counts[0] + counts[1] + counts[2] + counts[3];
        cached_poses.assign(poses, poses + 6 * nb);
        cache_valid = true;
        return RIGC_OK;
    }

    // Calls a potential entry point, growing the off-diagonal capacity once;
    // the time spent in the call is added to `spent`. Touches nothing shared
    // but `S` (the friction term runs on a helper thread).
    template <typename F>
    rigc_status potential(F&& call, Staging& S, int32_t flags, double& e, std::vector<double>* g, BlockHess* H,
                          double& spent) {
        if (g) g->assign(6 * nb, 0.0);
        rigc_block_hessian h{};
        if (H) {
            // the backend writes every diagonal block straight into H
            H->diag.resize(nb);
            H->off.clear();
        }
        for (int attempt = 0; attempt < 2; attempt++) {
            if (H) {
                S.hpairs.resize(2 * (size_t)std::max<int64_t>(S.hcap, 1));
                S.hoff.resize(36 * (size_t)std::max<int64_t>(S.hcap, 1));
                h.diag = H->diag[0].data();
                h.pairs = S.hpairs.data();
                h.off = S.hoff.data();
                h.capacity = S.hcap;
            }
            const double t0 = now_ms();
            rigc_status s = call(flags, &e, g ? g->data() : nullptr, H ? &h : nullptr);
            spent += now_ms() - t0;
            if (s == RIGC_ERR_CAPACITY && H && h.count > S.hcap) {
                S.hcap = h.count;
                continue;
            }
            if (s != RIGC_OK) return s;
            break;
        }
        if (H) {
            for (int64_t k = 0; k < h.count; k++) {
                Block blk;
                std::copy(&S.hoff[36 * k], &S.hoff[36 * k] + 36, blk.begin());
                H->off[{ S.hpairs[2 * k], S.hpairs[2 * k + 1] }] = blk;
            }
        }
        return RIGC_OK;
    }

    // A device term at the first trial point of a line search, evaluated
    // with its gradient and Hessian (prefetch): when the trial is accepted
    // the next Newton iteration starts there and takes them from here
    // instead of making a second call. Entries are keyed by the poses; the
    // barrier term is a function of the poses alone (its constraint set is
    // built at them), the friction term also of the lagged friction set,
    // so the friction entry is dropped when that set changes.
    struct TermCache {
        bool valid = false;
        bool derivs = false;  // g and H still held (they are handed out once, by move)
        std::vector<double> x;
        double e = 0;
        std::vector<double> g;
        BlockHess H;
    };
    TermCache bcache, fcache;

    template <typename F>
    rigc_status cached_term(TermCache& c, F&& eval, const double* x, int32_t flags, bool prefetch, double& e,
                            std::vector<double>* g, BlockHess* H) {
        const int32_t full = RIGC_GRADIENT | RIGC_HESSIAN;
        // an energy-only request is served by a full entry: the energy is
        // the same sum of the same per-constraint values
        // the Newton iteration that starts at the entry's point is the only
        // user of its gradient and Hessian: they are moved out, not copied
        if (c.valid && (flags == 0 || (flags == full && c.derivs)) &&
            std::memcmp(c.x.data(), x, sizeof(double) * 6 * (size_t)nb) == 0) {
            e = c.e;
            if (flags == full) {
                if (g) *g = std::move(c.g);
                if (H) *H = std::move(c.H);
                c.derivs = false;
            }
            return RIGC_OK;
        }
        if (prefetch && flags == 0) {
            c.valid = false;
            rigc_status s = eval(full, c.e, &c.g, &c.H);
            if (s != RIGC_OK) return s;
            c.x.assign(x, x + 6 * (size_t)nb);
            c.valid = true;
            c.derivs = true;
            e = c.e;
            return RIGC_OK;
        }
        return eval(flags, e, g, H);
    }

    rigc_status barrier_term(const double* x, int32_t flags, double& e, std::vector<double>* g, BlockHess* H,
                             bool prefetch = false) {
        rigc_status s = ensure_constraints(x);
        if (s != RIGC_OK) return s;
        auto call = [&](int32_t f, double* pe, double* pg, rigc_block_hessian* ph) {
            return be.compute_barrier_term(be.self, x, P.dhat, P.minimum_separation_distance, f, pe, pg, ph);
        };
        double spent = 0;
        auto eval = [&](int32_t f, double& pe, std::vector<double>* pg, BlockHess* pH) {
            return potential(call, stg_barrier, f, pe, pg, pH, spent);
        };
        s = cached_term(bcache, eval, x, flags, prefetch, e, g, H);
        stats.contact_ms += spent;
        stats.barrier_ms += spent;
        return s == RIGC_OK ? s : backend_fail(s, "compute_barrier_term");
    }

    // The friction term through fbe; touches only fcache, stg_friction, the
    // outputs and `spent` / `msg`, so it may run on a helper thread.
    rigc_status friction_eval(const double* x, int32_t flags, double& e, std::vector<double>* g, BlockHess* H,
                              bool prefetch, double& spent, std::string& msg) {
        auto call = [&](int32_t f, double* pe, double* pg, rigc_block_hessian* ph) {
            return fbe.compute_friction_term(fbe.self, x, poses_t0.data(),
                                             P.static_friction_speed_bound * P.timestep, f, pe, pg, ph);
        };
        auto eval = [&](int32_t f, double& pe, std::vector<double>* pg, BlockHess* pH) {
            return potential(call, stg_friction, f, pe, pg, pH, spent);
        };
        rigc_status s = cached_term(fcache, eval, x, flags, prefetch, e, g, H);
        if (s != RIGC_OK)
            msg = std::string("compute_friction_term: ") + (fbe.last_error ? fbe.last_error(fbe.self) : "backend error");
        return s;
    }

    rigc_status friction_term(const double* x, int32_t flags, double& e, std::vector<double>* g, BlockHess* H,
                              bool prefetch = false) {
        double spent = 0;
        std::string msg;
        rigc_status s = friction_eval(x, flags, e, g, H, prefetch, spent, msg);
        stats.contact_ms += spent;
        stats.friction_ms += spent;
        if (s != RIGC_OK) err = msg;
        return s;
    }

    rigc_status update_friction(const double* poses) {
        if (P.coefficient_friction <= 0) return RIGC_OK;
        fcache.valid = false;
        int64_t n = 0;
        rigc_status s = RIGC_OK;
        const double t0 = now_ms();
        if (fric_ctx) {
            // the friction context builds its own constraint set at poses
            s = fbe.update_friction_constraints(fbe.self, poses, P.dhat, P.minimum_separation_distance, kappa,
                                                P.coefficient_friction, RIGC_BUILD_CONSTRAINTS, &n);
        } else {
            s = ensure_constraints(poses);
            if (s != RIGC_OK) return s;
            s = be.update_friction_constraints(be.self, poses, P.dhat, P.minimum_separation_distance, kappa,
                                               P.coefficient_friction, 0, &n);
        }
        const double dt = now_ms() - t0;
        stats.contact_ms += dt;
        stats.friction_ms += dt;
        if (s != RIGC_OK) {
            err = std::string("update_friction_constraints: ") +
                  (fbe.last_error ? fbe.last_error(fbe.self) : "backend error");
            return s;
        }
        return RIGC_OK;
    }

    // compute_min_distance: sqrt of the minimum, -1 without constraints
    rigc_status min_distance(const double* x, double& out) {
        rigc_status s = ensure_constraints(x);
        if (s != RIGC_OK) return s;
        const double t0 = now_ms();
        double m = 0;
        s = be.compute_minimum_distance(be.self, x, P.dhat, P.minimum_separation_distance, 0, &m);
        const double dt = now_ms() - t0;
        stats.contact_ms += dt;
        stats.distance_ms += dt;
        if (s != RIGC_OK) return backend_fail(s, "compute_minimum_distance");
        out = std::isfinite(m) ? m : -1;
        return RIGC_OK;
    }

    // RigidBodyProblem::detect_intersections through the backend (n = -1
    // when the backend has no such entry point).
    rigc_status intersections(const double* x, int64_t& n) {
        n = -1;
        if (!be.detect_intersections) return RIGC_OK;
        const double t0 = now_ms();
        rigc_status s = be.detect_intersections(be.self, x, &n);
        stats.contact_ms += now_ms() - t0;
        if (s != RIGC_OK) return backend_fail(s, "detect_intersections");
        return RIGC_OK;
    }

    rigc_status earliest_toi(const double* x0, const double* x1, double& toi) {
        const double t0 = now_ms();
        // compute_earliest_toi dispatches on the trajectory type (ccd.cpp:271-279)
        auto fn = P.trajectory_type == RIGS_TRAJECTORY_PIECEWISE_LINEAR ? be.compute_earliest_toi_piecewise_linear
                                                                         : be.compute_earliest_toi;
        rigc_status s = fn(be.self, x0, x1, P.minimum_separation_distance, &toi);
        const double dt = now_ms() - t0;
        stats.contact_ms += dt;
        stats.ccd_ms += dt;
        stats.ccd_calls++;
        if (s != RIGC_OK) return backend_fail(s, "compute_earliest_toi");
        if (be.get_stats) {
            rigc_stats cs{};
            if (be.get_stats(be.self, &cs) == RIGC_OK) {
                stats.ccd_candidates += cs.num_candidates;
                stats.ccd_broad_phase_ms += cs.broad_phase_ms;
                stats.ccd_narrow_phase_ms += cs.narrow_phase_ms;
                stats.ccd_box_evaluations += cs.box_evaluations;
                stats.ccd_kernel_ms += cs.ccd_kernel_ms;
                stats.ccd_launches += cs.ccd_launches;
            }
        }
        return RIGC_OK;
    }

    // ------------------------------------------------------------ body energy
    // compute_body_energy<T> for IMPLICIT_EULER without external forces:
    //   1/2 m q.q - m q.(q_t0 + h (qdot_t0 + h g))
    //   + 1/2 tr(Q J Q^T) - tr(Q J (Q_t0 + h Qdot_t0)^T)
    // Inertial targets of the step starting at the bodies' current state.
    // With an external force f and torque tau (compute_body_energy,
    // :859-904): the target adds h^2 f / m, and the energy h^2 tr(Q Tau)
    // with Tau = Q_t0^T [tau]_x.
    void step_targets() {
        const double h = P.timestep;
        for (auto& B : bodies) {
            for (int k = 0; k < 3; k++) {
                const double acc = B.mass > 0 ? P.gravity[k] + B.force[k] / B.mass : P.gravity[k];
                B.target[k] = B.pose[k] + h * (B.velocity[k] + h * acc);
            }
            double Q0[9];
            rigc::rotation_d(B.pose + 3, Q0);
            for (int e = 0; e < 9; e++) B.Pm[e] = Q0[e] + h * B.Qdot[e];
            const double* t = B.force + 3;
            const double K[9] = { 0, -t[2], t[1], t[2], 0, -t[0], -t[1], t[0], 0 };
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    double v = 0;
                    for (int k = 0; k < 3; k++) v += Q0[3 * k + i] * K[3 * k + j];
                    B.Tau[3 * i + j] = h * h * v;
                }
        }
    }

    double energy_term(const double* x, std::vector<double>* g, BlockHess* H) const {
        double E = 0;
        if (g) g->assign(6 * nb, 0.0);
        if (H) H->reset(nb);
        for (int b = 0; b < nb; b++) {
            const BodyState& B = bodies[b];
            if (!B.dynamic) continue;
            const double* q = x + 6 * b;
            const double* target = B.target;
            double qq = 0, qt = 0;
            for (int k = 0; k < 3; k++) {
                qq += q[k] * q[k];
                qt += q[k] * target[k];
            }
            E += 0.5 * B.mass * qq - B.mass * qt;
            double Q[9];
            const double* Pm = B.Pm;
            rigc::rotation_d(q + 3, Q);
            double tQJQ = 0, tQJP = 0;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    tQJQ += Q[3 * i + j] * B.J[j] * Q[3 * i + j];
                    tQJP += Q[3 * i + j] * B.J[j] * Pm[3 * i + j];
                }
            E += 0.5 * tQJQ - tQJP;
            const bool torque = B.force[3] != 0 || B.force[4] != 0 || B.force[5] != 0;
            if (torque) {
                double tQT = 0;  // tr(Q Tau)
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) tQT += Q[3 * i + j] * B.Tau[3 * j + i];
                E += tQT;
            }
            if (!g && !H) continue;
            rigc::BodyFrame F;
            rigc::body_frame(q + 3, F);
            // dE/dQ_ij = J_j (Q_ij - P_ij) (+ Tau_ji)
            double G[9];
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    G[3 * i + j] = B.J[j] * (Q[3 * i + j] - Pm[3 * i + j]);
                    if (torque) G[3 * i + j] += B.Tau[3 * j + i];
                }
            if (g) {
                for (int k = 0; k < 3; k++) (*g)[6 * b + k] = B.mass * q[k] - B.mass * target[k];
                for (int c = 0; c < 3; c++) {
                    double s = 0;
                    for (int e = 0; e < 9; e++) s += F.dR[c][e] * G[e];
                    (*g)[6 * b + 3 + c] = s;
                }
            }
            if (H) {
                Block& D = H->diag[b];
                for (int k = 0; k < 3; k++) D[6 * k + k] = B.mass;
                for (int c = 0; c < 3; c++)
                    for (int d = 0; d < 3; d++) {
                        const double* d2 = F.d2R[rigc::sym_index(c, d)];
                        double s = 0;
                        for (int i = 0; i < 3; i++)
                            for (int j = 0; j < 3; j++)
                                s += d2[3 * i + j] * G[3 * i + j] + F.dR[c][3 * i + j] * B.J[j] * F.dR[d][3 * i + j];
                        D[6 * (3 + c) + 3 + d] = s;
                    }
            }
        }
        return E;
    }

    // ------------------------------------------------------------ augmented Lagrangian
    bool dof_free(int b, int k) const { return dyn_index[b] >= 0 && !dof_sat[6 * b + k]; }

    // init_augmented_lagrangian (:190-235), at the start of a step: kinematic
    // bodies whose time ran out become static; the targets are the next
    // scripted pose, else the pose moved by the velocity over a timestep.
    void init_al() {
        al_kq = al_kQ = 1e3;
        std::fill(al_lam.begin(), al_lam.end(), 0.0);
        std::fill(al_Lam.begin(), al_Lam.end(), 0.0);
        for (auto& B : bodies)
            if (B.kinematic && B.kin_max_time < 0) convert_to_static(B);
        const double h = P.timestep;
        for (int b = 0; b < nb; b++) {
            const BodyState& B = bodies[b];
            for (int k = 0; k < 6; k++)
                x_pred[6 * b + k] = B.kin_poses.empty() ? B.pose[k] + h * B.velocity[k] : B.kin_poses.front()[k];
            // is_dof_satisfied: the static bodies (free blocks of bodies that
            // became static stay, with every dof out of the free set)
            const bool fixed = !B.dynamic && !B.kinematic;
            for (int k = 0; k < 6; k++) dof_sat[6 * b + k] = fixed ? 1 : 0;
        }
    }
    // RigidBody::convert_to_static: fixed, at rest
    static void convert_to_static(BodyState& B) {
        B.kinematic = false;
        B.dynamic = false;
        for (int k = 0; k < 6; k++) B.velocity[k] = 0, B.force[k] = 0;
        for (int e = 0; e < 9; e++) B.Qdot[e] = 0;
    }
    // step_kinematic_bodies (:237-251), after a step
    void step_kinematic() {
        for (auto& B : bodies) {
            if (!B.kinematic) continue;
            if (B.kin_max_time < 0) {
                convert_to_static(B);
            } else {
                B.kin_max_time -= P.timestep;
                if (!B.kin_poses.empty()) B.kin_poses.pop_front();
            }
        }
    }
    // compute_{linear,angular}_augment_lagrangian_progress (:253-337)
    void al_progress(const double* x, double& eta_q, double& eta_Q) const {
        double aq = 0, bq = 0, aQ = 0, bQ = 0;
        for (int b = 0; b < nb; b++) {
            if (!bodies[b].kinematic) continue;
            const double* xp = &x_pred[6 * b];
            for (int k = 0; k < 3; k++) {
                aq += (xp[k] - x[6 * b + k]) * (xp[k] - x[6 * b + k]);
                bq += (xp[k] - poses_t0[6 * b + k]) * (xp[k] - poses_t0[6 * b + k]);
            }
            double Qp[9], Q[9], Q0[9];
            rigc::rotation_d(xp + 3, Qp);
            rigc::rotation_d(x + 6 * b + 3, Q);
            rigc::rotation_d(&poses_t0[6 * b + 3], Q0);
            for (int e = 0; e < 9; e++) {
                aQ += (Q[e] - Qp[e]) * (Q[e] - Qp[e]);
                bQ += (Q0[e] - Qp[e]) * (Q0[e] - Qp[e]);
            }
        }
        eta_q = (aq == 0 && bq == 0) ? 1.0 : 1 - std::sqrt(aq / (bq == 0 ? 1 : bq));
        eta_Q = (aQ == 0 && bQ == 0) ? 1.0 : 1 - std::sqrt(aQ / (bQ == 0 ? 1 : bQ));
    }
    // are_equality_constraints_satisfied (:425-433)
    bool al_satisfied(const double* x) const {
        if (!n_kin) return true;
        double eq, eQ;
        al_progress(x, eq, eQ);
        return eq >= 0.999 && eQ >= 0.999;
    }
    // update_augmented_lagrangian (:339-423)
    void al_update(const double* x) {
        double eq, eQ;
        al_progress(x, eq, eQ);
        if (eq >= 0.999) {
            for (int b = 0; b < nb; b++)
                if (bodies[b].kinematic)
                    for (int k = 0; k < 3; k++) dof_sat[6 * b + k] = 1;
        } else if (eq < 0.99 && al_kq < 1e8) {
            al_kq *= 2;
        } else {
            for (int b = 0; b < nb; b++) {
                if (!bodies[b].kinematic) continue;
                const double sm = std::sqrt(bodies[b].mass);
                for (int k = 0; k < 3; k++) al_lam[3 * b + k] -= al_kq * sm * (x[6 * b + k] - x_pred[6 * b + k]);
            }
        }
        if (eQ >= 0.999) {
            for (int b = 0; b < nb; b++)
                if (bodies[b].kinematic)
                    for (int k = 3; k < 6; k++) dof_sat[6 * b + k] = 1;
        } else if (eQ < 0.99 && al_kQ < 1e8) {
            al_kQ *= 2;
        } else {
            for (int b = 0; b < nb; b++) {
                const BodyState& B = bodies[b];
                if (!B.kinematic) continue;
                double Qp[9], Q[9];
                rigc::rotation_d(&x_pred[6 * b + 3], Qp);
                rigc::rotation_d(x + 6 * b + 3, Q);
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++)
                        al_Lam[9 * b + 3 * i + j] -= al_kQ * (Q[3 * i + j] - Qp[3 * i + j]) * B.Jsqrt[j];
            }
        }
        if (verbose)
            std::fprintf(stderr, "rigs: augmented Lagrangian kappa_q=%g kappa_Q=%g eta_q=%g eta_Q=%g\n", al_kq, al_kQ,
                         eq, eQ);
    }
    // compute_augmented_lagrangian (:946-1109): for each kinematic body
    //   kq / 2 m |q - q_pred|^2 - sqrt(m) lam . (q - q_pred)
    //   + kQ / 2 tr((Q - Q_pred) J (Q - Q_pred)^T) - tr(Lam^T (Q - Q_pred) Jsqrt)
    // zero when every kinematic dof is satisfied. Gradient and Hessian are
    // added into g and H (which hold other terms already).
    double al_term(const double* x, std::vector<double>* g, BlockHess* H) const {
        bool all_sat = true;
        for (int b = 0; b < nb && all_sat; b++)
            if (bodies[b].kinematic)
                for (int k = 0; k < 6; k++) all_sat = all_sat && dof_sat[6 * b + k];
        if (all_sat) return 0.0;
        double A = 0;
        for (int b = 0; b < nb; b++) {
            const BodyState& B = bodies[b];
            if (!B.kinematic) continue;
            const double* q = x + 6 * b;
            const double* qp = &x_pred[6 * b];
            const double m = B.mass, sm = std::sqrt(m);
            double dd = 0, ld = 0;
            for (int k = 0; k < 3; k++) {
                dd += (q[k] - qp[k]) * (q[k] - qp[k]);
                ld += al_lam[3 * b + k] * (q[k] - qp[k]);
            }
            A += al_kq / 2 * m * dd - sm * ld;
            if (g)
                for (int k = 0; k < 3; k++) (*g)[6 * b + k] += al_kq * m * (q[k] - qp[k]) - sm * al_lam[3 * b + k];
            if (H)
                for (int k = 0; k < 3; k++) H->diag[b][7 * k] += al_kq * m;
            double Q[9], Qp[9];
            rigc::rotation_d(q + 3, Q);
            rigc::rotation_d(qp + 3, Qp);
            double tDJD = 0, tLDJ = 0;
            double G[9];  // d/dQ_ij = kQ J_j D_ij - Lam_ij Jsqrt_j
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) {
                    const double D = Q[3 * i + j] - Qp[3 * i + j], L = al_Lam[9 * b + 3 * i + j];
                    tDJD += D * B.J[j] * D;
                    tLDJ += L * D * B.Jsqrt[j];
                    G[3 * i + j] = al_kQ * B.J[j] * D - L * B.Jsqrt[j];
                }
            A += al_kQ / 2 * tDJD - tLDJ;
            if (!g && !H) continue;
            rigc::BodyFrame F;
            rigc::body_frame(q + 3, F);
            if (g)
                for (int c = 0; c < 3; c++) {
                    double sum = 0;
                    for (int e = 0; e < 9; e++) sum += F.dR[c][e] * G[e];
                    (*g)[6 * b + 3 + c] += sum;
                }
            if (H)
                for (int c = 0; c < 3; c++)
                    for (int d = 0; d < 3; d++) {
                        const double* d2 = F.d2R[rigc::sym_index(c, d)];
                        double sum = 0;
                        for (int i = 0; i < 3; i++)
                            for (int j = 0; j < 3; j++)
                                sum += d2[3 * i + j] * G[3 * i + j] +
                                       al_kQ * F.dR[c][3 * i + j] * B.J[j] * F.dR[d][3 * i + j];
                        H->diag[b][6 * (3 + c) + 3 + d] += sum;
                    }
        }
        return A;
    }

    // ------------------------------------------------------------ objective
    // f = E / m + AL / m + kappa / m B + D / m (compute_objective :633-703)
    // The body energy at a line search's first trial point, with gradient
    // and Hessian: computed on the helper thread (when `overlap`) while the
    // device terms are evaluated, and taken by the next Newton iteration
    // when that trial is accepted. Its energy is the same value an
    // energy-only call gives.
    struct EnergyCache {
        bool valid = false;
        bool derivs = false;  // g and H still held (handed out once, by move)
        std::vector<double> x;
        double E = 0;
        std::vector<double> g;
        BlockHess H;
    };
    EnergyCache ecache;
    rigs::Helper helper, fric_helper;
    bool overlap = true;
    // A backend that serves several steppers' calls in rounds (the batch
    // stepper) counts the threads that may call it: the friction helper
    // joins (fork_join) before its call and parts (fork_part) after it, and
    // the stepper's own thread parts while it waits for the helper.
    void (*fork_join)(void*) = nullptr;
    void (*fork_part)(void*) = nullptr;
    void* fork_arg = nullptr;

    rigc_status objective(const double* x, bool want_grad, bool want_hess, double& f, std::vector<double>& g,
                          BlockHess& H, bool prefetch = false) {
        stats.objective_evaluations++;
        const double t0 = now_ms();
        std::vector<double> gB, gD;
        BlockHess HB, HD;
        double E = 0;
        bool pending = false;
        if (ecache.valid && want_grad == want_hess && (!want_grad || ecache.derivs) &&
            std::memcmp(ecache.x.data(), x, sizeof(double) * 6 * (size_t)nb) == 0) {
            E = ecache.E;
            if (want_grad) {
                g = std::move(ecache.g);
                H = std::move(ecache.H);
                ecache.derivs = false;
            }
        } else if (prefetch && !want_grad && !want_hess) {
            ecache.valid = false;
            ecache.x.assign(x, x + 6 * (size_t)nb);
            auto task = [this] { ecache.E = energy_term(ecache.x.data(), &ecache.g, &ecache.H); };
            if (overlap) {
                helper.run(task);
                pending = true;
            } else {
                task();
                ecache.valid = ecache.derivs = true;
                E = ecache.E;
            }
        } else {
            E = energy_term(x, want_grad ? &g : nullptr, want_hess ? &H : nullptr);
        }
        stats.host_ms += now_ms() - t0;
        const int32_t flags = (want_grad ? RIGC_GRADIENT : 0) | (want_hess ? RIGC_HESSIAN : 0);
        double B = 0, D = 0;
        // with a friction context of its own, the trial point's friction
        // term runs on a helper thread beside the constraint set and
        // barrier term
        const bool fric = P.coefficient_friction > 0;
        const bool fric_beside = fric && prefetch && ((overlap && fric_ctx) || fork_join);
        rigc_status fs = RIGC_OK;
        double fspent = 0;
        std::string fmsg;
        if (fric_beside) {
            if (fork_join) fork_join(fork_arg);
            fric_helper.run([&] {
                fs = friction_eval(x, flags, D, want_grad ? &gD : nullptr, want_hess ? &HD : nullptr, prefetch,
                                   fspent, fmsg);
                if (fork_part) fork_part(fork_arg);
            });
        }
        rigc_status s = barrier_term(x, flags, B, want_grad ? &gB : nullptr, want_hess ? &HB : nullptr, prefetch);
        if (fric_beside) {
            if (fork_part) fork_part(fork_arg);
            fric_helper.wait();
            if (fork_join) fork_join(fork_arg);
            stats.contact_ms += fspent;
            stats.friction_ms += fspent;
            if (s == RIGC_OK && fs != RIGC_OK) {
                s = fs;
                err = fmsg;
            }
        } else if (s == RIGC_OK && fric) {
            s = friction_term(x, flags, D, want_grad ? &gD : nullptr, want_hess ? &HD : nullptr, prefetch);
        }
        if (pending) {
            const double tw = now_ms();
            helper.wait();
            stats.host_ms += now_ms() - tw;
            ecache.valid = ecache.derivs = true;
            E = ecache.E;
        }
        if (s != RIGC_OK) return s;
        const double t1 = now_ms();
        // the augmented Lagrangian (kinematic bodies) joins the body energy
        if (n_kin) E += al_term(x, want_grad ? &g : nullptr, want_hess ? &H : nullptr);
        const double im = 1.0 / average_mass, km = kappa / average_mass;
        f = E * im + km * B + D * im;
        last_terms[0] = E * im;
        last_terms[1] = km * B;
        last_terms[2] = D * im;
        if (want_grad)
            for (int i = 0; i < 6 * nb; i++)
                g[i] = g[i] * im + km * gB[i] + (P.coefficient_friction > 0 ? gD[i] * im : 0.0);
        if (want_hess) {
            for (int b = 0; b < nb; b++)
                for (int e = 0; e < 36; e++) {
                    double v = H.diag[b][e] * im + km * HB.diag[b][e];
                    if (P.coefficient_friction > 0) v += HD.diag[b][e] * im;
                    H.diag[b][e] = v;
                }
            for (auto& kv : HB.off) {
                Block& o = H.off[kv.first];
                for (int e = 0; e < 36; e++) o[e] += km * kv.second[e];
            }
            if (P.coefficient_friction > 0)
                for (auto& kv : HD.off) {
                    Block& o = H.off[kv.first];
                    for (int e = 0; e < 36; e++) o[e] += im * kv.second[e];
                }
        }
        stats.host_ms += now_ms() - t1;
        return RIGC_OK;
    }

    // ------------------------------------------------------------ Newton
    void world_vertices(const double* x, std::vector<double>& W) const {
        W.resize(3 * (size_t)nv);
        for (int b = 0; b < nb; b++) {
            double R[9];
            rigc::rotation_d(x + 6 * b + 3, R);
            for (int i = vstart[b]; i < vstart[b + 1]; i++) rigc::world_vertex_d(R, x + 6 * b, &V[3 * i], &W[3 * i]);
        }
    }

    // H_free d = -g_free (compute_direction with SimplicialLDLT)
    // A dof out of the free set (is_dof_satisfied: a kinematic body's dofs
    // once its augmented Lagrangian has converged) keeps its row and column
    // out of the system: both are zeroed and its diagonal is 1 with a zero
    // right-hand side, so its direction is 0 and the other rows are those
    // of the reduced system.
    bool solve_direction(const std::vector<double>& g, const BlockHess& H, double coeff, std::vector<double>& dir) {
        std::vector<Block> diag(n_free);
        std::vector<std::pair<int, int>> pairs;
        std::vector<Block> off;
        bool masked = false;
        for (int b = 0; b < nb; b++)
            if (dyn_index[b] >= 0) {
                Block& D = diag[dyn_index[b]];
                D = H.diag[b];
                for (int k = 0; k < 6; k++) D[7 * k] += coeff;
                for (int k = 0; k < 6; k++)
                    if (dof_sat[6 * b + k]) {
                        masked = true;
                        for (int e = 0; e < 6; e++) D[6 * k + e] = D[6 * e + k] = 0.0;
                        D[7 * k] = 1.0;
                    }
            }
        for (const auto& kv : H.off) {
            const int i = dyn_index[kv.first.first], j = dyn_index[kv.first.second];
            if (i < 0 || j < 0) continue;
            pairs.push_back({ i, j });
            off.push_back(kv.second);
            if (masked) {
                Block& O = off.back();
                for (int k = 0; k < 6; k++) {
                    if (dof_sat[6 * kv.first.first + k])
                        for (int e = 0; e < 6; e++) O[6 * k + e] = 0.0;
                    if (dof_sat[6 * kv.first.second + k])
                        for (int e = 0; e < 6; e++) O[6 * e + k] = 0.0;
                }
            }
        }
        std::vector<double> rhs(6 * (size_t)n_free), d(6 * (size_t)n_free);
        for (int b = 0; b < nb; b++)
            if (dyn_index[b] >= 0)
                for (int k = 0; k < 6; k++) rhs[6 * dyn_index[b] + k] = dof_sat[6 * b + k] ? 0.0 : -g[6 * b + k];
        dir.assign(6 * nb, 0.0);
        rigs::BlockLDLT ldlt;
        std::vector<Block> diag_copy;
        if (masked) diag_copy = diag;
        bool ok = ldlt.factorize(n_free, std::move(diag), pairs, off);
        if (ok) {
            ldlt.solve(rhs.data(), d.data());
            // residual must be finite (compute_direction :540-556); the
            // diagonal blocks are H's plus coeff on the diagonal
            double res = 0;
            std::vector<double> Hd(6 * (size_t)n_free, 0.0);
            for (int b = 0; b < nb; b++) {
                const int i = dyn_index[b];
                if (i < 0) continue;
                for (int r = 0; r < 6; r++)
                    for (int c = 0; c < 6; c++) {
                        const double hv = masked ? diag_copy[i][6 * r + c]
                                                 : (r == c ? H.diag[b][6 * r + c] + coeff : H.diag[b][6 * r + c]);
                        Hd[6 * i + r] += hv * d[6 * i + c];
                    }
            }
            for (size_t e = 0; e < pairs.size(); e++) {
                const int i = pairs[e].first, j = pairs[e].second;
                for (int r = 0; r < 6; r++)
                    for (int c = 0; c < 6; c++) {
                        Hd[6 * i + r] += off[e][6 * r + c] * d[6 * j + c];
                        Hd[6 * j + c] += off[e][6 * r + c] * d[6 * i + r];
                    }
            }
            for (int i = 0; i < 6 * n_free; i++) res += (Hd[i] - rhs[i]) * (Hd[i] - rhs[i]);
            ok = std::isfinite(res);
        }
        if (ok) {
            for (int b = 0; b < nb; b++)
                if (dyn_index[b] >= 0)
                    for (int k = 0; k < 6; k++) dir[6 * b + k] = d[6 * dyn_index[b] + k];
        } else {
            for (int b = 0; b < nb; b++)
                for (int k = 0; k < 6; k++)
                    if (dof_free(b, k)) dir[6 * b + k] = -g[6 * b + k];
        }
        return ok;
    }

    double dot_free(const std::vector<double>& a, const std::vector<double>& b) const {
        double s = 0;
        for (int i = 0; i < nb; i++)
            if (dyn_index[i] >= 0)
                for (int k = 0; k < 6; k++)
                    if (!dof_sat[6 * i + k]) s += a[6 * i + k] * b[6 * i + k];
        return s;
    }

    // NewtonSolver::converged, VELOCITY criterion (newton_solver.cpp:95-128):
    // max |V(x + dir) - V(x)|_inf / h <= tol. Evaluated body by body with an
    // early exit; bodies that do not move contribute zero.
    bool converged(const std::vector<double>& x, const std::vector<double>& dir) {
        const double limit = P.velocity_conv_tol * bbox_diagonal;
        for (int b = 0; b < nb; b++) {
            const double* d = &dir[6 * b];
            if (d[0] == 0 && d[1] == 0 && d[2] == 0 && d[3] == 0 && d[4] == 0 && d[5] == 0) continue;
            double x1[6], R0[9], R1[9];
            for (int k = 0; k < 6; k++) x1[k] = x[6 * b + k] + d[k];
            rigc::rotation_d(&x[6 * b + 3], R0);
            rigc::rotation_d(x1 + 3, R1);
            for (int i = vstart[b]; i < vstart[b + 1]; i++) {
                double w0[3], w1[3];
                rigc::world_vertex_d(R0, &x[6 * b], &V[3 * i], w0);
                rigc::world_vertex_d(R1, x1, &V[3 * i], w1);
                for (int k = 0; k < 3; k++)
                    if (!(std::fabs(w1[k] - w0[k]) / P.timestep <= limit)) return false;
            }
        }
        return true;
    }

    // NewtonSolver::line_search (:262-360)
    rigc_status line_search(const std::vector<double>& x, const std::vector<double>& dir, double fx,
                            const std::vector<double>& g, double& step, bool& success) {
        success = false;
        double lower = P.line_search_lower_bound / std::sqrt(std::fabs(dot_free(g, dir)));
        lower = std::min(lower, 1e-1);
        std::vector<double> x1(x);
        for (size_t i = 0; i < x.size(); i++) x1[i] += dir[i];
        double toi = 0;
        rigc_status s = earliest_toi(x.data(), x1.data(), toi);
        if (s != RIGC_OK) return s;
        step = std::min(step, std::min(toi, 1.0));
        std::vector<double> xi(x.size()), gdummy;
        BlockHess Hdummy;
        bool first = true;
        while (std::isfinite(lower) && step >= lower) {
            stats.line_search_iterations++;
            for (size_t i = 0; i < x.size(); i++) xi[i] = x[i] + step * dir[i];
            double fxi = 0;
            // the first trial is usually accepted: its device terms come
            // with the gradient and Hessian the next iteration needs
            s = objective(xi.data(), false, false, fxi, gdummy, Hdummy, first);
            first = false;
            if (s != RIGC_OK) return s;
            if (verbose >= 2)
                std::fprintf(stderr, "rigs:   alpha=%g f=%.17g (E %.17g, B %.17g, D %.17g)\n", step, fxi,
                             last_terms[0], last_terms[1], last_terms[2]);
            if (fxi < fx) {
                success = true;
                break;
            }
            step /= 2.0;
        }
        return RIGC_OK;
    }

    // IPCSolver::init_solve: kappa from the energy / barrier gradients
    // (ipc::initial_barrier_stiffness), previous minimum distance.
    rigc_status init_solve(const std::vector<double>& x0) {
        std::vector<double> gE, gB;
        energy_term(x0.data(), &gE, nullptr);
        double B = 0;
        rigc_status s = barrier_term(x0.data(), RIGC_GRADIENT, B, &gB, nullptr);
        if (s != RIGC_OK) return s;
        const double dhat = P.dhat;
        double d0 = 1e-8 * bbox_diagonal;
        d0 *= d0;
        if (d0 >= dhat * dhat) d0 = 0.5 * dhat * dhat;
        double min_kappa = 4 * d0 * rigc::ipc_barrier_hessian(d0, dhat * dhat);
        min_kappa = P.min_barrier_stiffness_scale * average_mass / min_kappa;
        max_kappa = 100 * min_kappa;
        double k = 1.0, gg = 0, ge = 0;
        for (int i = 0; i < 6 * nb; i++) {
            gg += gB[i] * gB[i];
            ge += gB[i] * gE[i];
        }
        if (gg > 0) k = -ge / gg;
        kappa = std::min(max_kappa, std::max(min_kappa, k));
        return min_distance(x0.data(), prev_min_distance);
    }

    // IPCSolver::post_step_update (ipc::update_barrier_stiffness)
    rigc_status post_step_update(const std::vector<double>& x) {
        double md = 0;
        rigc_status s = min_distance(x.data(), md);
        if (s != RIGC_OK) return s;
        double eps = P.dhat_epsilon * bbox_diagonal;
        eps *= eps;
        if (prev_min_distance < eps && md < eps && md < prev_min_distance) kappa = std::min(max_kappa, 2 * kappa);
        prev_min_distance = md;
        return RIGC_OK;
    }

    // P.newton_polish full Newton steps from a converged x (dir: the Newton
    // direction at x), each taken only when the CCD allows the whole step;
    // the barrier stiffness is held.
    rigc_status polish(std::vector<double>& x, std::vector<double>& dir, std::vector<double>& g, BlockHess& H) {
        for (int k = 0; k < P.newton_polish; k++) {
            std::vector<double> x1(x);
            for (size_t i = 0; i < x.size(); i++) x1[i] += dir[i];
            double toi = 0;
            rigc_status s = earliest_toi(x.data(), x1.data(), toi);
            if (s != RIGC_OK) return s;
            if (!(toi >= 1.0)) break;
            x.swap(x1);
            stats.newton_iterations++;
            double fx = 0;
            s = objective(x.data(), true, true, fx, g, H);
            if (s != RIGC_OK) return s;
            if (!solve_direction(g, H, 0.0, dir)) break;
        }
        return RIGC_OK;
    }

    // IPCSolver::solve = init_solve + NewtonSolver::solve
    rigc_status newton_solve(std::vector<double>& x, bool& success) {
        rigc_status s = init_solve(x);
        if (s != RIGC_OK) return s;
        success = false;
        double coeff = 0;
        std::vector<double> g(6 * nb), dir(6 * nb), gdir;
        BlockHess H;
        for (int it = 0; it < P.max_iterations; it++) {
            double fx = 0;
            s = objective(x.data(), true, true, fx, g, H);
            if (s != RIGC_OK) return s;
            // compute_regularized_direction (:420-470)
            const double t0 = now_ms();
            for (;;) {
                const bool ok = solve_direction(g, H, coeff, dir) && dot_free(g, dir) <= 0;
                if (coeff > 0) stats.regularizations++;
                if (ok) {
                    coeff /= 2;
                    if (coeff < 1e-8) coeff = 0;
                    break;
                }
                coeff = std::max(2 * coeff, 1e-8);
                if (verbose && coeff == 1e-8)
                    std::fprintf(stderr, "rigs: iter=%d solve failed or not descent (g.d=%g); regularizing\n", it,
                                 dot_free(g, dir));
                if (!std::isfinite(coeff)) {
                    stats.host_ms += now_ms() - t0;
                    return fail(RIGC_ERR_UNRESOLVED, "Newton regularization failed");
                }
            }
            stats.host_ms += now_ms() - t0;
            // converged (newton_solver.cpp:95-157): the step criterion and the
            // equality constraints of the kinematic bodies
            const bool energy_converged = it > 0 && converged(x, dir);
            if (energy_converged && al_satisfied(x.data())) {
                success = true;
                s = polish(x, dir, g, H);
                if (s != RIGC_OK) return s;
                break;
            }
            double step = 1;
            bool found = false;
            if (verbose >= 2)
                std::fprintf(stderr, "rigs: iter=%d at x: E %.17g, B %.17g, D %.17g\n", it, last_terms[0],
                             last_terms[1], last_terms[2]);
            s = line_search(x, dir, fx, g, step, found);
            if (s != RIGC_OK) return s;
            if (verbose)
                std::fprintf(stderr, "rigs: iter=%d f=%.17g g.d=%.6g coeff=%g step=%g found=%d kappa=%g\n", it, fx,
                             dot_free(g, dir), coeff, step, (int)found, kappa);
            if (!found) {
                gdir.assign(6 * nb, 0.0);
                for (int b = 0; b < nb; b++)
                    for (int k = 0; k < 6; k++)
                        if (dof_free(b, k)) gdir[6 * b + k] = -g[6 * b + k];
                dir = gdir;
                step = 1;
                s = line_search(x, dir, fx, g, step, found);
                if (s != RIGC_OK) return s;
                if (!found) {
                    if (it == 0 && converged(x, dir) && al_satisfied(x.data())) success = true;
                    break;
                }
            }
            for (int i = 0; i < 6 * nb; i++) x[i] += step * dir[i];
            stats.newton_iterations++;
            // NewtonSolver::post_step_update (:606-613)
            if (energy_converged && !al_satisfied(x.data())) al_update(x.data());
            s = post_step_update(x);
            if (s != RIGC_OK) return s;
        }
        return RIGC_OK;
    }

    // ------------------------------------------------------------ state
    // Body b at pose / velocity (linear in world coordinates, angular in
    // the body frame); Qdot = Q [omega]_x (rigid_body.cpp:159-160).
    void set_body_state(int b, const double* pose, const double* velocity) {
        BodyState& B = bodies[b];
        std::memcpy(B.pose, pose, sizeof(B.pose));
        std::memcpy(B.pose_prev, B.pose, sizeof(B.pose));
        std::memcpy(B.velocity, velocity, sizeof(B.velocity));
        qdot_from_omega(B.pose + 3, B.velocity + 3, B.Qdot);
        if (!B.dynamic && !B.kinematic) {
            for (int k = 0; k < 6; k++) B.velocity[k] = 0;
            for (int e = 0; e < 9; e++) B.Qdot[e] = 0;
        }
    }
    static void qdot_from_omega(const double* r, const double* w, double* Qdot) {
        double Q[9];
        rigc::rotation_d(r, Q);
        const double K[9] = { 0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0 };
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double v = 0;
                for (int k = 0; k < 3; k++) v += Q[3 * i + k] * K[3 * k + j];
                Qdot[3 * i + j] = v;
            }
    }
    // init_bbox_diagonal over the current world vertices; poses_t0 = poses
    void init_bbox() {
        std::vector<double> x0(6 * (size_t)nb), W;
        for (int b = 0; b < nb; b++) std::memcpy(&x0[6 * b], bodies[b].pose, sizeof(double) * 6);
        world_vertices(x0.data(), W);
        double lo[3] = { HUGE_VAL, HUGE_VAL, HUGE_VAL }, hi[3] = { -HUGE_VAL, -HUGE_VAL, -HUGE_VAL };
        for (int i = 0; i < nv; i++)
            for (int k = 0; k < 3; k++) {
                lo[k] = std::min(lo[k], W[3 * i + k]);
                hi[k] = std::max(hi[k], W[3 * i + k]);
            }
        double d2 = 0;
        for (int k = 0; k < 3; k++) d2 += (hi[k] - lo[k]) * (hi[k] - lo[k]);
        bbox_diagonal = std::sqrt(d2);
        poses_t0 = x0;
    }
    // Back to the state rigs_create leaves (barrier stiffness, caches).
    void reset_solver_state() {
        kappa = 1;
        max_kappa = 0;
        prev_min_distance = -1;
        cache_valid = false;
        bcache.valid = fcache.valid = false;
        ecache.valid = false;
    }

    // ------------------------------------------------------------ differentiable rollouts
    // A step's start state and the barrier stiffnesses its solve used: the
    // lagged friction set's (built at the step start) and the final one.
    struct Checkpoint {
        std::vector<double> poses, velocities, Qdot;  // [nb][6], [nb][6], [nb][9]
        double kappa_friction = 0, kappa = 0;
        bool differentiable = true;  // converged, one lagged friction set
    };
    bool recording = false;
    std::vector<Checkpoint> tape;
    std::vector<double> tape_end;  // poses after the last recorded step

    rigc_status rollout_backward(const double* dLdx, const double* dLdv, double* dLdv0, double* dLdx0);

    // ------------------------------------------------------------ step
    rigc_status step() {
        const double t_start = now_ms();
        for (auto& B : bodies) {
            std::memcpy(B.pose_prev, B.pose, sizeof(B.pose));
        }
        for (int b = 0; b < nb; b++) std::memcpy(&poses_t0[6 * b], bodies[b].pose, sizeof(double) * 6);
        bcache.valid = fcache.valid = false;  // the friction term depends on poses_t0
        ecache.valid = false;                 // and the body energy on the step's targets
        if (recording) {
            Checkpoint cp;
            cp.poses = poses_t0;
            cp.velocities.resize(6 * (size_t)nb);
            cp.Qdot.resize(9 * (size_t)nb);
            for (int b = 0; b < nb; b++) {
                std::memcpy(&cp.velocities[6 * b], bodies[b].velocity, sizeof(double) * 6);
                std::memcpy(&cp.Qdot[9 * b], bodies[b].Qdot, sizeof(double) * 9);
            }
            cp.kappa_friction = kappa;  // update_friction below uses it
            cp.differentiable = n_kin == 0;  // the adjoint has no augmented Lagrangian
            tape.push_back(std::move(cp));
        }
        step_targets();
        // update_constraints: barrier gradient at t0 (kept by the reference for
        // its integrators) and the lagged friction set at t0
        {
            double B = 0;
            std::vector<double> gB;
            rigc_status s = barrier_term(poses_t0.data(), RIGC_GRADIENT, B, &gB, nullptr);
            if (s != RIGC_OK) return s;
            s = update_friction(poses_t0.data());
            if (s != RIGC_OK) return s;
        }
        init_al();
        // solve_constraints with lagged friction
        std::vector<double> x(poses_t0);
        const double eps_d = 1e-2 * bbox_diagonal;
        int i = 0;
        bool success = false;
        for (;;) {
            rigc_status s = newton_solve(x, success);
            if (s != RIGC_OK) return s;
            if (recording) {
                Checkpoint& cp = tape.back();
                if (i == 0) cp.kappa = kappa;
                else cp.differentiable = false;  // a second lagged friction set
                if (!success) cp.differentiable = false;
            }
            if (!success) {
                stats.failed_solves++;
                break;
            }
            s = update_friction(x.data());
            if (s != RIGC_OK) return s;
            std::vector<double> gE, gB, gD;
            energy_term(x.data(), &gE, nullptr);
            double B = 0, D = 0;
            s = barrier_term(x.data(), RIGC_GRADIENT, B, &gB, nullptr);
            if (s != RIGC_OK) return s;
            if (P.coefficient_friction > 0) {
                s = friction_term(x.data(), RIGC_GRADIENT, D, &gD, nullptr);
                if (s != RIGC_OK) return s;
            }
            double bal = 0;
            for (int b = 0; b < nb; b++) {
                if (dyn_index[b] < 0) continue;
                for (int k = 0; k < 6; k++) {
                    const int e = 6 * b + k;
                    const double t = gE[e] + kappa * gB[e] + (P.coefficient_friction > 0 ? gD[e] : 0.0);
                    bal += t * t;
                }
            }
            i++;
            if (!((P.friction_iterations < 0 || i < P.friction_iterations) && std::sqrt(bal) > eps_d)) break;
        }
        // take_step
        double md = 0;
        rigc_status s = min_distance(x.data(), md);
        if (s != RIGC_OK) return s;
        stats.min_distance = md;
        const double h = P.timestep;
        for (int b = 0; b < nb; b++) {
            BodyState& B = bodies[b];
            std::memcpy(B.pose, &x[6 * b], sizeof(double) * 6);
            if (!B.dynamic) continue;
            for (int k = 0; k < 3; k++) B.velocity[k] = (B.pose[k] - B.pose_prev[k]) / h;
            double Q[9], Qp[9];
            rigc::rotation_d(B.pose + 3, Q);
            rigc::rotation_d(B.pose_prev + 3, Qp);
            for (int e = 0; e < 9; e++) B.Qdot[e] = (Q[e] - Qp[e]) / h;
            // omega in the body frame: Q^T Qdot ~ [omega]_x
            double W[9] = { 0 };
            for (int r = 0; r < 3; r++)
                for (int c = 0; c < 3; c++)
                    for (int k = 0; k < 3; k++) W[3 * r + c] += Q[3 * k + r] * B.Qdot[3 * k + c];
            B.velocity[3] = 0.5 * (W[7] - W[5]);
            B.velocity[4] = 0.5 * (W[2] - W[6]);
            B.velocity[5] = 0.5 * (W[3] - W[1]);
        }
        if (recording) tape_end = x;
        step_kinematic();
        if (P.do_intersection_check) {
            int64_t n = -1;
            s = intersections(x.data(), n);
            if (s != RIGC_OK) return s;
            stats.intersections = n;
            if (n > 0) stats.intersecting_steps++;
        }
        stats.steps++;
        stats.kappa = kappa;
        stats.total_ms += now_ms() - t_start;
        return RIGC_OK;
    }
};

// Implicit-function adjoint of a recorded rollout. Step n maps its start
// state s_n = (x_n, v_n, Qdot_n) to x_{n+1} = argmin_x F(x; s_n) with
//   F = E(x; p_n + h (v_n + h g), Q(r_n) + h Qdot_n) / m + kappa / m B(x)
//       + D(x; x_n) / m
// (compute_objective, distance_barrier_rb_problem.cpp:633-703), then sets
// v_{n+1} = (p_{n+1} - p_n) / h and Qdot_{n+1} = (Q(r_{n+1}) - Q(r_n)) / h
// (take_step, :488-626 for IMPLICIT_EULER). At a converged step
// grad_x F(x_{n+1}; s_n) = 0 on the free dofs, so with H = hess_x F there
// (unprojected) and lambda = H^-1 dL/dx_{n+1}:
//   dL/ds_n = -lambda^T d(grad_x F)/ds_n.
// The body energy's dependence on s_n is through its inertial targets; the
// friction potential's through x_n (rigc_friction_adjoint); the barrier
// potential does not depend on s_n. The barrier stiffnesses are taken as
// the forward solve left them.
rigc_status rigs_sim::rollout_backward(const double* dLdx, const double* dLdv, double* dLdv0, double* dLdx0) {
    if (tape.empty()) return fail(RIGC_ERR_INVALID_ARGUMENT, "rollout_backward: no recorded rollout");
    for (size_t n = 0; n < tape.size(); n++)
        if (!tape[n].differentiable)
            return fail(RIGC_ERR_UNRESOLVED, "rollout_backward: step " + std::to_string(n) +
                                                 " did not converge, took several lagged friction sets or "
                                                 "moved kinematic bodies");
    const bool fric = P.coefficient_friction > 0;
    if (fric && !be.friction_adjoint)
        return fail(RIGC_ERR_INVALID_ARGUMENT, "rollout_backward: the backend has no friction adjoint");
    const double h = P.timestep, im = 1.0 / average_mass;
    const double eh = P.static_friction_speed_bound * P.timestep;
    const int T = (int)tape.size();
    // the forward end state is restored afterwards
    const std::vector<BodyState> keep_bodies = bodies;
    const std::vector<double> keep_t0 = poses_t0;
    const double keep_k = kappa, keep_mk = max_kappa, keep_pm = prev_min_distance;

    std::vector<double> ax(dLdx, dLdx + 6 * (size_t)nb), av(3 * (size_t)nb, 0.0), aQ(9 * (size_t)nb, 0.0);
    if (dLdv) {
        // v_T (linear) and omega_T = vee of the skew part of Q_T^T Qdot_T
        const std::vector<double>& x1 = tape_end;
        const std::vector<double>& x0 = tape[T - 1].poses;
        for (int b = 0; b < nb; b++) {
            if (!bodies[b].dynamic) continue;
            for (int k = 0; k < 3; k++) av[3 * b + k] = dLdv[6 * b + k];
            double Q1[9], Q0[9], Qd[9];
            rigc::rotation_d(&x1[6 * b + 3], Q1);
            rigc::rotation_d(&x0[6 * b + 3], Q0);
            for (int e = 0; e < 9; e++) Qd[e] = (Q1[e] - Q0[e]) / h;
            // omega = 0.5 (W21 - W12, W02 - W20, W10 - W01), W = Q^T Qdot
            const double* aw = dLdv + 6 * b + 3;
            double dW[9] = { 0 };
            dW[7] += 0.5 * aw[0]; dW[5] -= 0.5 * aw[0];
            dW[2] += 0.5 * aw[1]; dW[6] -= 0.5 * aw[1];
            dW[3] += 0.5 * aw[2]; dW[1] -= 0.5 * aw[2];
            // dL/dQdot = Q dW; dL/dQ = Qdot dW^T
            double dQ[9] = { 0 };
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++)
                    for (int k = 0; k < 3; k++) {
                        aQ[9 * b + 3 * i + j] += Q1[3 * i + k] * dW[3 * k + j];
                        dQ[3 * i + j] += Qd[3 * i + k] * dW[3 * j + k];
                    }
            rigc::BodyFrame F;
            rigc::body_frame(&x1[6 * b + 3], F);
            for (int c = 0; c < 3; c++)
                for (int e = 0; e < 9; e++) ax[6 * b + 3 + c] += F.dR[c][e] * dQ[e];
        }
    }
    rigc_status s = RIGC_OK;
    std::vector<double> lambda(6 * (size_t)nb), fadj(6 * (size_t)nb);
    for (int n = T - 1; n >= 0 && s == RIGC_OK; n--) {
        const Checkpoint& cp = tape[n];
        const std::vector<double>& x1 = n == T - 1 ? tape_end : tape[n + 1].poses;
        const std::vector<double>& x0 = cp.poses;
        // the step's velocity and Qdot updates
        std::vector<double> g(ax), b0(6 * (size_t)nb, 0.0);
        for (int b = 0; b < nb; b++) {
            if (!bodies[b].dynamic) continue;
            for (int k = 0; k < 3; k++) {
                g[6 * b + k] += av[3 * b + k] / h;
                b0[6 * b + k] -= av[3 * b + k] / h;
            }
            rigc::BodyFrame F1, F0;
            rigc::body_frame(&x1[6 * b + 3], F1);
            rigc::body_frame(&x0[6 * b + 3], F0);
            for (int c = 0; c < 3; c++) {
                double s1 = 0, s0 = 0;
                for (int e = 0; e < 9; e++) {
                    s1 += F1.dR[c][e] * aQ[9 * b + e];
                    s0 += F0.dR[c][e] * aQ[9 * b + e];
                }
                g[6 * b + 3 + c] += s1 / h;
                b0[6 * b + 3 + c] -= s0 / h;
            }
        }
        // the step's objective at x_{n+1}, rebuilt from the checkpoint
        for (int b = 0; b < nb; b++) {
            BodyState& B = bodies[b];
            std::memcpy(B.pose, &x0[6 * b], sizeof(double) * 6);
            std::memcpy(B.velocity, &cp.velocities[6 * b], sizeof(double) * 6);
            std::memcpy(B.Qdot, &cp.Qdot[9 * b], sizeof(double) * 9);
        }
        step_targets();
        poses_t0 = x0;
        reset_solver_state();
        if (fric) {
            int64_t nf = 0;
            const double t0 = now_ms();
            s = be.update_friction_constraints(be.self, x0.data(), P.dhat, P.minimum_separation_distance,
                                               cp.kappa_friction, P.coefficient_friction, RIGC_BUILD_CONSTRAINTS, &nf);
            stats.contact_ms += now_ms() - t0;
            if (s != RIGC_OK) {
                backend_fail(s, "update_friction_constraints");
                break;
            }
        }
        kappa = cp.kappa;
        s = ensure_constraints(x1.data());
        if (s != RIGC_OK) break;
        std::vector<double> gE, gB, gD;
        BlockHess H, HB, HD;
        energy_term(x1.data(), &gE, &H);
        const int32_t flags = RIGC_HESSIAN | RIGC_UNPROJECTED;
        double B = 0, D = 0, spent = 0;
        auto bcall = [&](int32_t f, double* pe, double* pg, rigc_block_hessian* ph) {
            return be.compute_barrier_term(be.self, x1.data(), P.dhat, P.minimum_separation_distance, f, pe, pg, ph);
        };
        s = potential(bcall, stg_barrier, flags, B, &gB, &HB, spent);
        if (s != RIGC_OK) {
            backend_fail(s, "compute_barrier_term");
            break;
        }
        if (fric) {
            auto fcall = [&](int32_t f, double* pe, double* pg, rigc_block_hessian* ph) {
                return be.compute_friction_term(be.self, x1.data(), x0.data(), eh, f, pe, pg, ph);
            };
            s = potential(fcall, stg_friction, flags, D, &gD, &HD, spent);
            if (s != RIGC_OK) {
                backend_fail(s, "compute_friction_term");
                break;
            }
        }
        stats.contact_ms += spent;
        const double km = kappa * im;
        for (int b = 0; b < nb; b++)
            for (int e = 0; e < 36; e++) {
                double v = H.diag[b][e] * im + km * HB.diag[b][e];
                if (fric) v += HD.diag[b][e] * im;
                H.diag[b][e] = v;
            }
        for (auto& kv : HB.off) {
            Block& o = H.off[kv.first];
            for (int e = 0; e < 36; e++) o[e] += km * kv.second[e];
        }
        if (fric)
            for (auto& kv : HD.off) {
                Block& o = H.off[kv.first];
                for (int e = 0; e < 36; e++) o[e] += im * kv.second[e];
            }
        // lambda = H^-1 dL/dx_{n+1} on the free dofs (solve_direction solves
        // H d = -rhs)
        std::vector<double> neg(6 * (size_t)nb);
        for (int i = 0; i < 6 * nb; i++) neg[i] = -g[i];
        if (!solve_direction(neg, H, 0.0, lambda)) {
            s = fail(RIGC_ERR_UNRESOLVED, "rollout_backward: the Hessian of step " + std::to_string(n) +
                                              " could not be factorized");
            break;
        }
        // -lambda^T d(grad F)/ds_n
        std::vector<double> nax(b0), nav(3 * (size_t)nb, 0.0), naQ(9 * (size_t)nb, 0.0);
        for (int b = 0; b < nb; b++) {
            const BodyState& Bs = bodies[b];
            if (!Bs.dynamic) continue;
            const double* lb = &lambda[6 * b];
            for (int k = 0; k < 3; k++) {
                nax[6 * b + k] += Bs.mass * im * lb[k];
                nav[3 * b + k] += h * Bs.mass * im * lb[k];
            }
            rigc::BodyFrame F1, F0;
            rigc::body_frame(&x1[6 * b + 3], F1);
            rigc::body_frame(&x0[6 * b + 3], F0);
            double dP[9];
            for (int e = 0; e < 9; e++) {
                double sum = 0;
                for (int c = 0; c < 3; c++) sum += lb[3 + c] * F1.dR[c][e];
                dP[e] = sum * Bs.J[e % 3] * im;
            }
            for (int e = 0; e < 9; e++) naQ[9 * b + e] += h * dP[e];
            for (int c = 0; c < 3; c++) {
                double sum = 0;
                for (int e = 0; e < 9; e++) sum += F0.dR[c][e] * dP[e];
                nax[6 * b + 3 + c] += sum;
            }
        }
        if (fric) {
            const double t0 = now_ms();
            s = be.friction_adjoint(be.self, x1.data(), x0.data(), eh, P.dhat, P.minimum_separation_distance,
                                    cp.kappa_friction, lambda.data(), fadj.data());
            stats.contact_ms += now_ms() - t0;
            if (s != RIGC_OK) {
                backend_fail(s, "friction_adjoint");
                break;
            }
            for (int i = 0; i < 6 * nb; i++) nax[i] -= im * fadj[i];
        }
        ax.swap(nax);
        av.swap(nav);
        aQ.swap(naQ);
    }
    if (s == RIGC_OK) {
        const std::vector<double>& x0 = tape[0].poses;
        for (int b = 0; b < nb; b++) {
            double* o = dLdv0 + 6 * b;
            for (int k = 0; k < 6; k++) o[k] = 0;
            if (dLdx0)
                for (int k = 0; k < 6; k++) dLdx0[6 * b + k] = bodies[b].dynamic ? ax[6 * b + k] : 0.0;
            if (!bodies[b].dynamic) continue;
            for (int k = 0; k < 3; k++) o[k] = av[3 * b + k];
            // Qdot_0 = Q_0 [omega_0]_x
            double Q0[9];
            rigc::rotation_d(&x0[6 * b + 3], Q0);
            for (int c = 0; c < 3; c++) {
                double sum = 0;
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) {
                        double qk = 0;  // (Q_0 [e_c]_x)_ij
                        for (int k = 0; k < 3; k++) qk += Q0[3 * i + k] * rigc::v3_skew_unit(c, k, j);
                        sum += aQ[9 * b + 3 * i + j] * qk;
                    }
                o[3 + c] = sum;
            }
            if (dLdx0) {
                const double* w = &tape[0].velocities[6 * b + 3];
                const double K[9] = { 0, -w[2], w[1], w[2], 0, -w[0], -w[1], w[0], 0 };
                double dQ[9] = { 0 };  // dL/dQ_0 = aQ K^T
                for (int i = 0; i < 3; i++)
                    for (int k = 0; k < 3; k++)
                        for (int j = 0; j < 3; j++) dQ[3 * i + k] += aQ[9 * b + 3 * i + j] * K[3 * k + j];
                rigc::BodyFrame F0;
                rigc::body_frame(&x0[6 * b + 3], F0);
                for (int c = 0; c < 3; c++)
                    for (int e = 0; e < 9; e++) dLdx0[6 * b + 3 + c] += F0.dR[c][e] * dQ[e];
            }
        }
    }
    bodies = keep_bodies;
    poses_t0 = keep_t0;
    reset_solver_state();
    kappa = keep_k;
    max_kappa = keep_mk;
    prev_min_distance = keep_pm;
    return s;
}

namespace rigs {
void set_fork_hooks(rigs_sim* sim, void (*join)(void*), void (*part)(void*), void* arg) {
    if (!sim) return;
    sim->fork_join = join;
    sim->fork_part = part;
    sim->fork_arg = arg;
}

// The batch stepper runs every scene's stepper on its own thread already.
void set_host_overlap(rigs_sim* sim, bool on) {
    if (!sim) return;
    sim->overlap = on;
    if (!on && sim->fric_ctx) {  // friction back on the main context
        rigc_destroy(sim->fric_ctx);
        sim->fric_ctx = nullptr;
        sim->fbe = sim->be;
    }
}
}  // namespace rigs

extern "C" {

void rigs_default_params(rigs_params* p) {
    if (!p) return;
    p->timestep = 0.01;
    p->gravity[0] = 0;
    p->gravity[1] = -9.81;
    p->gravity[2] = 0;
    p->dhat = 1e-3;
    p->minimum_separation_distance = 0;
    p->coefficient_friction = 0;
    p->friction_iterations = 1;
    p->static_friction_speed_bound = 1e-3;
    p->velocity_conv_tol = 1e-2;
    p->line_search_lower_bound = 1e-12;
    p->max_iterations = 3000;
    p->dhat_epsilon = 1e-9;
    p->min_barrier_stiffness_scale = 1e11;
    p->do_intersection_check = 0;
    p->trajectory_type = RIGS_TRAJECTORY_RIGID;
    p->newton_polish = 0;
    p->detection_method = RIGC_DETECTION_HASH_GRID;
}

rigc_status rigs_create(rigs_sim** out, const rigs_backend* backend, int32_t device, const rigc_scene* scene,
                        const rigs_bodies* bodies, const rigs_params* params) {
    if (!out || !scene || !bodies || !params || bodies->num_bodies != scene->num_bodies || !bodies->poses
        || !bodies->velocities || !bodies->mass || !bodies->inertia || !(params->timestep > 0) || !(params->dhat > 0)
        || (params->trajectory_type != RIGS_TRAJECTORY_RIGID
            && params->trajectory_type != RIGS_TRAJECTORY_PIECEWISE_LINEAR))
        return RIGC_ERR_INVALID_ARGUMENT;
    if (params->trajectory_type == RIGS_TRAJECTORY_PIECEWISE_LINEAR && backend
        && !backend->compute_earliest_toi_piecewise_linear)
        return RIGC_ERR_INVALID_ARGUMENT;
    if (params->detection_method != RIGC_DETECTION_HASH_GRID
        && (params->detection_method != RIGC_DETECTION_BVH || (backend && !backend->set_detection_method)))
        return RIGC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    rigs_sim* sim = new rigs_sim();
    sim->P = *params;
    const char* vb = std::getenv("RIGS_VERBOSE");
    sim->verbose = vb ? std::atoi(vb) : 0;
    if (backend) {
        sim->be = *backend;
    } else {
        rigc_status s = rigc_create(&sim->own_ctx, device);
        if (s != RIGC_OK) {
            delete sim;
            return s;
        }
        sim->be = rigs_backend{ sim->own_ctx, gb_set_scene, gb_toi, gb_build, gb_barrier, gb_mind,
                                gb_fupdate, gb_friction, gb_error, gb_stats, gb_isect, gb_toi_pl, gb_fadjoint,
                                gb_detection };
    }
    sim->fbe = sim->be;
    rigc_status s = sim->be.set_scene(sim->be.self, scene);
    if (s == RIGC_OK && params->detection_method != RIGC_DETECTION_HASH_GRID)
        s = sim->be.set_detection_method(sim->be.self, params->detection_method);
    if (s != RIGC_OK) {
        if (sim->own_ctx) rigc_destroy(sim->own_ctx);
        delete sim;
        return s;
    }
    // Helper threads pay a wake-up of some tens of microseconds per use:
    // below OVERLAP_MIN_BODIES bodies the work they would hide is smaller.
    sim->overlap = scene->num_bodies >= OVERLAP_MIN_BODIES;
    if (sim->overlap && sim->own_ctx && params->coefficient_friction > 0) {
        // a second context for the friction term (without one the friction
        // term just runs through the main context)
        if (rigc_create(&sim->fric_ctx, device) == RIGC_OK && rigc_set_scene(sim->fric_ctx, scene) == RIGC_OK) {
            sim->fbe = rigs_backend{ sim->fric_ctx, gb_set_scene, gb_toi, gb_build, gb_barrier, gb_mind,
                                     gb_fupdate, gb_friction, gb_error, gb_stats, gb_isect, gb_toi_pl, gb_fadjoint,
                                gb_detection };
        } else if (sim->fric_ctx) {
            rigc_destroy(sim->fric_ctx);
            sim->fric_ctx = nullptr;
        }
    }
    const int nb = scene->num_bodies;
    sim->nb = nb;
    sim->nv = scene->num_vertices;
    sim->V.assign(scene->vertices, scene->vertices + 3 * (size_t)sim->nv);
    sim->vstart.assign(scene->body_vertex_start, scene->body_vertex_start + nb + 1);
    sim->bodies.resize(nb);
    sim->dyn_index.assign(nb, -1);
    double mass_sum = 0;
    int n_free_dof = 0;
    const int32_t* ks = bodies->kinematic_pose_start;
    for (int b = 0; b < nb; b++) {
        BodyState& B = sim->bodies[b];
        const bool kin = bodies->kinematic && bodies->kinematic[b] != 0;
        if (kin && scene->is_static[b]) {
            rigs_destroy(sim);
            return RIGC_ERR_INVALID_ARGUMENT;  // a kinematic body is not static in the scene
        }
        B.kinematic = kin;
        B.dynamic = !scene->is_static[b] && !kin;
        B.mass = bodies->mass[b];
        for (int k = 0; k < 3; k++) B.inertia[k] = bodies->inertia[3 * b + k];
        B.J[0] = 0.5 * (-B.inertia[0] + B.inertia[1] + B.inertia[2]);
        B.J[1] = 0.5 * (B.inertia[0] - B.inertia[1] + B.inertia[2]);
        B.J[2] = 0.5 * (B.inertia[0] + B.inertia[1] - B.inertia[2]);
        for (int k = 0; k < 3; k++) B.Jsqrt[k] = std::sqrt(std::max(B.J[k], 0.0));
        if (bodies->force && !scene->is_static[b])
            for (int k = 0; k < 6; k++) B.force[k] = bodies->force[6 * b + k];
        if (kin) {
            const double mt = bodies->kinematic_max_time ? bodies->kinematic_max_time[b] : -1.0;
            B.kin_max_time = mt < 0 ? HUGE_VAL : mt;
            if (ks && bodies->kinematic_poses) {
                if (ks[b] < 0 || ks[b + 1] < ks[b]) {
                    rigs_destroy(sim);
                    return RIGC_ERR_INVALID_ARGUMENT;
                }
                for (int i = ks[b]; i < ks[b + 1]; i++) {
                    std::array<double, 6> pz;
                    for (int k = 0; k < 6; k++) pz[k] = bodies->kinematic_poses[6 * (size_t)i + k];
                    B.kin_poses.push_back(pz);
                }
            }
            sim->n_kin++;
        }
        sim->set_body_state(b, bodies->poses + 6 * b, bodies->velocities + 6 * b);
        if (!scene->is_static[b]) {
            // the free set: dynamic and kinematic bodies (is_rb_dof_fixed)
            sim->dyn_index[b] = sim->n_free++;
            mass_sum += 3 * B.mass + B.inertia[0] + B.inertia[1] + B.inertia[2];
            n_free_dof += 6;
        }
    }
    sim->bodies_init = sim->bodies;
    sim->x_pred.assign(6 * (size_t)nb, 0.0);
    sim->dof_sat.assign(6 * (size_t)nb, 0);
    sim->al_lam.assign(3 * (size_t)nb, 0.0);
    sim->al_Lam.assign(9 * (size_t)nb, 0.0);
    sim->average_mass = n_free_dof ? mass_sum / n_free_dof : 1.0;
    sim->init_bbox();
    std::vector<double> x0(sim->poses_t0);
    // the initial state is checked for intersections
    // (RigidBodyProblem::init, rigid_body_problem.cpp:100)
    int64_t n0 = -1;
    s = sim->intersections(x0.data(), n0);
    if (s != RIGC_OK) {
        rigs_destroy(sim);
        return s;
    }
    sim->stats.initial_intersections = n0;
    sim->stats.intersections = -1;
    sim->stats.contact_ms = 0;
    *out = sim;
    return RIGC_OK;
}

rigc_status rigs_destroy(rigs_sim* sim) {
    if (!sim) return RIGC_OK;
    if (sim->fric_ctx) rigc_destroy(sim->fric_ctx);
    if (sim->own_ctx) rigc_destroy(sim->own_ctx);
    delete sim;
    return RIGC_OK;
}

const char* rigs_last_error(const rigs_sim* sim) { return sim ? sim->err.c_str() : "null simulation"; }

rigc_status rigs_step(rigs_sim* sim, int32_t num_steps) {
    if (!sim || num_steps < 0) return RIGC_ERR_INVALID_ARGUMENT;
    for (int i = 0; i < num_steps; i++) {
        rigc_status s = sim->step();
        if (s != RIGC_OK) return s;
    }
    return RIGC_OK;
}

rigc_status rigs_set_state(rigs_sim* sim, const double* poses, const double* velocities) {
    if (!sim || !poses || !velocities) return RIGC_ERR_INVALID_ARGUMENT;
    for (int b = 0; b < sim->nb; b++) {
        // kinematic scripts and external forces as rigs_create left them
        BodyState& B = sim->bodies[b];
        const BodyState& I = sim->bodies_init[b];
        B.dynamic = I.dynamic;
        B.kinematic = I.kinematic;
        B.kin_max_time = I.kin_max_time;
        B.kin_poses = I.kin_poses;
        std::memcpy(B.force, I.force, sizeof(B.force));
        sim->set_body_state(b, poses + 6 * b, velocities + 6 * b);
    }
    sim->init_bbox();
    sim->reset_solver_state();
    sim->tape.clear();
    sim->tape_end.clear();
    return RIGC_OK;
}

rigc_status rigs_rollout(rigs_sim* sim, int32_t num_steps) {
    if (!sim || num_steps < 0) return RIGC_ERR_INVALID_ARGUMENT;
    sim->tape.clear();
    sim->tape_end.clear();
    sim->recording = true;
    rigc_status s = RIGC_OK;
    for (int i = 0; i < num_steps && s == RIGC_OK; i++) s = sim->step();
    sim->recording = false;
    return s;
}

rigc_status rigs_rollout_backward(rigs_sim* sim, const double* dL_dposes, const double* dL_dvelocities,
                                  double* dL_dvelocities0, double* dL_dposes0) {
    if (!sim || !dL_dposes || !dL_dvelocities0) return RIGC_ERR_INVALID_ARGUMENT;
    return sim->rollout_backward(dL_dposes, dL_dvelocities, dL_dvelocities0, dL_dposes0);
}

rigc_status rigs_get_state(const rigs_sim* sim, double* poses, double* velocities) {
    if (!sim) return RIGC_ERR_INVALID_ARGUMENT;
    for (int b = 0; b < sim->nb; b++) {
        if (poses) std::memcpy(poses + 6 * b, sim->bodies[b].pose, sizeof(double) * 6);
        if (velocities) std::memcpy(velocities + 6 * b, sim->bodies[b].velocity, sizeof(double) * 6);
    }
    return RIGC_OK;
}

rigc_status rigs_get_stats(const rigs_sim* sim, rigs_stats* stats) {
    if (!sim || !stats) return RIGC_ERR_INVALID_ARGUMENT;
    *stats = sim->stats;
    return RIGC_OK;
}

rigc_status rigs_block_ldlt_solve(int32_t n, const double* diag, int64_t m, const int32_t* pairs, const double* off,
                                  const double* b, double* x) {
    if (n < 0 || m < 0 || (n && (!diag || !b || !x)) || (m && (!pairs || !off))) return RIGC_ERR_INVALID_ARGUMENT;
    std::vector<Block> D(n), O(m);
    std::vector<std::pair<int, int>> P(m);
    for (int i = 0; i < n; i++) std::copy(diag + 36 * i, diag + 36 * i + 36, D[i].begin());
    for (int64_t e = 0; e < m; e++) {
        std::copy(off + 36 * e, off + 36 * e + 36, O[e].begin());
        P[e] = { pairs[2 * e], pairs[2 * e + 1] };
        if (P[e].first < 0 || P[e].first >= n || P[e].second < 0 || P[e].second >= n) return RIGC_ERR_INVALID_ARGUMENT;
    }
    rigs::BlockLDLT L;
    if (!L.factorize(n, D, P, O)) return RIGC_ERR_UNRESOLVED;
    L.solve(b, x);
    return RIGC_OK;
}

}  // extern "C"
