// ORACLE — test infrastructure only (see oracle_interval.hpp).
// extern "C" entry points that the Python tests load with ctypes.
#include <atomic>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <cstring>

#include "oracle_ccd.hpp"
#include "oracle_barrier.hpp"
#include "oracle_scene.hpp"

using namespace oracle;

namespace {
void unpack_ee(const double* q, EEQuery& e) {
    std::memcpy(e.a0, q + 0, 3 * sizeof(double));
    std::memcpy(e.a1, q + 3, 3 * sizeof(double));
    std::memcpy(e.b0, q + 6, 3 * sizeof(double));
    std::memcpy(e.b1, q + 9, 3 * sizeof(double));
    std::memcpy(e.poseA_t0, q + 12, 6 * sizeof(double));
    std::memcpy(e.poseA_t1, q + 18, 6 * sizeof(double));
    std::memcpy(e.poseB_t0, q + 24, 6 * sizeof(double));
    std::memcpy(e.poseB_t1, q + 30, 6 * sizeof(double));
}
void unpack_fv(const double* q, FVQuery& f) {
    std::memcpy(f.v, q + 0, 3 * sizeof(double));
    std::memcpy(f.f0, q + 3, 3 * sizeof(double));
    std::memcpy(f.f1, q + 6, 3 * sizeof(double));
    std::memcpy(f.f2, q + 9, 3 * sizeof(double));
    std::memcpy(f.poseA_t0, q + 12, 6 * sizeof(double));
    std::memcpy(f.poseA_t1, q + 18, 6 * sizeof(double));
    std::memcpy(f.poseB_t0, q + 24, 6 * sizeof(double));
    std::memcpy(f.poseB_t1, q + 30, 6 * sizeof(double));
}
}  // namespace

extern "C" {

// ---- interval primitives
void oracle_interval_cos(double lo, double hi, double* out) {
    Interval r = cos(Interval(lo, hi));
    out[0] = r.lo; out[1] = r.hi;
}
void oracle_interval_sin(double lo, double hi, double* out) {
    Interval r = sin(Interval(lo, hi));
    out[0] = r.lo; out[1] = r.hi;
}
void oracle_interval_sinc(double lo, double hi, double* out) {
    Interval r = sinc(Interval(lo, hi));
    out[0] = r.lo; out[1] = r.hi;
}
double oracle_sinc(double x) { return sinc(x); }
double oracle_cos_rn(double x) { return cos_rn(x); }
double oracle_sin_rn(double x) { return sin_rn_full(x); }

// op: 0 add, 1 sub, 2 mul, 3 div
void oracle_interval_binop(int op, double xl, double xu, double yl, double yu, double* out) {
    Interval x(xl, xu), y(yl, yu), r;
    switch (op) {
    case 0: r = x + y; break;
    case 1: r = x - y; break;
    case 2: r = x * y; break;
    default: r = x / y; break;
    }
    out[0] = r.lo; out[1] = r.hi;
}

// Interval rotation matrix from an interval rotation vector r (3x2 doubles),
// out: 3x3x2 doubles row-major.
void oracle_rotation_interval(const double* r, double* out) {
    Vec3I rv;
    for (int i = 0; i < 3; i++) rv[i] = Interval(r[2 * i], r[2 * i + 1]);
    Mat3I R = rotation_matrix(rv);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            out[2 * (3 * i + j)] = R.m[i][j].lo;
            out[2 * (3 * i + j) + 1] = R.m[i][j].hi;
        }
}
void oracle_rotation_double(const double* r, double* out) {
    Mat3d R = rotation_matrix(r);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[3 * i + j] = R.m[i][j];
}

// Root of a*x^2 + b*x + c on [0,1] (tests/interval/test_interval_root_finder.cpp:9-33),
// optional constraint x.lower() > lower_gt.
int oracle_root_quadratic(double a, double b, double c, double tol, int use_constraint,
                          double lower_gt, double* out) {
    auto f = [&](const Interval& x) { return a * x * x + b * x + c; };
    Interval sol;
    bool found;
    if (use_constraint)
        found = interval_root_finder_1d(f, [&](const Interval& x) { return x.lo > lower_gt; },
                                        Interval(0, 1), tol, sol);
    else
        found = interval_root_finder_1d(f, [](const Interval&) { return true; }, Interval(0, 1),
                                        tol, sol);
    out[0] = sol.lo; out[1] = sol.hi;
    return found ? 1 : 0;
}
// Root of cos(2*pi*x) + yshift on [0,1] (test_interval_root_finder.cpp:35-52).
int oracle_root_trig(double two_pi, double yshift, double tol, double* out) {
    auto f = [&](const Interval& x) { return cos(two_pi * x) + yshift; };
    Interval sol;
    bool found = interval_root_finder_1d(f, [](const Interval&) { return true; }, Interval(0, 1),
                                         tol, sol);
    out[0] = sol.lo; out[1] = sol.hi;
    return found ? 1 : 0;
}

// ---- rigid CCD per query. q layout: 4 body-local vertices (12 doubles),
// then poseA_t0, poseA_t1, poseB_t0, poseB_t1 (4 x 6 doubles).
// Returns 1 (impact), 0 (no impact) or -1 (evaluation budget exceeded).
void oracle_set_eval_budget(long budget) { eval_budget() = budget; }
int oracle_ee_toi(const double* q, double earliest_toi, double toi_tol, double* toi, long* n_evals) {
    EEQuery e;
    unpack_ee(q, e);
    bool hit = edge_edge_toi(e, *toi, earliest_toi, toi_tol, n_evals);
    if (budget_exceeded()) return -1;
    return hit ? 1 : 0;
}
int oracle_fv_toi(const double* q, double earliest_toi, double toi_tol, double* toi, long* n_evals) {
    FVQuery f;
    unpack_fv(q, f);
    bool hit = face_vertex_toi(f, *toi, earliest_toi, toi_tol, n_evals);
    if (budget_exceeded()) return -1;
    return hit ? 1 : 0;
}
void oracle_ee_aabb(const double* q, const double* box, double* out) {
    EEQuery e;
    unpack_ee(q, e);
    Vec3I d = edge_edge_aabb(e, Interval(box[0], box[1]), Interval(box[2], box[3]),
                             Interval(box[4], box[5]));
    for (int i = 0; i < 3; i++) { out[2 * i] = d[i].lo; out[2 * i + 1] = d[i].hi; }
}
void oracle_fv_aabb(const double* q, const double* box, double* out) {
    FVQuery f;
    unpack_fv(q, f);
    Vec3I d = face_vertex_aabb(f, Interval(box[0], box[1]), Interval(box[2], box[3]),
                               Interval(box[4], box[5]));
    for (int i = 0; i < 3; i++) { out[2 * i] = d[i].lo; out[2 * i + 1] = d[i].hi; }
}

// ---- FP64 op counts (meaningful in the -DORACLE_COUNT_FLOPS build only)
long oracle_flops_get() {
#ifdef ORACLE_COUNT_FLOPS
    return flop_counter();
#else
    return -1;
#endif
}
void oracle_flops_reset() {
#ifdef ORACLE_COUNT_FLOPS
    flop_counter() = 0;
#endif
}

// ---- scene + broad phase
// DetectionMethod of the broad phase and of the earliest TOI's candidates:
// 0 HASH_GRID, 1 BVH (src/ccd/ccd.hpp:20-24). Loading a scene resets it to
// the hash grid; a caller that wants the BVH sets it after the scene.
static int g_detection = 0;

void oracle_set_scene(int nb, int nv, int ne, int nf, const double* V, const int* E, const int* F,
                      const int* FE, const int* bv, const int* be, const int* bf, const int* group,
                      const unsigned char* is_static) {
    g_detection = 0;
    Scene& s = scene();
    s.nb = nb; s.nv = nv; s.ne = ne; s.nf = nf;
    s.V.assign(V, V + 3 * (size_t)nv);
    s.E.assign(E, E + 2 * (size_t)ne);
    s.F.assign(F, F + 3 * (size_t)nf);
    s.FE.assign(FE, FE + 3 * (size_t)nf);
    s.bv.assign(bv, bv + nb + 1);
    s.be.assign(be, be + nb + 1);
    s.bf.assign(bf, bf + nb + 1);
    s.group.assign(group, group + nb);
    s.is_static.assign(is_static, is_static + nb);
    scene_derive(s);
}

static Candidates g_candidates;
void oracle_set_detection_method(int m) { g_detection = m; }
int oracle_bvh_supported() { return bvh_supported(scene()) ? 1 : 0; }

// continuous != 0: detect_collision_candidates_rigid(poses_t0, poses_t1);
// otherwise the discrete version at poses_t0.
void oracle_broadphase(int continuous, const double* p0, const double* p1, double r, long* n_ee, long* n_fv) {
    if (g_detection == 1) broadphase_bvh(scene(), p0, p1, r, continuous != 0, g_candidates);
    else if (continuous) broadphase_ccd(scene(), p0, p1, r, g_candidates);
    else broadphase_static(scene(), p0, r, g_candidates);
    *n_ee = (long)g_candidates.ee.size();
    *n_fv = (long)g_candidates.fv.size();
}
void oracle_get_candidates(int* ee, int* fv) {
    for (size_t i = 0; i < g_candidates.ee.size(); i++) { ee[2 * i] = g_candidates.ee[i][0]; ee[2 * i + 1] = g_candidates.ee[i][1]; }
    for (size_t i = 0; i < g_candidates.fv.size(); i++) { fv[2 * i] = g_candidates.fv[i][0]; fv[2 * i + 1] = g_candidates.fv[i][1]; }
}
// [nv][7]: included flag, AABB min xyz, AABB max xyz of the last broad phase.
void oracle_last_vertex_aabbs(double* out) {
    const std::vector<double>& v = last_vertex_aabbs();
    std::copy(v.begin(), v.end(), out);
}
// DistanceBarrierConstraint::compute_earliest_toi
// (src/opt/distance_barrier_constraint.cpp:130-265), parallel over
// candidates as the reference's tbb::parallel_for; the narrow phase per
// trajectory type as edge_edge_ccd / face_vertex_ccd dispatch it
// (src/ccd/ccd.cpp:271-279, 339-347): 0 RIGID, 1 PIECEWISE_LINEAR (with the
// caller's minimum separation).
//
// As in the reference, the queries share the running earliest TOI
// (:185-244), but deterministically: every rigid query searches t in [0,1]
// and drops the boxes whose t.lo is at or above the running minimum
// (interval_root_finder_nd), so the call returns the minimum over queries
// of their [0,1] results, bit for bit, in any schedule (the reference
// instead shrinks each query's t-domain to a racily read earliest_toi,
// which makes its trees depend on the schedule).
// oracle_set_ccd_pruning(0) turns the sharing off (each query runs its
// whole tree); oracle_last_box_evaluations() gives the inclusion-function
// evaluations of the last call, summed over its queries.
namespace {
bool g_ccd_pruning = true;
long g_last_box_evaluations = 0;
}
void oracle_set_ccd_pruning(int on) { g_ccd_pruning = on != 0; }
long oracle_last_box_evaluations() { return g_last_box_evaluations; }
void oracle_set_num_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}
int oracle_max_threads() {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
double oracle_earliest_toi_trajectory(const double* p0, const double* p1, double msd, int trajectory,
                                      long* n_candidates) {
    const Scene& s = scene();
    Candidates c;
    if (g_detection == 1) broadphase_bvh(s, p0, p1, msd / 2.0, true, c);
    else broadphase_ccd(s, p0, p1, msd / 2.0, c);
    const long ne = (long)c.ee.size(), nf = (long)c.fv.size();
    *n_candidates = ne + nf;
    std::atomic<double> shared(INFINITY);
    const std::atomic<double>* prune = g_ccd_pruning ? &shared : nullptr;
    long boxes = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : boxes)
    for (long i = 0; i < ne + nf; i++) {
        double toi = INFINITY;
        long evals = 0;
        bool hit;
        if (i < ne) {
            const int ea = c.ee[i][0], eb = c.ee[i][1];
            const int A = s.e_body[ea], B = s.e_body[eb];
            EEQuery q;
            for (int k = 0; k < 3; k++) {
                q.a0[k] = s.V[3 * s.E[2 * ea] + k]; q.a1[k] = s.V[3 * s.E[2 * ea + 1] + k];
                q.b0[k] = s.V[3 * s.E[2 * eb] + k]; q.b1[k] = s.V[3 * s.E[2 * eb + 1] + k];
            }
            for (int k = 0; k < 6; k++) {
                q.poseA_t0[k] = p0[6 * A + k]; q.poseA_t1[k] = p1[6 * A + k];
                q.poseB_t0[k] = p0[6 * B + k]; q.poseB_t1[k] = p1[6 * B + k];
            }
            hit = trajectory == 1 ? pl_edge_edge_toi(q, toi, 1.0, msd, nullptr, nullptr)
                                  : edge_edge_toi(q, toi, 1.0, 1e-4, &evals, prune);
        } else {
            const int f = c.fv[i - ne][0], v = c.fv[i - ne][1];
            const int A = s.v_body[v], B = s.f_body[f];
            FVQuery q;
            for (int k = 0; k < 3; k++) {
                q.v[k] = s.V[3 * v + k];
                q.f0[k] = s.V[3 * s.F[3 * f] + k]; q.f1[k] = s.V[3 * s.F[3 * f + 1] + k];
                q.f2[k] = s.V[3 * s.F[3 * f + 2] + k];
            }
            for (int k = 0; k < 6; k++) {
                q.poseA_t0[k] = p0[6 * A + k]; q.poseA_t1[k] = p1[6 * A + k];
                q.poseB_t0[k] = p0[6 * B + k]; q.poseB_t1[k] = p1[6 * B + k];
            }
            hit = trajectory == 1 ? pl_face_vertex_toi(q, toi, 1.0, msd, nullptr, nullptr)
                                  : face_vertex_toi(q, toi, 1.0, 1e-4, &evals, prune);
        }
        boxes += evals;
        if (hit) {
            double cur = shared.load(std::memory_order_relaxed);
            while (toi < cur && !shared.compare_exchange_weak(cur, toi, std::memory_order_relaxed)) {
            }
        }
    }
    g_last_box_evaluations = boxes;
    return shared.load();
}
double oracle_earliest_toi(const double* p0, const double* p1, double msd, long* n_candidates) {
    return oracle_earliest_toi_trajectory(p0, p1, msd, 0, n_candidates);
}

double oracle_average_edge_length() { return scene().average_edge_length; }

// Double-precision world vertices of every body at the given poses.
void oracle_world_vertices(const double* poses, double* out) {
    const Scene& s = scene();
    std::vector<double> w;
    for (int b = 0; b < s.nb; b++) {
        world_vertices(s, b, poses + 6 * b, w);
        std::copy(w.begin(), w.end(), out + 3 * (size_t)s.bv[b]);
    }
}

// ---- constraint set, barrier, friction
static std::vector<Constraint> g_constraints;
static std::vector<FrictionConstraint> g_friction;

void oracle_set_psd_projection(int on) { psd_projection_enabled() = on != 0; }

// counts: vv, ev, ee, fv
void oracle_constraint_set(const double* poses, double dhat, double dmin, long* counts) {
    construct_constraint_set(scene(), poses, dhat, dmin, g_constraints);
    for (int k = 0; k < 4; k++) counts[k] = 0;
    for (const auto& c : g_constraints) counts[c.type]++;
}
// per constraint: type, a, b, multiplicity, body0, body1 (int64) and eps_x
void oracle_get_constraints(long* rec, double* eps_x) {
    for (size_t i = 0; i < g_constraints.size(); i++) {
        const Constraint& c = g_constraints[i];
        long* r = rec + 6 * i;
        r[0] = c.type; r[1] = c.a; r[2] = c.b; r[3] = c.multiplicity; r[4] = c.body[0]; r[5] = c.body[1];
        eps_x[i] = c.eps_x;
    }
}
// Barrier term on the current constraint set: energy, grad (6 nb, zeroed
// here) and per-constraint projected 12x12 (optional, n x 144).
double oracle_barrier(const double* poses, double dhat, double dmin, double* grad, double* local_hess) {
    const Scene& s = scene();
    for (int i = 0; i < 6 * s.nb; i++) grad[i] = 0;
    std::vector<std::array<double, 144>> lh;
    const double e = barrier_term(s, poses, g_constraints, dhat, dmin, grad, local_hess ? &lh : nullptr);
    if (local_hess)
        for (size_t i = 0; i < lh.size(); i++) std::copy(lh[i].begin(), lh[i].end(), local_hess + 144 * i);
    return e;
}
double oracle_min_distance(const double* poses) {
    return minimum_distance(scene(), poses, g_constraints);
}
long oracle_friction_update(const double* poses, double dhat, double dmin, double kappa, double mu) {
    construct_friction_constraint_set(scene(), poses, g_constraints, dhat, dmin, kappa, mu, g_friction);
    return (long)g_friction.size();
}
void oracle_friction_adjoint(const double* poses, const double* poses_t0, double epsv_times_h, double dhat,
                             double dmin, double kappa, const double* lambda, double* out) {
    friction_adjoint(scene(), poses, poses_t0, g_friction, epsv_times_h, dhat, dmin, kappa, lambda, out);
}
// per friction constraint: normal force, coords (2), basis (3x2 row-major)
void oracle_get_friction(double* rec) {
    for (size_t i = 0; i < g_friction.size(); i++) {
        const FrictionConstraint& f = g_friction[i];
        double* r = rec + 9 * i;
        r[0] = f.normal_force; r[1] = f.coords[0]; r[2] = f.coords[1];
        for (int k = 0; k < 3; k++) { r[3 + 2 * k] = f.basis[k][0]; r[4 + 2 * k] = f.basis[k][1]; }
    }
}
double oracle_friction(const double* poses, const double* poses_t0, double epsv_times_h, double* grad,
                       double* local_hess) {
    const Scene& s = scene();
    for (int i = 0; i < 6 * s.nb; i++) grad[i] = 0;
    std::vector<std::array<double, 144>> lh;
    const double e = friction_term(s, poses, poses_t0, g_friction, epsv_times_h, grad, local_hess ? &lh : nullptr);
    if (local_hess)
        for (size_t i = 0; i < lh.size(); i++) std::copy(lh[i].begin(), lh[i].end(), local_hess + 144 * i);
    return e;
}
double oracle_constraint_local(int type, const double* X, double eps_x, double dhat, double dmin, long multiplicity,
                               double* g, double* H) {
    return constraint_local(type, X, eps_x, dhat, dmin, multiplicity, g, H);
}
void oracle_point_derivatives(const double* pose, const double* v, double* w, double* jac, double* hess) {
    point_derivatives(pose, v, w, jac, hess);
}
void oracle_vertex_derivatives(const double* pose, int vertex, double* w, double* jac, double* hess) {
    vertex_derivatives(scene(), pose, vertex, w, jac, hess);
}

// ---- intersection check
int oracle_orient3d(const double* a, const double* b, const double* c, const double* d) {
    return orient3d_sign(a, b, c, d);
}
int oracle_edge_triangle(const double* p, const double* q, const double* t0, const double* t1, const double* t2) {
    return edge_triangle_intersect(p, q, t0, t1, t2) ? 1 : 0;
}
// Intersecting (edge, face) pairs at poses, sorted; pairs [cap][2]; returns the count.
long oracle_detect_intersections(const double* poses, int* pairs, long cap) {
    const std::vector<std::array<int, 2>> r = detect_intersections(scene(), poses);
    for (long i = 0; i < (long)r.size() && i < cap; i++) {
        pairs[2 * i] = r[i][0];
        pairs[2 * i + 1] = r[i][1];
    }
    return (long)r.size();
}

}  // extern "C"
