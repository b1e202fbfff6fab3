// ORACLE — test infrastructure only (see oracle_interval.hpp).
//
// Restatement of Rigid IPC's piecewise-linear CCD (TrajectoryType::PIECEWISE_LINEAR):
//   compute_piecewise_linear_edge_edge_time_of_impact    src/ccd/piecewise_linear/time_of_impact.cpp:296-454
//   compute_piecewise_linear_face_vertex_time_of_impact  src/ccd/piecewise_linear/time_of_impact.cpp:458-649
//   constants                                            src/ccd/piecewise_linear/time_of_impact.cpp:30-53
//   dispatch from edge_edge_ccd / face_vertex_ccd        src/ccd/ccd.cpp:271-274, 339-342
// and of the linear CCD those functions call, Tight-Inclusion
// (inclusion_ccd::edgeEdgeCCD_double / vertexFaceCCD_double, git tag
// b913416aecf1bbb8a51726039dcb94f395849cab per cmake/recipes/tight_inclusion.cmake:8-12).
// Tight-Inclusion is an external dependency that is not in the reference
// tree; its published algorithm (Wang et al. 2021, "A Large Scale Benchmark
// and an Inclusion-Based Algorithm for Continuous Collision Detection") is
// restated here as the reference calls it: error filter computed from the
// vertices (err = -1), minimum separation ms, co-domain tolerance 1e-6,
// t_max = 1, max_itr = 1e6, ccd_type 1 (breadth-first "horizontal tree"),
// no_zero_toi = false. Its bits against the real library are unpinned; the
// closed-form TOIs of the reference's own tests
// (tests/ccd/test_rigid_body_time_of_impact.cpp:170-303) pin its behaviour.
//
// The breadth-first queue is std::priority_queue with Tight-Inclusion's
// comparator (level first, then the lower end of t), so boxes of equal level
// and equal t are served in libstdc++'s heap order.
#include "oracle_barrier.hpp"
#include "oracle_ccd.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <queue>
#include <stack>
#include <vector>

namespace oracle {

namespace {

// ---------------------------------------------------------------- Tight-Inclusion
// Interval endpoints of the (t, u, v) parameter box. Tight-Inclusion keeps
// them as dyadic rationals n / 2^p; every endpoint produced by bisecting
// [0, 1] is a dyadic rational that a double holds exactly, and so is every
// midpoint, so doubles give the same values and the same comparisons.
struct TIBox {
    double lo[3], hi[3];
};
struct TIItem {
    TIBox b;
    int level;
};
struct TICmp {
    bool operator()(const TIItem& a, const TIItem& b) const {
        if (a.level != b.level) return a.level >= b.level;
        return !(a.b.lo[0] < b.b.lo[0]);
    }
};

constexpr double kTolerance = 1e-6;      // LINEAR_CCD_TOL (time_of_impact.cpp:38)
constexpr long kMaxItr = 1000000;        // LINEAR_CCD_MAX_ITERATIONS (:37)

// Largest inf-norm of (q_k - p_k) over four pairs.
double max_linf_4(const double* p1, const double* p2, const double* p3, const double* p4, const double* q1,
                  const double* q2, const double* q3, const double* q4) {
    const double* P[4] = { p1, p2, p3, p4 };
    const double* Q[4] = { q1, q2, q3, q4 };
    double m = 0;
    for (int k = 0; k < 4; k++)
        for (int i = 0; i < 3; i++) m = std::max(m, std::fabs(Q[k][i] - P[k][i]));
    return m;
}

// Per-dimension domain tolerances (t, u, v): the co-domain tolerance over
// the largest extent of the function cube's edges along each parameter.
// p[c] is the function at cube corner c = 4 t + 2 u + v.
void domain_tolerances(const double p[8][3], double tol[3]) {
    const double dl = 3 * max_linf_4(p[0], p[1], p[3], p[2], p[4], p[5], p[7], p[6]);
    const double l0 = 3 * max_linf_4(p[0], p[4], p[5], p[1], p[2], p[6], p[7], p[3]);
    const double l1 = 3 * max_linf_4(p[0], p[4], p[6], p[2], p[1], p[5], p[7], p[3]);
    tol[0] = kTolerance / dl;
    tol[1] = kTolerance / l0;
    tol[2] = kTolerance / l1;
}

// Floating-point filter of the inclusion function (get_numerical_error):
// per-axis max(|x|, 1)^3 times a constant for the function type.
void numerical_error(const double (*V)[3], bool vf, bool use_ms, double err[3]) {
    const double filter = use_ms ? (vf ? 7.549516567451064e-15 : 7.105427357601002e-15)
                                 : (vf ? 6.661338147750939e-15 : 6.217248937900877e-15);
    for (int k = 0; k < 3; k++) {
        double m = std::fabs(V[0][k]);
        for (int i = 1; i < 8; i++) m = std::max(m, std::fabs(V[i][k]));
        const double d = m > 1 ? m : 1;
        err[k] = d * d * d * filter;
    }
}

// The linear-trajectory function at (t, u, v), one axis. V holds the eight
// vertices (start positions then end positions).
double ti_function(const double (*V)[3], bool vf, int k, double t, double u, double v) {
    if (vf) {
        const double p = (V[4][k] - V[0][k]) * t + V[0][k];
        const double t0 = (V[5][k] - V[1][k]) * t + V[1][k];
        const double t1 = (V[6][k] - V[2][k]) * t + V[2][k];
        const double t2 = (V[7][k] - V[3][k]) * t + V[3][k];
        const double q = ((t1 - t0) * u + (t2 - t0) * v) + t0;
        return p - q;
    }
    const double a0 = (V[4][k] - V[0][k]) * t + V[0][k];
    const double a1 = (V[5][k] - V[1][k]) * t + V[1][k];
    const double b0 = (V[6][k] - V[2][k]) * t + V[2][k];
    const double b1 = (V[7][k] - V[3][k]) * t + V[3][k];
    const double pa = (a1 - a0) * u + a0;
    const double pb = (b1 - b0) * v + b0;
    return pb - pa;
}

// Origin_in_function_bounding_box: the function's box over the parameter
// box, from its values at the eight corners, against the origin inflated by
// err + ms. box_in: the function box lies inside that inflated box;
// true_tol: its largest side.
bool origin_in_box(const double (*V)[3], bool vf, const TIBox& b, const double err[3], double ms, bool& box_in,
                   double& true_tol) {
    box_in = true;
    true_tol = 0;
    for (int k = 0; k < 3; k++) {
        double mn = std::numeric_limits<double>::infinity(), mx = -mn;
        for (int c = 0; c < 8; c++) {
            const double f = ti_function(V, vf, k, (c & 4) ? b.hi[0] : b.lo[0], (c & 2) ? b.hi[1] : b.lo[1],
                                         (c & 1) ? b.hi[2] : b.lo[2]);
            mn = std::min(mn, f);
            mx = std::max(mx, f);
        }
        const double eps_ms = err[k] + ms;
        if (mn > eps_ms || mx < -eps_ms) return false;
        if (!(mn >= -eps_ms && mx <= eps_ms)) box_in = false;
        true_tol = std::max(true_tol, mx - mn);
    }
    return true;
}

// interval_root_finder_double_horizontal_tree over [0,1]^3 (t_max = 1, so
// the impact ratio is 1 and every t half overlaps [0, t_max]).
bool ti_root_finder(const double (*V)[3], bool vf, const double tol[3], const double err[3], double ms,
                    double& toi, double& output_tolerance, long* n_boxes) {
    const double co_tol = kTolerance;
    output_tolerance = co_tol;
    double temp_output_tolerance = co_tol;
    std::priority_queue<TIItem, std::vector<TIItem>, TICmp> q;
    q.push(TIItem{ TIBox{ { 0, 0, 0 }, { 1, 1, 1 } }, -1 });
    long refine = 0;
    toi = std::numeric_limits<double>::infinity();
    double temp_toi = toi;
    double toi_skip = 4;  // TOI_SKIP starts above any t
    bool use_skip = false;
    int current_level = -2, box_in_level = -2;
    bool this_level_less_tol = true, find_level_root = false;
    while (!q.empty()) {
        const TIItem it = q.top();
        q.pop();
        TIBox cur = it.b;
        const int level = it.level;
        if (!(cur.lo[0] < toi_skip)) continue;
        if (box_in_level != level) {
            box_in_level = level;
            this_level_less_tol = true;
        }
        refine++;
        if (n_boxes) ++*n_boxes;
        bool box_in;
        double true_tol;
        if (!origin_in_box(V, vf, cur, err, ms, box_in, true_tol)) continue;
        const double w[3] = { cur.hi[0] - cur.lo[0], cur.hi[1] - cur.lo[1], cur.hi[2] - cur.lo[2] };
        const bool tol_condition = true_tol <= co_tol;
        const bool condition1 = w[0] <= tol[0] && w[1] <= tol[1] && w[2] <= tol[2];
        const bool condition2 = box_in && this_level_less_tol;
        if (!tol_condition) this_level_less_tol = false;
        const bool condition3 = this_level_less_tol;
        if (condition1 || condition2 || condition3) {
            toi = cur.lo[0];
            return true;
        }
        if (current_level != level) {
            current_level = level;
            find_level_root = false;
        }
        if (!find_level_root) {
            temp_toi = cur.lo[0];
            temp_output_tolerance = std::max(std::max(std::max(w[0], w[1]), w[2]), co_tol);
            find_level_root = true;
        }
        if (refine > kMaxItr) {
            toi = temp_toi;
            output_tolerance = temp_output_tolerance;
            return true;
        }
        if (tol_condition || box_in) {
            if (cur.lo[0] < toi_skip) toi_skip = cur.lo[0];
            use_skip = true;
            continue;
        }
        // split the dimension with the largest width relative to its tolerance
        const double r0 = w[0] / tol[0], r1 = w[1] / tol[1], r2 = w[2] / tol[2];
        int s;
        if (r0 >= r1 && r0 >= r2) s = 0;
        else if (r1 >= r0 && r1 >= r2) s = 1;
        else s = 2;
        const double mid = (cur.lo[s] + cur.hi[s]) / 2;
        if (!(cur.lo[s] < mid && mid < cur.hi[s])) {  // no dyadic midpoint left: report the box
            toi = cur.lo[0];
            return true;
        }
        TIBox first = cur, second = cur;
        first.hi[s] = mid;
        second.lo[s] = mid;
        if (vf && s != 0) {
            const int o = 3 - s;  // the other barycentric parameter
            if (second.lo[s] + cur.lo[o] <= 1) q.push(TIItem{ second, level + 1 });
            if (first.lo[s] + cur.lo[o] <= 1) q.push(TIItem{ first, level + 1 });
        } else {
            q.push(TIItem{ second, level + 1 });
            q.push(TIItem{ first, level + 1 });
        }
    }
    if (use_skip) {
        toi = toi_skip;
        return true;
    }
    return false;
}

// edgeEdgeCCD_double / vertexFaceCCD_double with err auto, t_max 1,
// ccd_type 1 and no_zero_toi false (a single root-finder pass).
bool ti_ccd(const double (*V)[3], bool vf, double ms, double& toi, double& output_tolerance, long* n_boxes) {
    double err[3];
    numerical_error(V, vf, ms > 0, err);
    // the function at the corners of the unit cube (t, u, v in {0,1}) as
    // differences of the start / end vertices (compute_*_tolerances)
    double p[8][3];
    for (int c = 0; c < 8; c++) {
        const int te = (c & 4) ? 4 : 0;
        const int u = (c >> 1) & 1, v = c & 1;
        for (int k = 0; k < 3; k++) {
            if (vf) {
                // p000 = vs - f0s, p001 = vs - f2s, p010 = vs - f1s, p011 = vs - (f1s + f2s - f0s)
                const double f0 = V[te + 1][k], f1 = V[te + 2][k], f2 = V[te + 3][k];
                double q;
                if (!u && !v) q = f0;
                else if (!u && v) q = f2;
                else if (u && !v) q = f1;
                else q = (f1 + f2) - f0;
                p[c][k] = V[te + 0][k] - q;
            } else {
                // p_uv = edge0 vertex u - edge1 vertex v
                p[c][k] = V[te + u][k] - V[te + 2 + v][k];
            }
        }
    }
    double tol[3];
    domain_tolerances(p, tol);
    return ti_root_finder(V, vf, tol, err, ms, toi, output_tolerance, n_boxes);
}

// ---------------------------------------------------------------- piecewise-linear outer loop
constexpr double kTrajectoryDistanceFactor = 0.5;     // :40
constexpr double kDecreasingDistanceFactor = 0.2;     // :43
constexpr double kDecreasingDistanceMinTime = 1e-6;   // :44
constexpr long kMaxNumSubdivisions = 1000;            // :35
// Not in the reference: a linear CCD that reports an impact at the very
// start of the first piece makes the reference halve that piece again
// (time_of_impact.cpp:428-438, "impossible because distance_t0 > MS_DIST").
// A pair that rests just outside the separation distance (closer than the
// linear CCD's error filter) gets that answer for every piece, so the
// reference halves without end. After this many such halvings an impact at
// t = 0 is reported, the conservative answer.
constexpr int kMaxZeroToiSplits = 64;

// Pose<double>::interpolate (src/physics/pose.tpp:141-148).
void interpolate_pose(const double* p0, const double* p1, double t, double* out) {
    for (int i = 0; i < 6; i++) out[i] = (p1[i] - p0[i]) * t + p0[i];
}

void world_point(const double* pose, const double* v, double* w) {
    const Mat3d R = rotation_matrix(pose + 3);
    world_vertex(R, pose, v, w);
}

// Upper bound over s in [0,1] of |x(s) - lerp(x(0), x(1), s)| where x is
// the rigid trajectory of body-local point v between poses pa and pb
// (the interval bound of time_of_impact.cpp:534-560).
double deviation_bound(const double* pa, const double* pb, const double* v) {
    const Interval s(0, 1);
    const PoseI P = interpolate(pa, pb, s);
    const Mat3I R = rotation_matrix(P.r);
    const Vec3I x = world_vertex(R, P.p, v);
    double wa[3], wb[3];
    world_point(pa, v, wa);
    world_point(pb, v, wb);
    Interval d[3];
    for (int k = 0; k < 3; k++) d[k] = x[k] - ((Interval(wb[k]) - Interval(wa[k])) * s + Interval(wa[k]));
    // Eigen's norm of a fixed-size vector: sqrt(x0*x0 + (x1*x1 + x2*x2))
    const Interval sq = d[0] * d[0] + (d[1] * d[1] + d[2] * d[2]);
    return sqrt(sq).upper();
}

// Shared outer loop. nA / nB local vertices of body A / B; EE: A = edge a
// (2), B = edge b (2); FV: A = vertex (1), B = face (3).
bool pl_toi(bool vf, const double* const* LA, int nA, const double* const* LB, int nB, const double* pA0,
            const double* pA1, const double* pB0, const double* pB1, double earliest_toi, double msd, double& toi,
            long* n_pieces, long* n_boxes) {
    auto world_all = [&](const double* pa, const double* pb, double X[4][3]) {
        int s = 0;
        for (int i = 0; i < nA; i++) world_point(pa, LA[i], X[s++]);
        for (int i = 0; i < nB; i++) world_point(pb, LB[i], X[s++]);
    };
    const int type = vf ? FV : EE;
    double X0[4][3];
    world_all(pA0, pB0, X0);
    const double distance_t0 = std::sqrt(element_distance_sq(type, &X0[0][0]));
    if (distance_t0 <= msd) {
        toi = 0;
        return true;
    }
    toi = std::numeric_limits<double>::infinity();
    bool is_impacting = false;
    double pa_ti0[6], pb_ti0[6];
    std::memcpy(pa_ti0, pA0, sizeof(pa_ti0));
    std::memcpy(pb_ti0, pB0, sizeof(pb_ti0));
    double ti0 = 0;
    std::stack<double> ts;
    ts.push(earliest_toi);
    long num_subdivisions = 1;
    int zero_splits = 0;
    while (!ts.empty()) {
        const double ti1 = ts.top();
        if (n_pieces) ++*n_pieces;
        double pa_ti1[6], pb_ti1[6];
        interpolate_pose(pA0, pA1, ti1, pa_ti1);
        interpolate_pose(pB0, pB1, ti1, pb_ti1);
        // both element types measure the distance at the piece's start
        // (time_of_impact.cpp:335-339 and :529-533)
        double Xd[4][3];
        world_all(pa_ti0, pb_ti0, Xd);
        const double distance_ti0 = std::sqrt(element_distance_sq(type, &Xd[0][0]));
        if (distance_ti0 < kDecreasingDistanceFactor * distance_t0 && ti0 >= kDecreasingDistanceMinTime) {
            toi = ti0;
            is_impacting = true;
            break;
        }
        double dev_a = 0, dev_b = 0;
        for (int i = 0; i < nA; i++) dev_a = std::max(dev_a, deviation_bound(pa_ti0, pa_ti1, LA[i]));
        for (int i = 0; i < nB; i++) dev_b = std::max(dev_b, deviation_bound(pb_ti0, pb_ti1, LB[i]));
        double min_distance = dev_a + dev_b;
        if (min_distance >= kTrajectoryDistanceFactor * distance_ti0
            && (num_subdivisions < kMaxNumSubdivisions || ti0 == 0)) {
            ts.push((ti1 + ti0) / 2);
            num_subdivisions++;
            continue;
        }
        min_distance += msd;
        double V[8][3], Xs[4][3], Xe[4][3];
        world_all(pa_ti0, pb_ti0, Xs);
        world_all(pa_ti1, pb_ti1, Xe);
        for (int i = 0; i < 4; i++)
            for (int k = 0; k < 3; k++) {
                V[i][k] = Xs[i][k];
                V[4 + i][k] = Xe[i][k];
            }
        double ltoi, out_tol;
        is_impacting = ti_ccd(V, vf, min_distance, ltoi, out_tol, n_boxes);
        if (is_impacting) {
            toi = (ti1 - ti0) * ltoi + ti0;
            if (toi == 0) {
                if (++zero_splits > kMaxZeroToiSplits) break;  // impact at t = 0
                ts.push((ti1 + ti0) / 2);
                num_subdivisions++;
                continue;
            }
            break;
        }
        toi = std::numeric_limits<double>::infinity();
        ts.pop();
        ti0 = ti1;
        std::memcpy(pa_ti0, pa_ti1, sizeof(pa_ti0));
        std::memcpy(pb_ti0, pb_ti1, sizeof(pb_ti0));
    }
    return is_impacting;
}

}  // namespace

bool pl_edge_edge_toi(const EEQuery& q, double& toi, double earliest_toi, double msd, long* n_pieces,
                      long* n_boxes) {
    const double* LA[2] = { q.a0, q.a1 };
    const double* LB[2] = { q.b0, q.b1 };
    return pl_toi(false, LA, 2, LB, 2, q.poseA_t0, q.poseA_t1, q.poseB_t0, q.poseB_t1, earliest_toi, msd, toi,
                  n_pieces, n_boxes);
}

bool pl_face_vertex_toi(const FVQuery& q, double& toi, double earliest_toi, double msd, long* n_pieces,
                        long* n_boxes) {
    const double* LA[1] = { q.v };
    const double* LB[3] = { q.f0, q.f1, q.f2 };
    return pl_toi(true, LA, 1, LB, 3, q.poseA_t0, q.poseA_t1, q.poseB_t0, q.poseB_t1, earliest_toi, msd, toi,
                  n_pieces, n_boxes);
}

}  // namespace oracle

using namespace oracle;

namespace {
void unpack_query(const double* q, double* a, double* b, double* c, double* d, double* p) {
    std::memcpy(a, q + 0, 3 * sizeof(double));
    std::memcpy(b, q + 3, 3 * sizeof(double));
    std::memcpy(c, q + 6, 3 * sizeof(double));
    std::memcpy(d, q + 9, 3 * sizeof(double));
    std::memcpy(p, q + 12, 24 * sizeof(double));
}
}  // namespace

extern "C" {

double oracle_element_distance_sq(int type, const double* X) { return element_distance_sq(type, X); }

// Linear CCD of eight vertices V[8][3] (start a0 a1 b0 b1 / v f0 f1 f2, then
// end); returns 1 on impact with *toi in [0,1], else 0.
int oracle_ti_edge_edge(const double* V, double ms, double* toi, double* out_tol, long* n_boxes) {
    return ti_ccd(reinterpret_cast<const double(*)[3]>(V), false, ms, *toi, *out_tol, n_boxes) ? 1 : 0;
}
int oracle_ti_vertex_face(const double* V, double ms, double* toi, double* out_tol, long* n_boxes) {
    return ti_ccd(reinterpret_cast<const double(*)[3]>(V), true, ms, *toi, *out_tol, n_boxes) ? 1 : 0;
}

// Piecewise-linear TOI of one packed query (the layout of the rigid CCD
// query entry points: 4 body-local vertices, then poses A t0, A t1, B t0, B t1).
int oracle_pl_ee_toi(const double* q, double earliest_toi, double msd, double* toi, long* n_pieces,
                     long* n_boxes) {
    EEQuery e;
    double p[24];
    unpack_query(q, e.a0, e.a1, e.b0, e.b1, p);
    std::memcpy(e.poseA_t0, p, 6 * sizeof(double));
    std::memcpy(e.poseA_t1, p + 6, 6 * sizeof(double));
    std::memcpy(e.poseB_t0, p + 12, 6 * sizeof(double));
    std::memcpy(e.poseB_t1, p + 18, 6 * sizeof(double));
    return pl_edge_edge_toi(e, *toi, earliest_toi, msd, n_pieces, n_boxes) ? 1 : 0;
}
int oracle_pl_fv_toi(const double* q, double earliest_toi, double msd, double* toi, long* n_pieces,
                     long* n_boxes) {
    FVQuery f;
    double p[24];
    unpack_query(q, f.v, f.f0, f.f1, f.f2, p);
    std::memcpy(f.poseA_t0, p, 6 * sizeof(double));
    std::memcpy(f.poseA_t1, p + 6, 6 * sizeof(double));
    std::memcpy(f.poseB_t0, p + 12, 6 * sizeof(double));
    std::memcpy(f.poseB_t1, p + 18, 6 * sizeof(double));
    return pl_face_vertex_toi(f, *toi, earliest_toi, msd, n_pieces, n_boxes) ? 1 : 0;
}

}  // extern "C"
