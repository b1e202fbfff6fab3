// Piecewise-linear CCD (TrajectoryType::PIECEWISE_LINEAR) on the device.
//
// Replaces compute_piecewise_linear_{edge_edge,face_vertex}_time_of_impact
// (src/ccd/piecewise_linear/time_of_impact.cpp:296-649) reached from
// edge_edge_ccd / face_vertex_ccd (src/ccd/ccd.cpp:271-274, 339-342), and
// the linear CCD those call, Tight-Inclusion's edgeEdgeCCD_double /
// vertexFaceCCD_double (external, tag b913416) with the reference's
// arguments: error filter from the vertices, minimum separation = the
// piece's trajectory deviation bound + the caller's separation, co-domain
// tolerance 1e-6, t_max 1, at most 1e6 box refinements, breadth-first
// ("horizontal tree") traversal.
//
// One lane runs one query: the outer loop over trajectory pieces (a stack
// of piece ends), per piece an interval bound of how far the rigid
// trajectory strays from the straight one, then the breadth-first box
// traversal over (t, u, v). The traversal's priority queue (level first,
// then the lower end of t) is a binary heap with the sift-up / sift-down
// steps of libstdc++'s std::priority_queue, so boxes with equal keys are
// served in the same order as in the reference's build. The first pass
// runs one query per wavefront with its piece stack and heap in LDS; later
// passes run one query per lane with both in the lane's slice of an arena in
// HBM. A query whose heap or stack would outgrow its space stops with
// status 2 and is run again in the next pass with more space. A heap of 2^21
// entries holds any traversal: at most 1e6 refinements push at most two
// boxes each.
#include "rigc_internal.hpp"
#include "rigc_barrier.hpp"
#include "rigc_geom.hpp"
#include "rigc_rigid_iv.hpp"

namespace rigc {

namespace {

constexpr double kTiTolerance = 1e-6;           // LINEAR_CCD_TOL (time_of_impact.cpp:38)
constexpr long long kTiMaxItr = 1000000;        // LINEAR_CCD_MAX_ITERATIONS (:37)
constexpr double kTrajectoryDistanceFactor = 0.5;
constexpr double kDecreasingDistanceFactor = 0.2;
constexpr double kDecreasingDistanceMinTime = 1e-6;
constexpr long long kMaxNumSubdivisions = 1000;
// A linear CCD impact at the very start of the first piece halves that piece
// again (time_of_impact.cpp:428-438). A pair resting just outside the
// separation distance gets that answer for every piece, and the reference
// halves without end; after this many halvings an impact at t = 0 is
// reported (the conservative answer; not in the reference).
constexpr int kMaxZeroToiSplits = 64;
constexpr int kItem = 7;  // t_lo t_hi u_lo u_hi v_lo v_hi level

struct Lane {
    double* tstack;
    int tcap;
    double* heap;
    long long hcap;
    unsigned long long boxes, pieces;
    bool wave;  // the whole wavefront runs this query (uniform control flow)
};

// Tight-Inclusion's queue order: comp(a, b) = a is served after b.
RIGC_DI bool ti_after(const double* a, const double* b) {
    if (a[6] != b[6]) return a[6] >= b[6];
    return !(a[0] < b[0]);
}
RIGC_DI void copy_item(double* d, const double* s) {
#pragma unroll
    for (int i = 0; i < kItem; i++) d[i] = s[i];
}
// std::push_heap's sift-up of `value` from `hole`.
RIGC_DI void sift_up(double* h, long long hole, const double* value) {
    while (hole > 0) {
        const long long parent = (hole - 1) / 2;
        if (!ti_after(h + kItem * parent, value)) break;
        copy_item(h + kItem * hole, h + kItem * parent);
        hole = parent;
    }
    copy_item(h + kItem * hole, value);
}
RIGC_DI void heap_push(double* h, long long& n, const double* item) {
    const long long hole = n++;
    sift_up(h, hole, item);
}
// std::pop_heap + pop_back: top -> `top`, heap of n - 1 entries.
RIGC_DI void heap_pop(double* h, long long& n, double* top) {
    copy_item(top, h);
    const long long len = n - 1;
    if (len > 0) {
        double value[kItem];
        copy_item(value, h + kItem * len);
        long long hole = 0, second = 0;
        while (second < (len - 1) / 2) {
            second = 2 * (second + 1);
            if (ti_after(h + kItem * second, h + kItem * (second - 1))) second--;
            copy_item(h + kItem * hole, h + kItem * second);
            hole = second;
        }
        if ((len & 1) == 0 && second == (len - 2) / 2) {
            second = 2 * (second + 1);
            copy_item(h + kItem * hole, h + kItem * (second - 1));
            hole = second - 1;
        }
        sift_up(h, hole, value);
    }
    n = len;
}

// The linear-trajectory function at (t, u, v), axis k; V = 8 vertices
// (start, then end): EE pb - pa, FV p - q.
RIGC_DI double ti_function(const double (*V)[3], bool vf, int k, double t, double u, double v) {
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

// Function box over the parameter box (corner values) against the origin
// inflated by err + ms.
RIGC_DI bool origin_in_box(const double (*V)[3], bool vf, const double* b, const double err[3], double ms,
                           bool& box_in, double& true_tol) {
    box_in = true;
    true_tol = 0;
    for (int k = 0; k < 3; k++) {
        double mn = HUGE_VAL, mx = -HUGE_VAL;
        for (int c = 0; c < 8; c++) {
            const double f = ti_function(V, vf, k, (c & 4) ? b[1] : b[0], (c & 2) ? b[3] : b[2],
                                         (c & 1) ? b[5] : b[4]);
            mn = fmin(mn, f);
            mx = fmax(mx, f);
        }
        const double eps_ms = err[k] + ms;
        if (mn > eps_ms || mx < -eps_ms) return false;
        if (!(mn >= -eps_ms && mx <= eps_ms)) box_in = false;
        true_tol = fmax(true_tol, mx - mn);
    }
    return true;
}

// origin_in_box with the wavefront's lanes sharing the 24 function values:
// lane l takes value j = l mod 24 (axis j / 8, corner j mod 8), so every
// aligned group of 8 lanes holds one axis's corners; a butterfly over each
// group gives that axis's minimum and maximum (fmin / fmax are exact and
// order-free), and lanes 0, 8 and 16 hand them to all. Same answer as the
// loop above. Every lane of the wavefront must call it with the same box.
RIGC_DI bool origin_in_box_wave(const double (*V)[3], bool vf, const double* b, const double err[3], double ms,
                                bool& box_in, double& true_tol) {
    const int j = (int)(threadIdx.x & 63) % 24;
    const int kk = j >> 3, c = j & 7;
    const double f = ti_function(V, vf, kk, (c & 4) ? b[1] : b[0], (c & 2) ? b[3] : b[2], (c & 1) ? b[5] : b[4]);
    double mn = f, mx = f;
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) {
        mn = fmin(mn, __shfl_xor(mn, o));
        mx = fmax(mx, __shfl_xor(mx, o));
    }
    double amn[3], amx[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        amn[k] = __shfl(mn, 8 * k);
        amx[k] = __shfl(mx, 8 * k);
    }
    box_in = true;
    true_tol = 0;
    for (int k = 0; k < 3; k++) {
        const double eps_ms = err[k] + ms;
        if (amn[k] > eps_ms || amx[k] < -eps_ms) return false;
        if (!(amn[k] >= -eps_ms && amx[k] <= eps_ms)) box_in = false;
        true_tol = fmax(true_tol, amx[k] - amn[k]);
    }
    return true;
}

RIGC_DI double max_linf_4(const double (*p)[3], int a0, int a1, int a2, int a3, int b0, int b1, int b2, int b3) {
    const int A[4] = { a0, a1, a2, a3 }, B[4] = { b0, b1, b2, b3 };
    double m = 0;
    for (int j = 0; j < 4; j++)
        for (int i = 0; i < 3; i++) m = fmax(m, fabs(p[B[j]][i] - p[A[j]][i]));
    return m;
}

// Linear CCD of one piece. Returns 1 impact (toi set), 0 none, 2 heap full.
RIGC_DI int ti_ccd(Lane& L, const double (*V)[3], bool vf, double ms, double& toi) {
    // error filter (get_numerical_error)
    const double filter = ms > 0 ? (vf ? 7.549516567451064e-15 : 7.105427357601002e-15)
                                 : (vf ? 6.661338147750939e-15 : 6.217248937900877e-15);
    double err[3];
    for (int k = 0; k < 3; k++) {
        double m = fabs(V[0][k]);
        for (int i = 1; i < 8; i++) m = fmax(m, fabs(V[i][k]));
        const double d = m > 1 ? m : 1;
        err[k] = d * d * d * filter;
    }
    // function at the unit cube's corners c = 4t + 2u + v, then the
    // per-parameter domain tolerances (compute_*_tolerances)
    double p[8][3];
    for (int c = 0; c < 8; c++) {
        const int te = (c & 4) ? 4 : 0;
        const int u = (c >> 1) & 1, v = c & 1;
        for (int k = 0; k < 3; k++) {
            if (vf) {
                const double f0 = V[te + 1][k], f1 = V[te + 2][k], f2 = V[te + 3][k];
                const double q = (!u && !v) ? f0 : (!u ? f2 : (!v ? f1 : (f1 + f2) - f0));
                p[c][k] = V[te][k] - q;
            } else {
                p[c][k] = V[te + u][k] - V[te + 2 + v][k];
            }
        }
    }
    double tol[3];
    tol[0] = kTiTolerance / (3 * max_linf_4(p, 0, 1, 3, 2, 4, 5, 7, 6));
    tol[1] = kTiTolerance / (3 * max_linf_4(p, 0, 4, 5, 1, 2, 6, 7, 3));
    tol[2] = kTiTolerance / (3 * max_linf_4(p, 0, 4, 6, 2, 1, 5, 7, 3));

    const double co_tol = kTiTolerance;
    double* h = L.heap;
    long long n = 0;
    {
        const double root[kItem] = { 0, 1, 0, 1, 0, 1, -1 };
        heap_push(h, n, root);
    }
    long long refine = 0;
    double temp_toi = HUGE_VAL;
    double toi_skip = 4;
    bool use_skip = false;
    double current_level = -2, box_in_level = -2;
    bool this_level_less_tol = true, find_level_root = false;
    toi = HUGE_VAL;
    while (n > 0) {
        double cur[kItem];
        heap_pop(h, n, cur);
        const double level = cur[6];
        if (!(cur[0] < toi_skip)) continue;
        if (box_in_level != level) {
            box_in_level = level;
            this_level_less_tol = true;
        }
        refine++;
        L.boxes++;
        bool box_in;
        double true_tol;
        const bool in = L.wave ? origin_in_box_wave(V, vf, cur, err, ms, box_in, true_tol)
                               : origin_in_box(V, vf, cur, err, ms, box_in, true_tol);
        if (!in) continue;
        const double w0 = cur[1] - cur[0], w1 = cur[3] - cur[2], w2 = cur[5] - cur[4];
        const bool tol_condition = true_tol <= co_tol;
        const bool condition1 = w0 <= tol[0] && w1 <= tol[1] && w2 <= tol[2];
        const bool condition2 = box_in && this_level_less_tol;
        if (!tol_condition) this_level_less_tol = false;
        const bool condition3 = this_level_less_tol;
        if (condition1 || condition2 || condition3) {
            toi = cur[0];
            return 1;
        }
        if (current_level != level) {
            current_level = level;
            find_level_root = false;
        }
        if (!find_level_root) {
            temp_toi = cur[0];
            find_level_root = true;
        }
        if (refine > kTiMaxItr) {
            toi = temp_toi;
            return 1;
        }
        if (tol_condition || box_in) {
            if (cur[0] < toi_skip) toi_skip = cur[0];
            use_skip = true;
            continue;
        }
        const double r0 = w0 / tol[0], r1 = w1 / tol[1], r2 = w2 / tol[2];
        const int s = (r0 >= r1 && r0 >= r2) ? 0 : ((r1 >= r0 && r1 >= r2) ? 1 : 2);
        const double mid = (cur[2 * s] + cur[2 * s + 1]) / 2;
        if (!(cur[2 * s] < mid && mid < cur[2 * s + 1])) {
            toi = cur[0];
            return 1;
        }
        if (n + 2 > L.hcap) return 2;
        double first[kItem], second[kItem];
        copy_item(first, cur);
        copy_item(second, cur);
        first[2 * s + 1] = mid;
        second[2 * s] = mid;
        first[6] = second[6] = level + 1;
        if (vf && s != 0) {
            const int o = 3 - s;
            if (second[2 * s] + cur[2 * o] <= 1) heap_push(h, n, second);
            if (first[2 * s] + cur[2 * o] <= 1) heap_push(h, n, first);
        } else {
            heap_push(h, n, second);
            heap_push(h, n, first);
        }
    }
    if (use_skip) {
        toi = toi_skip;
        return 1;
    }
    return 0;
}

RIGC_DI void world_point(const double* pose, const double* v, double* w) {
    double R[9];
    rotation_d(pose + 3, R);
    world_vertex_d(R, pose, v, w);
}

// Upper bound of |x(s) - lerp(x(0), x(1), s)| over s in [0,1] for the rigid
// trajectory of local point v between poses pa and pb.
RIGC_DI double deviation_bound(const double* pa, const double* pb, const double* v) {
    const Iv s = iv(0.0, 1.0);
    Iv P[3], R[3][3], x[3];
    pose_iv(pa, pb, s, P, R);
    world_vertex_iv(R, P, v, x);
    double wa[3], wb[3];
    world_point(pa, v, wa);
    world_point(pb, v, wb);
    Iv d[3];
    for (int k = 0; k < 3; k++) d[k] = x[k] - ((iv(wb[k]) - iv(wa[k])) * s + iv(wa[k]));
    const Iv sq = d[0] * d[0] + (d[1] * d[1] + d[2] * d[2]);
    return isqrt(sq).hi;
}

RIGC_DI double element_distance(bool vf, const double (*X)[3]) {
    DistForm f;
    constraint_form(vf ? CON_FV : CON_EE, &X[0][0], f);
    return sqrt(f.d);
}

// One query. loc[0..3]: local vertices (A's first: EE a0 a1 b0 b1, FV v f0
// f1 f2); nA = 2 (EE) or 1 (FV). Returns 1 impact, 0 none, 2 slice full.
RIGC_DI int pl_query(Lane& L, bool vf, const double (*loc)[3], const double* pA0, const double* pA1,
                     const double* pB0, const double* pB1, double earliest_toi, double msd, double& toi) {
    const int nA = vf ? 1 : 2;
    auto world_all = [&](const double* pa, const double* pb, double (*X)[3]) {
        for (int i = 0; i < 4; i++) world_point(i < nA ? pa : pb, loc[i], X[i]);
    };
    double X[4][3];
    world_all(pA0, pB0, X);
    const double distance_t0 = element_distance(vf, X);
    if (distance_t0 <= msd) {
        toi = 0;
        return 1;
    }
    toi = HUGE_VAL;
    double pa0[6], pb0[6];
    for (int i = 0; i < 6; i++) {
        pa0[i] = pA0[i];
        pb0[i] = pB0[i];
    }
    double ti0 = 0;
    int ntop = 0;
    L.tstack[ntop++] = earliest_toi;
    long long num_subdivisions = 1;
    int zero_splits = 0;
    int result = 0;
    while (ntop > 0) {
        const double ti1 = L.tstack[ntop - 1];
        L.pieces++;
        double pa1[6], pb1[6];
        for (int i = 0; i < 6; i++) {
            pa1[i] = (pA1[i] - pA0[i]) * ti1 + pA0[i];
            pb1[i] = (pB1[i] - pB0[i]) * ti1 + pB0[i];
        }
        // both element types measure the distance at the piece's start
        world_all(pa0, pb0, X);
        const double distance_ti0 = element_distance(vf, X);
        if (distance_ti0 < kDecreasingDistanceFactor * distance_t0 && ti0 >= kDecreasingDistanceMinTime) {
            toi = ti0;
            return 1;
        }
        double dev_a = 0, dev_b = 0;
        for (int i = 0; i < 4; i++) {
            const double d = deviation_bound(i < nA ? pa0 : pb0, i < nA ? pa1 : pb1, loc[i]);
            if (i < nA) dev_a = fmax(dev_a, d);
            else dev_b = fmax(dev_b, d);
        }
        double min_distance = dev_a + dev_b;
        if (min_distance >= kTrajectoryDistanceFactor * distance_ti0
            && (num_subdivisions < kMaxNumSubdivisions || ti0 == 0)) {
            if (ntop == L.tcap) return 2;
            L.tstack[ntop++] = (ti1 + ti0) / 2;
            num_subdivisions++;
            continue;
        }
        min_distance += msd;
        double V[8][3];
        world_all(pa0, pb0, V);
        world_all(pa1, pb1, V + 4);
        double ltoi;
        result = ti_ccd(L, V, vf, min_distance, ltoi);
        if (result == 2) return 2;
        if (result == 1) {
            toi = (ti1 - ti0) * ltoi + ti0;
            if (toi == 0) {
                if (++zero_splits > kMaxZeroToiSplits) return 1;  // impact at t = 0
                if (ntop == L.tcap) return 2;
                L.tstack[ntop++] = (ti1 + ti0) / 2;
                num_subdivisions++;
                continue;
            }
            return 1;
        }
        toi = HUGE_VAL;
        ntop--;
        ti0 = ti1;
        for (int i = 0; i < 6; i++) {
            pa0[i] = pa1[i];
            pb0[i] = pb1[i];
        }
    }
    return result;
}

// Query i of the list: gathers its vertices and bodies, runs it, writes its
// result (writer lanes only) and keeps the lane's counters.
RIGC_DI void run_query(const PlccdArgs& a, Lane& L, long long i, bool writer) {
    const long long q = a.query_ids ? (long long)a.query_ids[i] : i;
    const bool vf = q >= a.n_ee;
    double loc[4][3];
    int bodyA, bodyB;
    if (!vf) {
        const int ea = a.ee[2 * q], eb = a.ee[2 * q + 1];
        const int vid[4] = { a.E[2 * ea], a.E[2 * ea + 1], a.E[2 * eb], a.E[2 * eb + 1] };
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 3; k++) loc[j][k] = a.V[3 * vid[j] + k];
        bodyA = a.e_body[ea];
        bodyB = a.e_body[eb];
    } else {
        const long long r = q - a.n_ee;
        const int f = a.fv[2 * r], v = a.fv[2 * r + 1];
        const int vid[4] = { v, a.F[3 * f], a.F[3 * f + 1], a.F[3 * f + 2] };
        for (int j = 0; j < 4; j++)
            for (int k = 0; k < 3; k++) loc[j][k] = a.V[3 * vid[j] + k];
        bodyA = a.v_body[v];
        bodyB = a.f_body[f];
    }
    double toi;
    const unsigned long long boxes0 = L.boxes, pieces0 = L.pieces;
    const int st = pl_query(L, vf, loc, a.poses_t0 + 6 * bodyA, a.poses_t1 + 6 * bodyA, a.poses_t0 + 6 * bodyB,
                            a.poses_t1 + 6 * bodyB, a.earliest_toi, a.msd, toi);
    if (st == 2) {  // counted when the query is run again with a larger slice
        L.boxes = boxes0;
        L.pieces = pieces0;
    }
    if (writer) {
        a.toi_out[q] = st == 1 ? toi : HUGE_VAL;
        a.hit_out[q] = (signed char)st;
    }
}

// One query per lane, piece stack and heap in the lane's arena slice (HBM).
__global__ __launch_bounds__(256) void plccd_kernel(PlccdArgs a) {
    const long long lane = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long lanes = a.n_lanes;  // the arena holds n_lanes slices
    if (lane >= lanes) return;
    Lane L;
    L.tcap = a.tstack_cap;
    L.tstack = a.arena + lane * ((long long)a.tstack_cap + (long long)kItem * a.heap_cap);
    L.heap = L.tstack + a.tstack_cap;
    L.hcap = a.heap_cap;
    L.boxes = 0;
    L.pieces = 0;
    L.wave = false;
    for (long long i = lane; i < a.n_queries; i += lanes) run_query(a, L, i, true);
    if (L.boxes) atomicAdd(a.counters + 0, L.boxes);
    if (L.pieces) atomicAdd(a.counters + 1, L.pieces);
}

// One query per wavefront, piece stack and heap in LDS: every lane runs the
// same query on the same LDS data (uniform control flow, identical values),
// the lanes share each box's 24 function values (origin_in_box_wave), lane
// 0 writes the result. A heap operation then waits on LDS instead of
// HBM; queries that outgrow the LDS heap go to the arena passes.
constexpr int kWaveStack = 64;
constexpr int kWaveHeap = 512;
__global__ __launch_bounds__(64) void plccd_wave_kernel(PlccdArgs a) {
    __shared__ double s_tstack[kWaveStack];
    __shared__ double s_heap[kItem * kWaveHeap];
    Lane L;
    L.tcap = kWaveStack;
    L.tstack = s_tstack;
    L.heap = s_heap;
    L.hcap = kWaveHeap;
    L.boxes = 0;
    L.pieces = 0;
    L.wave = true;
    const bool writer = threadIdx.x == 0;
    for (long long i = blockIdx.x; i < a.n_queries; i += gridDim.x) run_query(a, L, i, writer);
    if (writer && L.boxes) atomicAdd(a.counters + 0, L.boxes);
    if (writer && L.pieces) atomicAdd(a.counters + 1, L.pieces);
}

}  // namespace

hipError_t launch_plccd(const PlccdArgs& a, int grid, hipStream_t st) {
    if (a.n_queries <= 0) return hipSuccess;
    hipLaunchKernelGGL(plccd_kernel, dim3(grid), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_plccd_wave(const PlccdArgs& a, int grid, hipStream_t st) {
    if (a.n_queries <= 0) return hipSuccess;
    hipLaunchKernelGGL(plccd_wave_kernel, dim3(grid), dim3(64), 0, st, a);
    return hipGetLastError();
}

}  // namespace rigc
