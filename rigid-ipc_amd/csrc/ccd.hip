// Rigid-trajectory interval CCD narrow phase for gfx950.
//
// Restates, per candidate, compute_edge_edge_time_of_impact /
// compute_face_vertex_time_of_impact (src/ccd/rigid/time_of_impact.cpp:158-310)
// driving interval_root_finder (src/interval/interval_root_finder.cpp:97-203)
// over the inclusion functions of src/ccd/rigid/rigid_trajectory_aabb.cpp:70-102.
//
// The root finder's answer is the least t.lower over all accepted boxes of
// the bisection tree (boxes reached through ancestors whose enclosure holds
// zero, and not skipped because their t.lower is already no better than an
// accepted box). That minimum does not depend on the order in which the
// tree is walked, so the tree of one query may be walked by many lanes at
// once, and pruning with the running minimum over all queries
// (earliest-TOI mode) returns exactly the minimum of the per-query answers.
//
// Execution model: a persistent grid of wavefronts. Every lane walks one
// subtree depth-first without a box stack: a box is a dyadic sub-box of
// [0,1]^3 held as an integer (index, level) per dimension, and the path
// from the subtree root is three 192-bit masks (split dimension t / alpha
// per level, and whether the second half of that level is taken); popping
// the next box is a find-last-zero plus popcounts. The t-only part of the
// inclusion function (interval poses, interval Rodrigues matrices, world
// vertices) is computed by the wavefront together for the distinct
// (query, t-interval) pairs of a step, and cached per lane while only
// alpha/beta split. Idle lanes first take new queries from a global
// counter; once that pool is empty, each busy lane with an untaken second
// half on its path hands
// the shallowest one (its largest pending subtree) to an idle lane of the
// same wavefront through cross-lane shuffles, so a deep query ends up
// bisected by up to 64 lanes. Per-query minima live in global memory and
// are combined with 64-bit atomicMin on the bit pattern (t >= 0).
//
// To put more than one wavefront on a deep query, the walk runs in passes:
// once the pool of a pass is empty, a lane that has evaluated walk_budget
// boxes since it took its subtree writes every pending subtree of its path
// (the second halves still to visit, and its next box) to a task buffer
// and goes idle; the next pass takes those tasks as its pool and spreads
// them over the whole grid. Each pass ends when its lanes have finished or
// exported, so no lane ever waits for another.
#include <algorithm>

#include "rigc_interval.hpp"
#include "rigc_internal.hpp"
#include "rigc_rigid_iv.hpp"

namespace rigc {
namespace {

constexpr int MAX_DEPTH = 192;
constexpr int THR_REFRESH = 8;
constexpr int SPEC_LEVELS = 5;  // most speculative levels below an owner's box (62 boxes: one 64-bit mask)

// Three named words rather than an array: the masks are read and written at
// bit positions known only at run time, and a dynamically indexed array
// would live in scratch memory instead of registers.
struct Bits192 {
    unsigned long long w0, w1, w2;
};
RIGC_DI unsigned long long word(const Bits192& b, int wi) { return wi == 0 ? b.w0 : (wi == 1 ? b.w1 : b.w2); }
RIGC_DI void clear_bits(Bits192& b) { b.w0 = b.w1 = b.w2 = 0ULL; }
RIGC_DI void bset(Bits192& b, int k, bool v) {
    const unsigned long long m = 1ULL << (k & 63);
    if (k < 64) b.w0 = v ? (b.w0 | m) : (b.w0 & ~m);
    else if (k < 128) b.w1 = v ? (b.w1 | m) : (b.w1 & ~m);
    else b.w2 = v ? (b.w2 | m) : (b.w2 & ~m);
}
RIGC_DI bool bget(const Bits192& b, int k) {
    const unsigned long long w = k < 64 ? b.w0 : (k < 128 ? b.w1 : b.w2);
    return (w >> (k & 63)) & 1ULL;
}
// mask of bits [lo, hi) within word wi
RIGC_DI unsigned long long range_mask(int wi, int lo, int hi) {
    const int a = max(lo - 64 * wi, 0), z = min(hi - 64 * wi, 64);
    if (z <= a) return 0ULL;
    const unsigned long long top = (z == 64) ? ~0ULL : ((1ULL << z) - 1ULL);
    const unsigned long long bot = (a == 0) ? 0ULL : ((1ULL << a) - 1ULL);
    return top & ~bot;
}
RIGC_DI int popcount_range(const Bits192& b, int lo, int hi) {
    return __popcll(b.w0 & range_mask(0, lo, hi)) + __popcll(b.w1 & range_mask(1, lo, hi))
        + __popcll(b.w2 & range_mask(2, lo, hi));
}
// Largest level L < depth whose bit is clear, or -1.
RIGC_DI int last_zero_below(const Bits192& b, int depth) {
#pragma unroll
    for (int wi = 2; wi >= 0; wi--) {
        const unsigned long long m = ~word(b, wi) & range_mask(wi, 0, depth);
        if (m) return 64 * wi + 63 - __clzll(m);
    }
    return -1;
}

RIGC_DI double dyad(unsigned long long i, int l) { return ldexp((double)i, -l); }

// a[k] += v and a[k] = 2 a[k] + bit for a dimension k known only at run
// time, written as selects so that the three-element arrays stay in
// registers (a dynamic index would put them in scratch memory).
template <typename T>
RIGC_DI void add_at(T (&a)[3], int k, T v) {
#pragma unroll
    for (int i = 0; i < 3; i++) a[i] = (i == k) ? a[i] + v : a[i];
}
template <typename T>
RIGC_DI void descend_at(T (&a)[3], int k, T bit) {
#pragma unroll
    for (int i = 0; i < 3; i++) a[i] = (i == k) ? (a[i] << 1) + bit : a[i];
}

struct Query {
    int kind, bA, bB, sc;
    int g[4];  // EE: ea0 ea1 eb0 eb1; FV: v f0 f1 f2 (global vertex ids)
};

// t-dependent part of the inclusion function.
// EE: A0 = ea0(t), DA = ea1(t) - ea0(t), B0 = eb0(t), DB = eb1(t) - eb0(t)
// FV: A0 = v(t),   DA = f1(t) - f0(t),   B0 = f0(t),  DB = f2(t) - f0(t)
//
// It is computed by the wavefront together. The boxes a wavefront
// evaluates in one step share few distinct (query, t-interval) pairs (the
// speculative subtree of a box splits t on a few of its levels only), and
// every value of the t part is a pure function of that pair. The pairs
// needing a new t part are numbered (up to COOP_ITEMS at a time) and the
// work is spread over the lanes in stages, each a short instruction stream
// run once for all pairs: the interval rotation vector and its two norms per
// (pair, body); the monotone-piece ends of sinc per (pair, body, angle,
// piece) and the piece beyond its maximum where a norm reaches it; one
// rotation row per (pair, body, row); the position and one world vertex per
// (pair, vertex). Each lane then forms its pair's differences. The operations and their order per
// value are those of pose_iv, rotation_iv (sinc_iv) and world_vertex_iv,
// so the result is the same, bit for bit, as evaluating each box alone.
// The stages' scratch (rotation vectors, norms, sinc pieces) is dead once
// the rotation rows are formed, so the world vertices of the last stage
// share its space: 19.8 KB per wavefront, which lets two 256-thread blocks
// share a CU's LDS (two wavefronts per SIMD).
constexpr int COOP_ITEMS = 32;
struct CoopWave {
    int kind[COOP_ITEMS], body[COOP_ITEMS][2], g[COOP_ITEMS][4];
    double t_lo[COOP_ITEMS], t_hi[COOP_ITEMS];
    Iv R[COOP_ITEMS][2][9];       // rotation matrices, row-major
    union {
        struct {
            Iv r[COOP_ITEMS][2][3];       // interval rotation vector per body
            Iv x[COOP_ITEMS][2][2];       // its norm and the norm of r / 2
            double se[COOP_ITEMS][2][2][2];  // sinc ends [body][angle][piece]: upper at x.lo, lower at min(x.hi, mb)
            Iv gp[COOP_ITEMS][2][2];      // sin(x) / x beyond mb [body][angle]
        } s;
        Iv w[COOP_ITEMS][4][3];   // world vertices, in the order of Query::g (written after R)
        // after the t part: what each helper found for its speculative node,
        // at slot = its rank among the idle lanes (owner o reads node k at
        // o * per + k - 1); info bit 0 zero inclusion, 1 accepted if
        // reached, 2 outside the face-vertex domain, bits 3-4 split dimension
        struct {
            double tlo[64];
            unsigned int info[64];
        } n;
    } u;
};

template <int I>
RIGC_DI void rotation_row(const Iv r[3], Iv sa, Iv c2, Iv* row) {
#pragma unroll
    for (int j = 0; j < 3; j++) row[j] = rotation_entry(r, sa, c2, I, j);
}

// Value of lane `lane` (wavefront-uniform) in every lane.
RIGC_DI long long bcast_lane(long long x, int lane) {
    const int lo = __builtin_amdgcn_readlane((int)(unsigned long long)x, lane);
    const int hi = __builtin_amdgcn_readlane((int)((unsigned long long)x >> 32), lane);
    return (long long)(((unsigned long long)(unsigned int)hi << 32) | (unsigned int)lo);
}

// Orders the wavefront's LDS writes before its later LDS reads (lanes of one
// wavefront communicating through LDS; no other wavefront takes part).
RIGC_DI void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Wavefront-uniform call: lanes with `need` set get the t part of their
// (query key_q, t) in A0..DB; q is the lane's query record.
RIGC_DI void coop_t_part(const CcdArgs& a, CoopWave& L, bool need, long long key_q, const Query& q, Iv t,
                         Iv A0[3], Iv DA[3], Iv B0[3], Iv DB[3]) {
    const int lane = threadIdx.x & 63;
    unsigned long long pending = __ballot(need);
    while (pending) {
        // number up to COOP_ITEMS distinct pairs; the first lane of each
        // records its pair
        int item = -1, n = 0;
        while (pending && n < COOP_ITEMS) {
            const int leader = __ffsll((long long)pending) - 1;  // wavefront-uniform
            const long long lq = bcast_lane(key_q, leader);
            const double ll = __longlong_as_double(bcast_lane(__double_as_longlong(t.lo), leader));
            const double lh = __longlong_as_double(bcast_lane(__double_as_longlong(t.hi), leader));
            const bool same = ((pending >> lane) & 1ULL) && key_q == lq && t.lo == ll && t.hi == lh;
            if (same) item = n;
            if (lane == leader) {
                L.kind[n] = q.kind;
                L.body[n][0] = q.bA;
                L.body[n][1] = q.bB;
                for (int k = 0; k < 4; k++) L.g[n][k] = q.g[k];
                L.t_lo[n] = t.lo;
                L.t_hi[n] = t.hi;
            }
            pending &= ~__ballot(same);
            n++;
        }
        wave_lds_sync();
        // rotation vector and norms per (pair, body)
        for (int k = lane; k < 2 * n; k += 64) {
            const int u = k >> 1, side = k & 1;
            const int body = L.body[u][side];
            const Iv tt = Iv{ L.t_lo[u], L.t_hi[u] };
            Iv r[3], xa, xh;
            pose_rot_iv(a.poses_t0 + 6 * body, a.poses_t1 + 6 * body, tt, r);
            rotation_norms(r, xa, xh);
            for (int j = 0; j < 3; j++) L.u.s.r[u][side][j] = r[j];
            L.u.s.x[u][side][0] = xa;
            L.u.s.x[u][side][1] = xh;
        }
        wave_lds_sync();
        // monotone-piece ends of sinc per (pair, body, angle, piece)
        bool any_gt = false;
        for (int k = lane; k < 8 * n; k += 64) {
            const int u = k >> 3, side = (k >> 2) & 1, ang = (k >> 1) & 1, piece = k & 1;
            const Iv x = L.u.s.x[u][side][ang];
            if (x.lo <= RIGC_SINC_MB) {
                const double d = piece == 0 ? x.lo : (x.hi > RIGC_SINC_MB ? RIGC_SINC_MB : x.hi);
                L.u.s.se[u][side][ang][piece] = sinc_point_end(d, piece == 0);
            }
            any_gt = any_gt || (piece == 0 && sinc_has_gt(x));
        }
        // the piece beyond mb, where a norm reaches it
        if (__ballot(any_gt)) {
            for (int k = lane; k < 4 * n; k += 64) {
                const int u = k >> 2, side = (k >> 1) & 1, ang = k & 1;
                const Iv x = L.u.s.x[u][side][ang];
                if (sinc_has_gt(x)) {
                    const Iv arg = (x.lo <= RIGC_SINC_MB) ? Iv{ RIGC_SINC_MB, x.hi } : x;
                    L.u.s.gp[u][side][ang] = div_nonzero(sin_iv(arg), arg);
                }
            }
        }
        wave_lds_sync();
        // one rotation row per (pair, body, row)
        for (int k = lane; k < 6 * n; k += 64) {
            const int u = k / 6, side = (k / 3) & 1, i = k % 3;
            const Iv xa = L.u.s.x[u][side][0], xh = L.u.s.x[u][side][1];
            const Iv sa = sinc_combine(xa, L.u.s.se[u][side][0][0], L.u.s.se[u][side][0][1], L.u.s.gp[u][side][0]);
            const Iv sh = sinc_combine(xh, L.u.s.se[u][side][1][0], L.u.s.se[u][side][1][1], L.u.s.gp[u][side][1]);
            const Iv c2 = rotation_c2(sh);
            Iv r[3];
#pragma unroll
            for (int j = 0; j < 3; j++) r[j] = L.u.s.r[u][side][j];
            // the row as a compile-time index (the entries index r by it)
            if (i == 0) rotation_row<0>(r, sa, c2, &L.R[u][side][0]);
            else if (i == 1) rotation_row<1>(r, sa, c2, &L.R[u][side][3]);
            else rotation_row<2>(r, sa, c2, &L.R[u][side][6]);
        }
        wave_lds_sync();
        // position and one world vertex per (pair, vertex)
        for (int k = lane; k < 4 * n; k += 64) {
            const int u = k >> 2, m = k & 3;
            const int side = L.kind[u] == 0 ? (m >= 2) : (m >= 1);
            const int body = L.body[u][side];
            const Iv tt = Iv{ L.t_lo[u], L.t_hi[u] };
            const double* v = a.V + 3 * L.g[u][m];
            for (int j = 0; j < 3; j++) {
                const Iv pj = pose_pos_component(a.poses_t0 + 6 * body, a.poses_t1 + 6 * body, tt, j);
                L.u.w[u][m][j] = world_coordinate(L.R[u][side][3 * j], L.R[u][side][3 * j + 1], L.R[u][side][3 * j + 2],
                                                pj, v);
            }
        }
        wave_lds_sync();
        if (item >= 0) {
            const int u = item;
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const Iv w0 = L.u.w[u][0][j], w1 = L.u.w[u][1][j], w2 = L.u.w[u][2][j], w3 = L.u.w[u][3][j];
                if (L.kind[u] == 0) {
                    A0[j] = w0;
                    DA[j] = w1 - w0;
                    B0[j] = w2;
                    DB[j] = w3 - w2;
                } else {
                    A0[j] = w0;
                    B0[j] = w1;
                    DA[j] = w2 - w1;
                    DB[j] = w3 - w1;
                }
            }
        }
        wave_lds_sync();  // the next batch overwrites the pairs
    }
}

// zero_in(f(t, alpha, beta)) given the t part.
RIGC_DI bool contains_zero(int kind, const Iv A0[3], const Iv DA[3], const Iv B0[3], const Iv DB[3],
                           Iv u, Iv v) {
    bool z = true;
#pragma unroll
    for (int j = 0; j < 3; j++) {
        Iv d;
        if (kind == 0) d = (DA[j] * u + A0[j]) - (DB[j] * v + B0[j]);
        else d = A0[j] - ((DA[j] * u + DB[j] * v) + B0[j]);
        z = z && zero_in(d);
    }
    return z;
}


// i-th set bit (0-based) of a 64-bit mask, i < popcount(mask).
RIGC_DI int nth_set_bit(unsigned long long m, int i) {
    int pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const unsigned long long low = m & ((w == 64 ? 0ULL : (1ULL << w)) - 1ULL);
        const int c = __popcll(low);
        if (i >= c) {
            i -= c;
            m >>= w;
            pos += w;
        } else {
            m = low;
        }
    }
    return pos;
}

// Lane src's query record and alpha / beta tolerances, in every lane (a
// receiver of a donated subtree or a helper takes them from the lane that
// holds the same query instead of loading them again).
RIGC_DI void shfl_query(const Query& q, const double tol[3], int src, Query& out, double& tol1, double& tol2) {
    out.kind = __shfl(q.kind, src);
    out.bA = __shfl(q.bA, src);
    out.bB = __shfl(q.bB, src);
    out.sc = __shfl(q.sc, src);
#pragma unroll
    for (int k = 0; k < 4; k++) out.g[k] = __shfl(q.g[k], src);
    tol1 = __shfl(tol[1], src);
    tol2 = __shfl(tol[2], src);
}

RIGC_DI void load_query(const CcdArgs& a, long long qi, Query& q, double tol[3]) {
    if (qi < a.n_ee) {
        const int ea = a.ee[2 * qi], eb = a.ee[2 * qi + 1];
        q.kind = 0;
        q.bA = a.e_body[ea];
        q.bB = a.e_body[eb];
        q.g[0] = a.E[2 * ea];
        q.g[1] = a.E[2 * ea + 1];
        q.g[2] = a.E[2 * eb];
        q.g[3] = a.E[2 * eb + 1];
        tol[1] = 1e-4 / a.edge_len[ea];
        tol[2] = 1e-4 / a.edge_len[eb];
    } else {
        const long long k = qi - a.n_ee;
        const int f = a.fv[2 * k], v = a.fv[2 * k + 1];
        q.kind = 1;
        q.bA = a.v_body[v];
        q.bB = a.f_body[f];
        q.g[0] = v;
        q.g[1] = a.F[3 * f];
        q.g[2] = a.F[3 * f + 1];
        q.g[3] = a.F[3 * f + 2];
        tol[1] = 1e-4 / a.edge_len[a.FE[3 * f]];
        tol[2] = 1e-4 / a.edge_len[a.FE[3 * f + 1]];
    }
    q.sc = a.body_scene[q.bA];
}

// Widths of a dyadic box at levels lev (rounded up, as the root finder's
// width()) and the bisection rule of interval_root_finder: split the
// dimension of largest width / tolerance among those still wider than their
// tolerance, or among all dimensions once every width is within tolerance.
RIGC_DI void box_widths(const int lev[3], double w[3]) {
#pragma unroll
    for (int i = 0; i < 3; i++) w[i] = next_up(ldexp(1.0, -lev[i]));
}
RIGC_DI int split_dim(const double w[3], const double tol[3], bool all_sat) {
    int sp = -1;
    double ws = 0.0, ts = 1.0;  // w[sp], tol[sp]
#pragma unroll
    for (int i = 0; i < 3; i++)
        if ((all_sat || w[i] > tol[i]) && (sp == -1 || w[i] * ts > ws * tol[i])) {
            sp = i;
            ws = w[i];
            ts = tol[i];
        }
    return sp;
}
RIGC_DI int split_dim_at(const int lev[3], const double tol[3]) {
    double w[3];
    box_widths(lev, w);
    return split_dim(w, tol, w[0] <= tol[0] && w[1] <= tol[1] && w[2] <= tol[2]);
}

// Per-query state, and the counters of the first pass (work counter,
// evaluation total, export count), so that pass needs no clears of its own.
__global__ void ccd_init_kernel(CcdArgs a) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        *a.counter = 0;
        *a.total_evals = 0;
        *a.n_tasks_out = 0;
    }
    if (i < a.n_ee + a.n_fv) {
        a.q_best[i] = 0x7ff0000000000000ULL;
        a.q_visits[i] = 0;
        a.q_flag[i] = 0;
    }
}

__global__ void ccd_finish_kernel(CcdArgs a) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < a.n_ee + a.n_fv) {
        const int f = a.q_flag[i];
        const unsigned long long b = a.q_best[i];
        if (a.toi_out) a.toi_out[i] = __longlong_as_double((long long)b);
        if (a.hit_out) a.hit_out[i] = f == 1 ? -1 : (f == 2 ? -2 : (b < 0x7ff0000000000000ULL ? 1 : 0));
    }
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void ccd_kernel(CcdArgs a) {
    const long long nq = a.n_ee + a.n_fv;
    const int lane = threadIdx.x & 63;
    const unsigned long long lt_mask = (1ULL << lane) - 1ULL;
    __shared__ CoopWave coop[4];  // one per wavefront of the 256-thread block

    long long qi = -1;  // query of the subtree this lane walks, -1 = idle
    int phase = 0;      // 1 = start-box check, 2 = bisection
    Query q;
    double tol[3];
    long long evals = 0;
    int ev_sc = -1;              // scene of the evaluations counted in ev_n
    unsigned long long ev_n = 0;
    int walk = 0;       // evaluations since this lane took its subtree
    double thr = 0.0;   // cached pruning bound of query thr_q
    long long thr_q = -1;
    int thr_age = 0;
    const long long npool = a.tasks_in ? a.n_tasks_in : nq;

    int depth = 0;
    Bits192 s, mt, ma;
    unsigned long long idx[3];
    int lev[3];

    bool cache_ok = false;
    unsigned long long c_idx = 0;
    int c_lev = -1;
    Iv A0[3], DA[3], B0[3], DB[3];

    // Queries are spread over all wavefronts of the grid: a wavefront takes
    // at most `quota` new queries at a time, and its remaining idle lanes
    // help with the subtrees of the queries it holds.
    const long long nwaves = (long long)gridDim.x * (blockDim.x / 64);
    const int quota = (int)max(1LL, min(64LL, (npool + nwaves - 1) / nwaves));
    bool pool_open = true;
    long long spec_q = -1;  // query whose data this idle lane holds from helping (-1: none)
    for (;;) {
        // A. idle lanes take new queries while the pool lasts
        if (pool_open) {
            unsigned long long im = __ballot(qi < 0);
            if (__popcll(im) > quota) {
                // keep the `quota` lowest idle lanes
                const int cut = nth_set_bit(im, quota);
                im &= (1ULL << cut) - 1ULL;
            }
            if (im) {
                const int leader = __ffsll((long long)im) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(a.counter, (unsigned long long)__popcll(im));
                base = __shfl(base, leader);
                if (base + __popcll(im) >= (unsigned long long)npool) pool_open = false;
                if ((im >> lane) & 1ULL) {
                    const long long mine = (long long)(base + __popcll(im & lt_mask));
                    if (mine < npool) {
                        walk = 0;
                        cache_ok = false;
                        spec_q = -1;
                        if (a.tasks_in) {
                            const CcdTask& tk = a.tasks_in[mine];
                            qi = tk.qi;
                            load_query(a, qi, q, tol);
                            tol[0] = tk.tol0;
                            idx[0] = tk.idx[0]; idx[1] = tk.idx[1]; idx[2] = tk.idx[2];
                            lev[0] = tk.lev & 255; lev[1] = (tk.lev >> 8) & 255; lev[2] = (tk.lev >> 16) & 255;
                            depth = 0;
                            clear_bits(s);
                            mt = s;
                            ma = s;
                            phase = 2;
                        } else {
                            qi = mine;
                            load_query(a, qi, q, tol);
                            tol[0] = a.toi_tol;
                            phase = 1;
                            if (a.scene_skip && a.scene_skip[q.sc]) qi = -1;
                        }
                    }
                }
            }
        }
        // B. idle lanes receive pending subtrees of busy lanes
        {
            const unsigned long long im = __ballot(qi < 0);
            if (im == ~0ULL && !pool_open) break;  // no lane of this wavefront has work left
            if (im) {
                const int first_open = (qi >= 0 && phase == 2) ? [&] {
#pragma unroll
                    for (int wi = 0; wi < 3; wi++) {
                        const unsigned long long m = ~word(s, wi) & range_mask(wi, 0, depth);
                        if (m) return 64 * wi + __ffsll((long long)m) - 1;
                    }
                    return -1;
                }()
                                                                : -1;
                const unsigned long long dm = __ballot(first_open >= 0);
                const int k = min(__popcll(im), __popcll(dm));
                // donor side: the second half of the shallowest open level
                long long d_q = -1;
                double d_tol0 = 0;
                unsigned long long d_idx[3] = { 0, 0, 0 };
                int d_lev = 0;
                const int drank = __popcll(dm & lt_mask);
                if (first_open >= 0 && drank < k) {
                    const int L = first_open;
                    const int ct = popcount_range(mt, L + 1, depth), ca = popcount_range(ma, L + 1, depth);
                    const int cb = depth - 1 - L - ct - ca;
                    d_idx[0] = idx[0] >> ct;
                    d_idx[1] = idx[1] >> ca;
                    d_idx[2] = idx[2] >> cb;
                    const int l0 = lev[0] - ct, l1 = lev[1] - ca, l2 = lev[2] - cb;
                    const int d = bget(mt, L) ? 0 : (bget(ma, L) ? 1 : 2);
                    add_at(d_idx, d, 1ULL);
                    bset(s, L, true);
                    d_q = qi;
                    d_tol0 = tol[0];
                    d_lev = l0 | (l1 << 8) | (l2 << 16);
                }
                // receiver side
                const int irank = __popcll(im & lt_mask);
                const int src = (qi < 0 && irank < k) ? nth_set_bit(dm, irank) : lane;
                const long long r_q = __shfl(d_q, src);
                const double r_tol0 = __shfl(d_tol0, src);
                const unsigned long long r0 = __shfl(d_idx[0], src);
                const unsigned long long r1 = __shfl(d_idx[1], src);
                const unsigned long long r2 = __shfl(d_idx[2], src);
                const int r_lev = __shfl(d_lev, src);
                // the donor holds the query's record
                Query r_rec;
                double r_tol1 = 0, r_tol2 = 0;
                if (k > 0) shfl_query(q, tol, src, r_rec, r_tol1, r_tol2);
                if (qi < 0 && irank < k) {
                    qi = r_q;
                    spec_q = -1;
                    q = r_rec;
                    tol[1] = r_tol1;
                    tol[2] = r_tol2;
                    tol[0] = r_tol0;
                    idx[0] = r0; idx[1] = r1; idx[2] = r2;
                    lev[0] = r_lev & 255; lev[1] = (r_lev >> 8) & 255; lev[2] = (r_lev >> 16) & 255;
                    depth = 0;
                    clear_bits(s);
                    mt = s;
                    ma = s;
                    phase = 2;
                    cache_ok = false;
                    walk = 0;
                }
            }
        }
        // C. one box per lane of its own query. Lanes left without work
        // evaluate, for lanes evaluating a bisection box, the descendants of
        // that box down to spec_levels levels (each box of the subtree is a
        // fixed function of its root: the split dimension depends only on
        // the box's widths); the owner then walks those levels from the
        // recorded results without waiting for further evaluations. Which
        // boxes are reached and accepted is unchanged, so is the minimum.
        bool done = false;
        bool own_eval = false;
        Iv t, u, v;
        // Pops to the next pending second half; done when none is left.
        auto pop = [&]() {
            const int L = last_zero_below(s, depth);
            if (L < 0) {
                done = true;
                return;
            }
            const int ct = popcount_range(mt, L + 1, depth), ca = popcount_range(ma, L + 1, depth);
            const int cb = depth - 1 - L - ct - ca;
            idx[0] >>= ct; lev[0] -= ct;
            idx[1] >>= ca; lev[1] -= ca;
            idx[2] >>= cb; lev[2] -= cb;
            const int d = bget(mt, L) ? 0 : (bget(ma, L) ? 1 : 2);
            add_at(idx, d, 1ULL);
            bset(s, L, true);
            depth = L + 1;
        };
        // Skips boxes pruned by the running minimum or outside the FV domain.
        auto skip_pruned = [&]() {
            while (!done) {
                if (a.eval_budget > 0) {
                    const unsigned int nv = atomicAdd(a.q_visits + qi, 1u) + 1u;
                    if ((long long)nv > a.eval_budget) {
                        atomicMax(a.q_flag + qi, 1);
                        done = true;
                        break;
                    }
                }
                bool skip = dyad(idx[0], lev[0]) >= thr;
                if (!skip && q.kind == 1) {
                    const Iv uu = Iv{ dyad(idx[1], lev[1]), dyad(idx[1] + 1, lev[1]) };
                    const Iv vv = Iv{ dyad(idx[2], lev[2]), dyad(idx[2] + 1, lev[2]) };
                    skip = !overlap(uu + vv, Iv{ 0.0, 1.0 });
                }
                if (!skip) break;
                pop();
            }
        };
        // An accepted box with t.lower tlo: the query's minimum (and the
        // scene's running minimum) and this lane's pruning bound.
        auto accept = [&](double tlo) {
            const unsigned long long bits = (unsigned long long)__double_as_longlong(tlo);
            atomicMin(a.q_best + qi, bits);
            thr = fmin(thr, tlo);
            if (a.use_global_best) atomicMin(a.global_best + q.sc, bits);
        };
        // Consumes the zero-inclusion result z of the current box.
        auto process = [&](bool z) {
            bool leaf = !z;
            if (z) {
                double w[3];
                box_widths(lev, w);
                const bool all_sat = w[0] <= tol[0] && w[1] <= tol[1] && w[2] <= tol[2];
                const bool all_zero = w[0] <= 1e-10 && w[1] <= 1e-10 && w[2] <= 1e-10;
                if ((idx[0] > 0 || all_zero) && all_sat) {
                    accept(dyad(idx[0], lev[0]));
                    leaf = true;
                } else {
                    const int sp = split_dim(w, tol, all_sat);
                    if (depth >= MAX_DEPTH) {
                        atomicMax(a.q_flag + qi, 2);
                        done = true;
                    } else {
                        bset(s, depth, false);
                        bset(mt, depth, sp == 0);
                        bset(ma, depth, sp == 1);
                        descend_at(idx, sp, 0ULL);
                        add_at(lev, sp, 1);
                        depth++;
                    }
                }
            }
            if (leaf && !done) pop();
        };

        if (qi >= 0) {
            if (phase == 1) {
                t = Iv{ 0.0, tol[0] };
                u = Iv{ 0.0, tol[1] };
                v = Iv{ 0.0, tol[2] };
                own_eval = true;
            } else {
                // The pruning bound only decreases; a stale (larger) bound
                // prunes less but never changes the minimum, so it is
                // refreshed from global memory every THR_REFRESH boxes (or
                // on a new subtree).
                if (thr_age <= 0 || thr_q != qi) {
                    thr = __longlong_as_double((long long)__hip_atomic_load(a.q_best + qi, __ATOMIC_RELAXED,
                                                                            __HIP_MEMORY_SCOPE_AGENT));
                    if (a.use_global_best) {
                        const unsigned long long gb =
                            __hip_atomic_load(a.global_best + q.sc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        thr = fmin(thr, __longlong_as_double((long long)gb));
                    }
                    thr_q = qi;
                    thr_age = THR_REFRESH;
                }
                thr_age--;
                if (a.eval_budget > 0 && a.q_flag[qi] != 0) done = true;
                skip_pruned();
                if (!done) {
                    t = Iv{ dyad(idx[0], lev[0]), dyad(idx[0] + 1, lev[0]) };
                    u = Iv{ dyad(idx[1], lev[1]), dyad(idx[1] + 1, lev[1]) };
                    v = Iv{ dyad(idx[2], lev[2]), dyad(idx[2] + 1, lev[2]) };
                    own_eval = true;
                }
            }
        }

        // speculation assignment: idle lane of rank h serves owner h / per,
        // subtree node h % per + 1 (heap order: root 0, children of node j
        // at 2j + 1 and 2j + 2)
        const unsigned long long owners = __ballot(own_eval && phase == 2 && a.eval_budget == 0);
        const unsigned long long idle = __ballot(qi < 0);
        const int n_own = __popcll(owners), n_idle = __popcll(idle);
        const int g = n_own ? n_idle / n_own : 0;
        int levels = 0;
        const int max_levels = min(a.spec_levels, SPEC_LEVELS);
        while (levels < max_levels && (2 << (levels + 1)) - 2 <= g) levels++;
        const int per = levels ? (2 << levels) - 2 : 0;
        int src = lane, node = 0, slot = 0;
        if (per && qi < 0) {
            const int h = __popcll(idle & lt_mask);
            const int o = h / per;
            if (o < n_own) {
                src = nth_set_bit(owners, o);
                node = h % per + 1;
                slot = h;
            }
        }
        const long long s_q = __shfl(qi, src);
        const double s_tol0 = __shfl(tol[0], src);
        const unsigned long long s_i0 = __shfl(idx[0], src), s_i1 = __shfl(idx[1], src), s_i2 = __shfl(idx[2], src);
        const int s_lev = __shfl(lev[0] | (lev[1] << 8) | (lev[2] << 16), src);
        bool spec = false;
        unsigned int spec_info = 0;  // the node's record without its zero-inclusion bit
        unsigned long long e_idx0 = idx[0];
        int e_lev0 = lev[0];
        // a helper that changes query takes the record from its owner
        const bool reload = node > 0 && s_q != spec_q;
        Query h_rec;
        double h_tol1 = 0, h_tol2 = 0;
        if (__ballot(reload)) shfl_query(q, tol, src, h_rec, h_tol1, h_tol2);
        if (node > 0) {
            // this lane is idle: its query registers are free until it takes work
            if (reload) {
                q = h_rec;
                tol[1] = h_tol1;
                tol[2] = h_tol2;
                spec_q = s_q;
            }
            tol[0] = s_tol0;
            unsigned long long bi[3] = { s_i0, s_i1, s_i2 };
            int bl[3] = { s_lev & 255, (s_lev >> 8) & 255, (s_lev >> 16) & 255 };
            const unsigned m = (unsigned)node + 1u;
            for (int j = 30 - __clz(m); j >= 0; j--) {
                const int sp = split_dim_at(bl, tol);
                descend_at(bi, sp, (unsigned long long)((m >> j) & 1u));
                add_at(bl, sp, 1);
            }
            t = Iv{ dyad(bi[0], bl[0]), dyad(bi[0] + 1, bl[0]) };
            u = Iv{ dyad(bi[1], bl[1]), dyad(bi[1] + 1, bl[1]) };
            v = Iv{ dyad(bi[2], bl[2]), dyad(bi[2] + 1, bl[2]) };
            e_idx0 = bi[0];
            e_lev0 = bl[0];
            spec = true;
            cache_ok = false;
            // what process / skip_pruned would decide at this box, for the owner
            double w[3];
            box_widths(bl, w);
            const bool all_sat = w[0] <= tol[0] && w[1] <= tol[1] && w[2] <= tol[2];
            const bool all_zero = w[0] <= 1e-10 && w[1] <= 1e-10 && w[2] <= 1e-10;
            const bool acc = (bi[0] > 0 || all_zero) && all_sat;
            const bool outside = q.kind == 1 && !overlap(u + v, Iv{ 0.0, 1.0 });
            spec_info = (acc ? 2u : 0u) | (outside ? 4u : 0u) | ((unsigned int)split_dim(w, tol, all_sat) << 3);
        }

        const bool evaluating = own_eval || spec;
        const bool need_t = evaluating && (spec || phase == 1 || !cache_ok || c_idx != e_idx0 || c_lev != e_lev0);
        coop_t_part(a, coop[threadIdx.x >> 6], need_t, spec ? s_q : qi, q, t, A0, DA, B0, DB);
        if (need_t) {
            cache_ok = !spec && phase == 2;
            c_idx = e_idx0;
            c_lev = e_lev0;
        }
        bool z = false;
        if (evaluating) {
            z = contains_zero(q.kind, A0, DA, B0, DB, u, v);
            evals++;
            if (a.scene_evals) {
                // per-scene counts: accumulated in the lane, added to the
                // scene's counter when the lane moves to another scene
                if (q.sc != ev_sc) {
                    if (ev_n) atomicAdd(a.scene_evals + ev_sc, ev_n);
                    ev_sc = q.sc;
                    ev_n = 0;
                }
                ev_n++;
            }
        }
        CoopWave& cw = coop[threadIdx.x >> 6];
        if (spec) {
            cache_ok = false;
            cw.u.n.tlo[slot] = t.lo;
            cw.u.n.info[slot] = spec_info | (z ? 1u : 0u);
        }
        wave_lds_sync();

        if (own_eval) {
            walk++;
            if (phase == 1) {
                if (z) tol[0] /= 1e2;
                phase = 2;
                depth = 0;
                clear_bits(s);
                mt = s;
                ma = s;
                idx[0] = idx[1] = idx[2] = 0;
                lev[0] = lev[1] = lev[2] = 0;
                cache_ok = false;
                continue;
            }
            const int k_levels = (per && ((owners >> lane) & 1ULL)) ? levels : 0;
            if (k_levels == 0) {
                process(z);
            } else {
                // The owner walks the known levels in heap order (root 0,
                // children of node j at 2j + 1 and 2j + 2) from its helpers'
                // records, deciding as process / pop / skip_pruned do on the
                // path masks, and writes the path masks once at the end. The
                // boxes reached and accepted are those of the box-by-box walk.
                const int base = __popcll(owners & lt_mask) * per - 1;  // record of node k: base + k
                const int root_d = depth;
                double w[3];
                box_widths(lev, w);
                const bool all_sat = w[0] <= tol[0] && w[1] <= tol[1] && w[2] <= tol[2];
                const bool all_zero = w[0] <= 1e-10 && w[1] <= 1e-10 && w[2] <= 1e-10;
                const int sp0 = split_dim(w, tol, all_sat);
                bool zc = z, acc = (idx[0] > 0 || all_zero) && all_sat;
                double tl = dyad(idx[0], lev[0]);
                int kc = 0, dc = 0;
                int ex = 0;  // 0: descended below the known levels, 1: left the subtree, 2: too deep
                for (;;) {
                    if (zc && !acc) {
                        if (root_d + dc >= MAX_DEPTH) {
                            atomicMax(a.q_flag + qi, 2);
                            done = true;
                            ex = 2;
                            break;
                        }
                        kc = 2 * kc + 1;
                        dc++;
                        if (dc > k_levels) break;
                    } else {
                        if (zc) accept(tl);
                        // next pending second half: up over the levels whose
                        // second half is taken, then across
                        const unsigned int m = (unsigned int)kc + 1u;
                        const int up = __builtin_ctz(~m);
                        if (up >= dc) {
                            ex = 1;
                            break;
                        }
                        kc = (int)(((m >> up) | 1u) - 1u);
                        dc -= up;
                    }
                    const unsigned int inf = cw.u.n.info[base + kc];
                    const double tk = cw.u.n.tlo[base + kc];
                    if (tk >= thr || (inf & 4u)) {  // pruned or outside the domain: a leaf
                        zc = false;
                        continue;
                    }
                    walk++;
                    zc = (inf & 1u) != 0u;
                    acc = (inf & 2u) != 0u;
                    tl = tk;
                }
                if (ex == 1) {
                    pop();
                } else if (ex == 0) {
                    // the path from the root to node kc (depth dc): split
                    // dimensions from the records; a pending second half
                    // whose box is known to be a leaf (no zero, pruned or
                    // outside the domain) is marked visited, which is what
                    // visiting it would conclude
                    const unsigned int path = (unsigned int)kc + 1u;  // a leading 1, then dc child bits
                    for (int i = 1; i <= dc; i++) {
                        const unsigned int above = path >> (dc - i + 1);  // node at depth i - 1, plus 1
                        const int sp = i == 1 ? sp0 : (int)((cw.u.n.info[base + (int)above - 1] >> 3) & 3u);
                        const unsigned int here = path >> (dc - i);
                        const bool second = (here & 1u) != 0u;
                        bool closed = second;
                        if (!second && i <= k_levels) {
                            const int sib = (int)here;  // (here | 1) - 1
                            const unsigned int si = cw.u.n.info[base + sib];
                            closed = !(si & 1u) || (si & 4u) || cw.u.n.tlo[base + sib] >= thr;
                        }
                        const int P = root_d + i - 1;
                        bset(s, P, closed);
                        bset(mt, P, sp == 0);
                        bset(ma, P, sp == 1);
                        descend_at(idx, sp, second ? 1ULL : 0ULL);
                        add_at(lev, sp, 1);
                    }
                    depth = root_d + dc;
                }
            }
        }
        // D. export the pending subtrees of a long walk to the next pass
        if (qi >= 0 && !done && phase == 2 && !pool_open && a.tasks_out && walk >= a.walk_budget) {
            int n_open = 0;
#pragma unroll
            for (int wi = 0; wi < 3; wi++) n_open += __popcll(~word(s, wi) & range_mask(wi, 0, depth));
            const unsigned long long base = atomicAdd(a.n_tasks_out, (unsigned long long)(n_open + 1));
            const unsigned long long cap = (unsigned long long)a.tasks_cap;
            int k = 0;
            for (int L = 0; L < depth; L++) {
                if (bget(s, L)) continue;
                const unsigned long long slot = base + (unsigned long long)(k++);
                if (slot >= cap) continue;
                const int ct = popcount_range(mt, L + 1, depth), ca = popcount_range(ma, L + 1, depth);
                const int cb = depth - 1 - L - ct - ca;
                CcdTask tk;
                tk.qi = qi;
                tk.tol0 = tol[0];
                tk.idx[0] = idx[0] >> ct;
                tk.idx[1] = idx[1] >> ca;
                tk.idx[2] = idx[2] >> cb;
                const int d = bget(mt, L) ? 0 : (bget(ma, L) ? 1 : 2);
                add_at(tk.idx, d, 1ULL);
                tk.lev = (lev[0] - ct) | ((lev[1] - ca) << 8) | ((lev[2] - cb) << 16);
                tk.pad = 0;
                a.tasks_out[slot] = tk;
                bset(s, L, true);
            }
            const unsigned long long slot = base + (unsigned long long)n_open;
            if (slot < cap) {
                CcdTask tk;
                tk.qi = qi;
                tk.tol0 = tol[0];
                tk.idx[0] = idx[0]; tk.idx[1] = idx[1]; tk.idx[2] = idx[2];
                tk.lev = lev[0] | (lev[1] << 8) | (lev[2] << 16);
                tk.pad = 0;
                a.tasks_out[slot] = tk;
                done = true;
            } else {
                walk = -(1 << 30);  // the buffer is full: finish this subtree here
            }
        }
        if (done) qi = -1;
    }
    if (evals) atomicAdd(a.total_evals, (unsigned long long)evals);
    if (ev_n) atomicAdd(a.scene_evals + ev_sc, ev_n);
}

}  // namespace

hipError_t launch_ccd_single(CcdArgs a, int grid, int block, hipStream_t stream) {
    const long long nq = a.n_ee + a.n_fv;
    if (nq <= 0) return hipSuccess;
    const unsigned eb = (unsigned)((nq + 255) / 256);
    a.tasks_in = nullptr;
    a.n_tasks_in = 0;
    a.tasks_out = nullptr;
    a.tasks_cap = 0;
    a.walk_budget = 0;
    hipLaunchKernelGGL(ccd_init_kernel, dim3(eb), dim3(256), 0, stream, a);
    hipLaunchKernelGGL(ccd_kernel, dim3(grid), dim3(block), 0, stream, a);
    if (a.toi_out || a.hit_out) hipLaunchKernelGGL(ccd_finish_kernel, dim3(eb), dim3(256), 0, stream, a);
    return hipGetLastError();
}

hipError_t launch_ccd(CcdArgs a, int grid, int block, hipStream_t stream, CcdWork& w, const CcdReadback* rb,
                      int n_rb) {
    const long long nq = a.n_ee + a.n_fv;
    w.passes = 0;
    w.kernel_ms = 0.0f;
    if (nq <= 0) return hipSuccess;
    const long long cap = 1LL << 18;
    hipError_t e;
    if (!w.k0 && (e = hipEventCreate(&w.k0)) != hipSuccess) return e;
    if (!w.k1 && (e = hipEventCreate(&w.k1)) != hipSuccess) return e;
    if ((e = w.tasks[0].ensure(sizeof(CcdTask) * cap)) != hipSuccess) return e;
    if ((e = w.tasks[1].ensure(sizeof(CcdTask) * cap)) != hipSuccess) return e;
    if ((e = w.count.ensure(sizeof(unsigned long long))) != hipSuccess) return e;
    if ((e = w.h_count.ensure(sizeof(unsigned long long))) != hipSuccess) return e;
    unsigned long long* h_n = w.h_count.as<unsigned long long>();
    const unsigned eb = (unsigned)((nq + 255) / 256);
    a.tasks_in = nullptr;
    a.n_tasks_in = 0;
    a.tasks_cap = cap;
    a.n_tasks_out = w.count.as<unsigned long long>();
    hipLaunchKernelGGL(ccd_init_kernel, dim3(eb), dim3(256), 0, stream, a);
    const int budget0 = a.walk_budget;
    for (int pass = 0;; pass++) {
        // later passes export only much longer walks, and from the fourth
        // pass on not at all, so the number of passes stays small
        a.walk_budget = budget0 << (2 * std::min(pass, 8));
        a.tasks_out = (budget0 > 0 && pass < 3) ? w.tasks[pass & 1].as<CcdTask>() : nullptr;
        if (pass > 0) {  // the init kernel cleared them for the first pass
            if ((e = hipMemsetAsync(a.counter, 0, sizeof(unsigned long long), stream)) != hipSuccess) return e;
            if ((e = hipMemsetAsync(a.n_tasks_out, 0, sizeof(unsigned long long), stream)) != hipSuccess) return e;
        }
        if ((e = hipEventRecord(w.k0, stream)) != hipSuccess) return e;
        hipLaunchKernelGGL(ccd_kernel, dim3(grid), dim3(block), 0, stream, a);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = hipEventRecord(w.k1, stream)) != hipSuccess) return e;
        w.passes = pass + 1;
        if ((e = hipMemcpyAsync(h_n, a.n_tasks_out, sizeof(*h_n), hipMemcpyDeviceToHost, stream)) != hipSuccess)
            return e;
        // the caller's results ride on the same synchronisation; a further
        // pass overwrites them
        for (int i = 0; i < n_rb; i++)
            if ((e = hipMemcpyAsync(rb[i].dst, rb[i].src, rb[i].bytes, hipMemcpyDeviceToHost, stream)) != hipSuccess)
                return e;
        if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
        float ms = 0.0f;
        if ((e = hipEventElapsedTime(&ms, w.k0, w.k1)) != hipSuccess) return e;
        w.kernel_ms += ms;
        const unsigned long long n = *h_n;
        if (n == 0) break;
        a.tasks_in = w.tasks[pass & 1].as<CcdTask>();
        a.n_tasks_in = (long long)std::min<unsigned long long>(n, (unsigned long long)cap);
    }
    if (a.toi_out || a.hit_out) hipLaunchKernelGGL(ccd_finish_kernel, dim3(eb), dim3(256), 0, stream, a);
    return hipGetLastError();
}

}  // namespace rigc
