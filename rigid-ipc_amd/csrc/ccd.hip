// Rigid-trajectory interval CCD narrow phase for gfx950.
//
// Restates, per candidate, compute_edge_edge_time_of_impact /
// compute_face_vertex_time_of_impact (src/ccd/rigid/time_of_impact.cpp:158-310)
// driving interval_root_finder (src/interval/interval_root_finder.cpp:97-203)
// over the inclusion functions of src/ccd/rigid/rigid_trajectory_aabb.cpp:70-102.
//
// Execution model: a persistent grid; every lane owns one query at a time
// and pulls the next query from a global counter when it finishes, so a
// wavefront keeps all 64 lanes busy while any queries remain. The
// bisection tree of a query is walked depth-first in the reference's order
// (first half before second half) without a box stack: every box is a
// dyadic sub-box of [0,1]^3 held as an integer (index, level) per
// dimension, and the path from the root is held as three 192-bit masks
// (split dimension t / alpha per level, and whether the second half of
// that level is already taken). Popping the next box is a find-last-zero
// plus popcounts. The t-only part of the inclusion function (interval
// poses, interval Rodrigues matrices, world vertices) is cached per lane
// and reused while only alpha/beta are being split.
//
// The root finder's answer is the least t.lower over all accepted boxes of
// the bisection tree, which is independent of traversal order, so pruning
// with the running minimum over all queries (earliest-TOI mode) returns
// exactly the minimum of the per-query answers.
#include "rigc_interval.hpp"
#include "rigc_internal.hpp"
#include "rigc_rigid_iv.hpp"

namespace rigc {
namespace {

constexpr int MAX_DEPTH = 192;

struct Bits192 {
    unsigned long long w[3];
};
RIGC_DI void bset(Bits192& b, int k, bool v) {
    const unsigned long long m = 1ULL << (k & 63);
    if (k < 64) b.w[0] = v ? (b.w[0] | m) : (b.w[0] & ~m);
    else if (k < 128) b.w[1] = v ? (b.w[1] | m) : (b.w[1] & ~m);
    else b.w[2] = v ? (b.w[2] | m) : (b.w[2] & ~m);
}
RIGC_DI bool bget(const Bits192& b, int k) {
    const unsigned long long w = k < 64 ? b.w[0] : (k < 128 ? b.w[1] : b.w[2]);
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
    return __popcll(b.w[0] & range_mask(0, lo, hi)) + __popcll(b.w[1] & range_mask(1, lo, hi))
        + __popcll(b.w[2] & range_mask(2, lo, hi));
}
// Largest level L < depth whose bit is clear, or -1.
RIGC_DI int last_zero_below(const Bits192& b, int depth) {
    for (int wi = 2; wi >= 0; wi--) {
        const unsigned long long m = ~b.w[wi] & range_mask(wi, 0, depth);
        if (m) return 64 * wi + 63 - __clzll(m);
    }
    return -1;
}

RIGC_DI double dyad(unsigned long long i, int l) { return ldexp((double)i, -l); }

struct Query {
    int kind, bA, bB;
    int g[4];  // EE: ea0 ea1 eb0 eb1; FV: v f0 f1 f2 (global vertex ids)
};

// t-dependent part of the inclusion function.
// EE: A0 = ea0(t), DA = ea1(t) - ea0(t), B0 = eb0(t), DB = eb1(t) - eb0(t)
// FV: A0 = v(t),   DA = f1(t) - f0(t),   B0 = f0(t),  DB = f2(t) - f0(t)
RIGC_DI void t_part(const CcdArgs& a, const Query& q, Iv t, Iv A0[3], Iv DA[3], Iv B0[3], Iv DB[3]) {
    Iv p[3], R[3][3];
    pose_iv(a.poses_t0 + 6 * q.bA, a.poses_t1 + 6 * q.bA, t, p, R);
    world_vertex_iv(R, p, a.V + 3 * q.g[0], A0);
    if (q.kind == 0) {
        Iv w1[3];
        world_vertex_iv(R, p, a.V + 3 * q.g[1], w1);
#pragma unroll
        for (int j = 0; j < 3; j++) DA[j] = w1[j] - A0[j];
        pose_iv(a.poses_t0 + 6 * q.bB, a.poses_t1 + 6 * q.bB, t, p, R);
        world_vertex_iv(R, p, a.V + 3 * q.g[2], B0);
        world_vertex_iv(R, p, a.V + 3 * q.g[3], w1);
#pragma unroll
        for (int j = 0; j < 3; j++) DB[j] = w1[j] - B0[j];
    } else {
        Iv f1[3], f2[3];
        pose_iv(a.poses_t0 + 6 * q.bB, a.poses_t1 + 6 * q.bB, t, p, R);
        world_vertex_iv(R, p, a.V + 3 * q.g[1], B0);
        world_vertex_iv(R, p, a.V + 3 * q.g[2], f1);
        world_vertex_iv(R, p, a.V + 3 * q.g[3], f2);
#pragma unroll
        for (int j = 0; j < 3; j++) {
            DA[j] = f1[j] - B0[j];
            DB[j] = f2[j] - B0[j];
        }
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

__global__ __launch_bounds__(256) void ccd_kernel(CcdArgs a) {
    const unsigned long long nq = (unsigned long long)(a.n_ee + a.n_fv);

    Query q;
    unsigned long long qi = 0;
    int phase = 0;  // 0 = fetch, 1 = start-box check, 2 = bisection
    double tol[3];
    double best = 0;
    int hit = 0;
    long long visits = 0, evals = 0;

    int depth = 0;
    Bits192 s, mt, ma;
    unsigned long long idx[3];
    int lev[3];

    bool cache_ok = false;
    unsigned long long c_idx = 0;
    int c_lev = -1;
    Iv A0[3], DA[3], B0[3], DB[3];

    for (;;) {
        if (phase == 0) {
            qi = atomicAdd(a.counter, 1ULL);
            if (qi >= nq) break;
            if (qi < (unsigned long long)a.n_ee) {
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
                const unsigned long long k = qi - a.n_ee;
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
            tol[0] = a.toi_tol;
            best = __longlong_as_double(0x7ff0000000000000LL);
            hit = 0;
            visits = 0;
            evals = 0;
            cache_ok = false;
            phase = 1;
        }

        bool finished = false;
        Iv t, u, v;
        if (phase == 1) {
            t = Iv{ 0.0, tol[0] };
            u = Iv{ 0.0, tol[1] };
            v = Iv{ 0.0, tol[2] };
            cache_ok = false;
        } else {
            double thr = best;
            if (a.use_global_best) {
                const unsigned long long gb =
                    __hip_atomic_load(a.global_best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                thr = fmin(thr, __longlong_as_double((long long)gb));
            }
            // Pop boxes until one needs an evaluation.
            for (;;) {
                if (a.eval_budget > 0 && ++visits > a.eval_budget) {
                    hit = -1;
                    finished = true;
                    break;
                }
                bool skip = dyad(idx[0], lev[0]) >= thr;
                if (!skip && q.kind == 1) {
                    const Iv uu = Iv{ dyad(idx[1], lev[1]), dyad(idx[1] + 1, lev[1]) };
                    const Iv vv = Iv{ dyad(idx[2], lev[2]), dyad(idx[2] + 1, lev[2]) };
                    skip = !overlap(uu + vv, Iv{ 0.0, 1.0 });
                }
                if (!skip) break;
                // advance to the next box in depth-first order
                const int L = last_zero_below(s, depth);
                if (L < 0) {
                    finished = true;
                    break;
                }
                const int ct = popcount_range(mt, L + 1, depth), ca = popcount_range(ma, L + 1, depth);
                const int cb = depth - 1 - L - ct - ca;
                idx[0] >>= ct; lev[0] -= ct;
                idx[1] >>= ca; lev[1] -= ca;
                idx[2] >>= cb; lev[2] -= cb;
                const int d = bget(mt, L) ? 0 : (bget(ma, L) ? 1 : 2);
                idx[d] += 1;
                bset(s, L, true);
                depth = L + 1;
            }
            if (!finished) {
                t = Iv{ dyad(idx[0], lev[0]), dyad(idx[0] + 1, lev[0]) };
                u = Iv{ dyad(idx[1], lev[1]), dyad(idx[1] + 1, lev[1]) };
                v = Iv{ dyad(idx[2], lev[2]), dyad(idx[2] + 1, lev[2]) };
            }
        }

        if (!finished) {
            if (!cache_ok || c_idx != idx[0] || c_lev != lev[0] || phase == 1) {
                t_part(a, q, t, A0, DA, B0, DB);
                cache_ok = phase == 2;
                c_idx = idx[0];
                c_lev = lev[0];
            }
            const bool z = contains_zero(q.kind, A0, DA, B0, DB, u, v);
            evals++;

            if (phase == 1) {
                if (z) tol[0] /= 1e2;
                phase = 2;
                depth = 0;
                s.w[0] = s.w[1] = s.w[2] = 0;
                mt = s;
                ma = s;
                idx[0] = idx[1] = idx[2] = 0;
                lev[0] = lev[1] = lev[2] = 0;
                cache_ok = false;
                continue;
            }

            bool leaf = !z;
            if (z) {
                const double w0 = next_up(ldexp(1.0, -lev[0]));
                const double w1 = next_up(ldexp(1.0, -lev[1]));
                const double w2 = next_up(ldexp(1.0, -lev[2]));
                const double w[3] = { w0, w1, w2 };
                const bool all_sat = w0 <= tol[0] && w1 <= tol[1] && w2 <= tol[2];
                const bool all_zero = w0 <= 1e-10 && w1 <= 1e-10 && w2 <= 1e-10;
                if ((idx[0] > 0 || all_zero) && all_sat) {
                    best = t.lo;
                    hit = 1;
                    if (a.use_global_best)
                        atomicMin(a.global_best, (unsigned long long)__double_as_longlong(best));
                    leaf = true;
                } else {
                    int sp = -1;
#pragma unroll
                    for (int i = 0; i < 3; i++)
                        if ((all_sat || w[i] > tol[i]) && (sp == -1 || w[i] * tol[sp] > w[sp] * tol[i])) sp = i;
                    if (depth >= MAX_DEPTH) {
                        hit = -2;
                        finished = true;
                    } else {
                        bset(s, depth, false);
                        bset(mt, depth, sp == 0);
                        bset(ma, depth, sp == 1);
                        idx[sp] <<= 1;
                        lev[sp] += 1;
                        depth++;
                    }
                }
            }
            if (leaf && !finished) {
                const int L = last_zero_below(s, depth);
                if (L < 0) {
                    finished = true;
                } else {
                    const int ct = popcount_range(mt, L + 1, depth), ca = popcount_range(ma, L + 1, depth);
                    const int cb = depth - 1 - L - ct - ca;
                    idx[0] >>= ct; lev[0] -= ct;
                    idx[1] >>= ca; lev[1] -= ca;
                    idx[2] >>= cb; lev[2] -= cb;
                    const int d = bget(mt, L) ? 0 : (bget(ma, L) ? 1 : 2);
                    idx[d] += 1;
                    bset(s, L, true);
                    depth = L + 1;
                }
            }
        }

        if (finished) {
            if (a.toi_out) a.toi_out[qi] = (hit == 1) ? best : __longlong_as_double(0x7ff0000000000000LL);
            if (a.hit_out) a.hit_out[qi] = (signed char)hit;
            atomicAdd(a.total_evals, (unsigned long long)evals);
            phase = 0;
        }
    }
}

}  // namespace

void launch_ccd(const CcdArgs& a, int grid, int block, hipStream_t stream) {
    hipLaunchKernelGGL(ccd_kernel, dim3(grid), dim3(block), 0, stream, a);
}

}  // namespace rigc
