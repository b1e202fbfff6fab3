// Rigid broad phase for gfx950: the candidate set of Rigid IPC's hash-grid
// detection (src/ccd/rigid/broad_phase.cpp:19-203,
// src/ccd/rigid/rigid_body_hash_grid.cpp:17-344) computed as
//
//   1. body boxes (RigidBody::compute_bounding_box, rigid_body.cpp:246-272),
//      one workgroup per body;
//   2. included bodies: those in any overlapping body-box pair of different
//      groups (close_bodies + body_pairs_to_body_ids);
//   3. the conservative domain over included bodies (resize, :90-119);
//   4. per-vertex AABBs: for continuous detection the interval enclosure of
//      every vertex trajectory, with the forced split of t at 1/2 and the
//      adaptive bisection of any t-piece whose enclosure leaves the domain
//      (compute_vertices_intervals, :164-246), one workgroup per body with
//      an LDS stack of t-pieces taken in batches; for discrete detection the world vertex
//      inflated by r;
//   5. element AABBs (vertex / edge / face) in the centre / half-extent form
//      used by the overlap test;
//   6. body pairs whose element hulls overlap (a conservative prefilter);
//   7. one workgroup per body pair: each body's elements that reach the
//      other body's hull are staged in LDS, every edge-edge and face-vertex
//      pair among them is tested, and hits are appended with one atomic per
//      wavefront;
//   8. a stable radix sort of the pair keys, giving the sorted, duplicate-
//      free candidate lists of the reference (rigc_broad_phase). The
//      earliest-TOI path, whose answer is a minimum over the candidates,
//      and the constraint build, which classifies the packed keys and
//      sorts per class, skip it.
// Every element pair of two included bodies with overlapping AABBs and
// different groups is found exactly once (each pair belongs to exactly one
// body pair), so the result equals the reference's hash-grid set, which is
// independent of its cell size.
#include <algorithm>
#include <vector>

#include "rigc_geom.hpp"
#include "rigc_internal.hpp"
#include "rigc_rigid_iv.hpp"

namespace rigc {
namespace {

constexpr int BP_THREADS = 256;
constexpr int LIST_CAP = 1024;

__device__ __forceinline__ void block_minmax6(double* v, double (*red)[BP_THREADS]) {
    // v[0..2] are minima, v[3..5] maxima
    const int tid = threadIdx.x;
    for (int k = 0; k < 6; k++) red[k][tid] = v[k];
    __syncthreads();
    for (int s = BP_THREADS / 2; s > 0; s >>= 1) {
        if (tid < s) {
            for (int k = 0; k < 3; k++) red[k][tid] = fmin(red[k][tid], red[k][tid + s]);
            for (int k = 3; k < 6; k++) red[k][tid] = fmax(red[k][tid], red[k][tid + s]);
        }
        __syncthreads();
    }
    for (int k = 0; k < 6; k++) v[k] = red[k][0];
    __syncthreads();
}

__global__ __launch_bounds__(BP_THREADS) void body_box_kernel(const int* bodies, const double* V, const int* bvs,
                                                             const uint8_t* is_static, const double* r_max,
                                                             const double* p0, const double* p1, double infl,
                                                             double* box_raw, double* box_inf, int* inc,
                                                             unsigned long long* counters) {
    __shared__ double R0[9], R1[9];
    __shared__ double red[6][BP_THREADS];
    const int b = bodies[blockIdx.x], tid = threadIdx.x;
    // clears for the kernels that follow: the inclusion flag of each
    // selected body and the pair counters
    if (tid == 0) inc[b] = 0;
    if (blockIdx.x == 0 && tid < 4) counters[tid] = 0;
    const double* q0 = p0 + 6 * b;
    const double* q1 = p1 + 6 * b;
    const bool same = q0[3] == q1[3] && q0[4] == q1[4] && q0[5] == q1[5];
    double v[6];
    if (is_static[b] || same) {
        if (tid == 0) rotation_d(q0 + 3, R0);
        if (tid == 1) rotation_d(q1 + 3, R1);
        __syncthreads();
        for (int k = 0; k < 3; k++) { v[k] = INFINITY; v[3 + k] = -INFINITY; }
        for (int i = bvs[b] + tid; i < bvs[b + 1]; i += BP_THREADS) {
            double w0[3], w1[3];
            world_vertex_d(R0, q0, V + 3 * i, w0);
            world_vertex_d(R1, q1, V + 3 * i, w1);
            for (int k = 0; k < 3; k++) {
                v[k] = fmin(v[k], fmin(w0[k], w1[k]));
                v[3 + k] = fmax(v[3 + k], fmax(w0[k], w1[k]));
            }
        }
        block_minmax6(v, red);
    } else {
        for (int k = 0; k < 3; k++) {
            v[k] = fmin(q0[k], q1[k]) - r_max[b];
            v[3 + k] = fmax(q0[k], q1[k]) + r_max[b];
        }
    }
    if (tid == 0)
        for (int k = 0; k < 3; k++) {
            box_raw[6 * b + k] = v[k];
            box_raw[6 * b + 3 + k] = v[3 + k];
            box_inf[6 * b + k] = v[k] - infl;
            box_inf[6 * b + 3 + k] = v[3 + k] + infl;
        }
}

// Body i is included when its box overlaps the box of a body of another
// group in its scene. One block per upper-triangular tile of a scene's
// body-pair grid; both bodies of an overlapping pair are marked.
// inc[] is cleared beforehand and only ever set to 1.
constexpr int PAIR_TILE = 64;
// Threads per tile: each of the four wavefronts takes a quarter of the
// tile's columns for all 64 rows (lane = row).
constexpr int TILE_THREADS = 256;
constexpr int TILE_COLS = PAIR_TILE / (TILE_THREADS / 64);

__global__ __launch_bounds__(TILE_THREADS) void included_kernel(const int4* tiles, const int* group,
                                                               const double* box_inf, int* inc) {
    __shared__ double tb[PAIR_TILE][6];
    __shared__ int tg[PAIR_TILE];
    const int4 T = tiles[blockIdx.x];
    const int end = T.z;
    const int row = threadIdx.x & 63, quarter = threadIdx.x >> 6;
    const int i = T.x + row;
    const int j0 = T.y;
    if (threadIdx.x < PAIR_TILE) {
        const int jt = j0 + threadIdx.x;
        if (jt < end) {
            for (int k = 0; k < 6; k++) tb[threadIdx.x][k] = box_inf[6 * jt + k];
            tg[threadIdx.x] = group[jt];
        }
    }
    __syncthreads();
    if (i >= end) return;
    double mi[6];
    for (int k = 0; k < 6; k++) mi[k] = box_inf[6 * i + k];
    const int gi = group[i];
    const int m = min(PAIR_TILE, end - j0);
    bool hit_i = false;
    for (int jj = quarter * TILE_COLS; jj < min(m, (quarter + 1) * TILE_COLS); jj++) {
        const int j = j0 + jj;
        // all six bounds are read before any is compared (no chain of
        // dependent LDS reads): bitwise &, not &&
        const double* o = tb[jj];
        const double o0 = o[0], o1 = o[1], o2 = o[2], o3 = o[3], o4 = o[4], o5 = o[5];
        const bool hit = (j > i) & (tg[jj] != gi) & (mi[0] <= o3) & (o0 <= mi[3]) & (mi[1] <= o4) & (o1 <= mi[4]) &
                         (mi[2] <= o5) & (o2 <= mi[5]);
        if (hit) {
            hit_i = true;
            inc[j] = 1;
        }
    }
    if (hit_i) inc[i] = 1;
}

// Domain of each selected scene: one block per scene over its bodies.
__global__ __launch_bounds__(BP_THREADS) void domain_kernel(const int* scenes, const int* scene_body, const int* inc,
                                                           const double* box_raw, double r, double* dom_all) {
    __shared__ double red[6][BP_THREADS];
    const int sc = scenes[blockIdx.x];
    double* dom = dom_all + 8 * sc;
    double v[6];
    for (int k = 0; k < 3; k++) { v[k] = INFINITY; v[3 + k] = -INFINITY; }
    for (int b = scene_body[sc] + threadIdx.x; b < scene_body[sc + 1]; b += BP_THREADS)
        if (inc[b]) {
            for (int k = 0; k < 3; k++) {
                v[k] = fmin(v[k], box_raw[6 * b + k]);
                v[3 + k] = fmax(v[3 + k], box_raw[6 * b + 3 + k]);
            }
        }
    block_minmax6(v, red);
    if (threadIdx.x == 0) {
        const double dx = v[3] - v[0], dy = v[4] - v[1], dz = v[5] - v[2];
        const double diag = sqrt((dx * dx + dy * dy) + dz * dz);
        for (int k = 0; k < 3; k++) {
            double mn = v[k] - 1e-8 * diag, mx = v[3 + k] + 1e-8 * diag;
            mn -= r;
            mn += r;
            dom[k] = mn;
            dom[3 + k] = mx;
        }
    }
}

// Interval enclosure of every vertex trajectory of an included body. The
// t-pieces still to examine are kept on an LDS stack and taken in batches
// of up to 64: thread k computes the interval pose of piece k, then every
// thread checks its vertices against the domain under each piece of the
// batch. A piece that fits (or is deeper than level 40) adds its vertex
// intervals; one that does not (or the forced root) is split in two. Each
// piece's outcome depends on that piece alone, so the kept pieces, and the
// enclosure (a min / max over them), are those of the reference's
// depth-first recursion (the fit test without the inflation radius, see
// below). A batch is a single piece whenever the stack is
// within VI_ROOM of full: one piece adds at most one entry per level
// below it, so the stack never overflows.
constexpr int VI_STACK = 512, VI_BATCH = 64, VI_ROOM = 2 * VI_BATCH;
// Pieces of levels 1..VI_TABLE_LEV have their interval poses evaluated
// together, one per thread, before the splitting starts (a pose is a
// function of its t-interval only): a round that reaches them reads them
// instead of waiting for its own evaluation.
constexpr int VI_TABLE_LEV = 4, VI_TABLE = (2 << VI_TABLE_LEV) - 2;

__global__ __launch_bounds__(BP_THREADS) void vertex_interval_kernel(const int* bodies, const int* body_scene,
                                                                    const double* V, const int* bvs, const int* inc,
                                                                    const double* p0, const double* p1,
                                                                    const double* dom_all, double r, double* vmin,
                                                                    double* vmax, double* bhull) {
    __shared__ unsigned long long s_idx[VI_STACK];
    __shared__ signed char s_lev[VI_STACK], s_force[VI_STACK];
    __shared__ int sp, nbatch;
    __shared__ unsigned long long c_idx[VI_BATCH];
    __shared__ int c_lev[VI_BATCH], c_force[VI_BATCH], c_bad[VI_BATCH];
    __shared__ Iv Rs[VI_BATCH][3][3], ps[VI_BATCH][3];
    __shared__ Iv Rt[VI_TABLE][3][3], pt[VI_TABLE][3];  // levels 1..VI_TABLE_LEV, piece (l, i) at 2^l - 2 + i
    __shared__ double red[6][BP_THREADS];
    const int b = bodies[blockIdx.x], tid = threadIdx.x;
    if (!inc[b]) return;
    const double* dom = dom_all + 8 * body_scene[b];
    const int v0 = bvs[b], v1 = bvs[b + 1];
    const double* q0 = p0 + 6 * b;
    const double* q1 = p1 + 6 * b;
    double dmn[3], dmx[3];
    for (int k = 0; k < 3; k++) { dmn[k] = dom[k]; dmx[k] = dom[3 + k]; }
    for (int i = v0 + tid; i < v1; i += BP_THREADS)
        for (int k = 0; k < 3; k++) {
            vmin[3 * i + k] = INFINITY;
            vmax[3 * i + k] = -INFINITY;
        }
    if (tid == 0) {
        s_idx[0] = 0; s_lev[0] = 0; s_force[0] = 1;
        sp = 1;
    }
    if (tid < VI_TABLE) {
        const int lev = 31 - __clz(tid + 2);
        const unsigned long long idx = (unsigned long long)(tid + 2 - (1 << lev));
        const Iv t = Iv{ ldexp((double)idx, -lev), ldexp((double)(idx + 1), -lev) };
        Iv p[3], R[3][3];
        pose_iv(q0, q1, t, p, R);
        for (int i = 0; i < 3; i++) {
            pt[tid][i] = p[i];
            for (int j = 0; j < 3; j++) Rt[tid][i][j] = R[i][j];
        }
    }
    __syncthreads();
    while (sp > 0) {
        // take the top k pieces
        const int top = sp;
        const int k = (top + VI_BATCH <= VI_STACK - VI_ROOM) ? min(top, VI_BATCH) : 1;
        if (tid < k) {
            c_idx[tid] = s_idx[top - 1 - tid];
            c_lev[tid] = s_lev[top - 1 - tid];
            c_force[tid] = s_force[top - 1 - tid];
            c_bad[tid] = 0;
            const bool evaluate = c_force[tid] <= 0 || c_lev[tid] > 40;
            if (evaluate) {
                const unsigned long long idx = c_idx[tid];
                const int lev = c_lev[tid];
                if (lev >= 1 && lev <= VI_TABLE_LEV) {
                    const int e = (1 << lev) - 2 + (int)idx;
                    for (int i = 0; i < 3; i++) {
                        ps[tid][i] = pt[e][i];
                        for (int j = 0; j < 3; j++) Rs[tid][i][j] = Rt[e][i][j];
                    }
                } else {
                    const Iv t = Iv{ ldexp((double)idx, -lev), ldexp((double)(idx + 1), -lev) };
                    Iv p[3], R[3][3];
                    pose_iv(q0, q1, t, p, R);
                    for (int i = 0; i < 3; i++) {
                        ps[tid][i] = p[i];
                        for (int j = 0; j < 3; j++) Rs[tid][i][j] = R[i][j];
                    }
                }
            }
        }
        if (tid == 0) nbatch = k;
        __syncthreads();
        const int nb = nbatch;
        // domain check of every evaluated piece
        for (int i = v0 + tid; i < v1; i += BP_THREADS)
            for (int c = 0; c < nb; c++) {
                if (!(c_force[c] <= 0 || c_lev[c] > 40) || c_bad[c]) continue;
                Iv w[3];
                world_vertex_iv(Rs[c], ps[c], V + 3 * i, w);
                bool bad = false;
                // the reference also subtracts / adds the inflation radius r
                // here; with r > 0 a vertex on the domain's edge then never
                // fits and the split never ends, so the test leaves r out
                // (identical for r = 0, the CCD's radius without separation)
                for (int d = 0; d < 3; d++) bad = bad || (w[d].lo < dmn[d]) || (w[d].hi > dmx[d]);
                if (bad) c_bad[c] = 1;
            }
        __syncthreads();
        // kept pieces add their vertex intervals
        for (int i = v0 + tid; i < v1; i += BP_THREADS) {
            double mn[3] = { vmin[3 * i], vmin[3 * i + 1], vmin[3 * i + 2] };
            double mx[3] = { vmax[3 * i], vmax[3 * i + 1], vmax[3 * i + 2] };
            bool any = false;
            for (int c = 0; c < nb; c++) {
                const bool evaluate = c_force[c] <= 0 || c_lev[c] > 40;
                if (!evaluate || !(c_bad[c] == 0 || c_lev[c] > 40)) continue;
                Iv w[3];
                world_vertex_iv(Rs[c], ps[c], V + 3 * i, w);
                for (int d = 0; d < 3; d++) {
                    mn[d] = fmin(mn[d], w[d].lo);
                    mx[d] = fmax(mx[d], w[d].hi);
                }
                any = true;
            }
            if (any)
                for (int d = 0; d < 3; d++) {
                    vmin[3 * i + d] = mn[d];
                    vmax[3 * i + d] = mx[d];
                }
        }
        // the others are split: children pushed in order, second half first
        if (tid == 0) {
            int s = top - nb;
            for (int c = nb - 1; c >= 0; c--) {
                const bool evaluate = c_force[c] <= 0 || c_lev[c] > 40;
                const bool keep = evaluate && (c_bad[c] == 0 || c_lev[c] > 40);
                if (keep) continue;
                s_idx[s] = 2 * c_idx[c] + 1; s_lev[s] = (signed char)(c_lev[c] + 1);
                s_force[s] = (signed char)(c_force[c] - 1);
                s_idx[s + 1] = 2 * c_idx[c]; s_lev[s + 1] = (signed char)(c_lev[c] + 1);
                s_force[s + 1] = (signed char)(c_force[c] - 1);
                s += 2;
            }
            sp = s;
        }
        __syncthreads();
    }
    double v[6];
    for (int k = 0; k < 3; k++) { v[k] = INFINITY; v[3 + k] = -INFINITY; }
    for (int i = v0 + tid; i < v1; i += BP_THREADS)
        for (int k = 0; k < 3; k++) {
            const double mn = vmin[3 * i + k] - r, mx = vmax[3 * i + k] + r;
            vmin[3 * i + k] = mn;
            vmax[3 * i + k] = mx;
            v[k] = fmin(v[k], mn);
            v[3 + k] = fmax(v[3 + k], mx);
        }
    block_minmax6(v, red);
    if (tid == 0)
        for (int k = 0; k < 6; k++) bhull[6 * b + k] = v[k];
}

// Discrete detection: world vertex +- r.
__global__ __launch_bounds__(BP_THREADS) void vertex_static_kernel(const int* bodies, const double* V, const int* bvs,
                                                                  const int* inc, const double* poses, double r,
                                                                  double* vmin, double* vmax, double* bhull) {
    __shared__ double R[9];
    __shared__ double red[6][BP_THREADS];
    const int b = bodies[blockIdx.x], tid = threadIdx.x;
    if (!inc[b]) return;
    const double* q = poses + 6 * b;
    if (tid == 0) rotation_d(q + 3, R);
    __syncthreads();
    double v[6];
    for (int k = 0; k < 3; k++) { v[k] = INFINITY; v[3 + k] = -INFINITY; }
    for (int i = bvs[b] + tid; i < bvs[b + 1]; i += BP_THREADS) {
        double w[3];
        world_vertex_d(R, q, V + 3 * i, w);
        for (int k = 0; k < 3; k++) {
            const double mn = w[k] - r, mx = w[k] + r;
            vmin[3 * i + k] = mn;
            vmax[3 * i + k] = mx;
            v[k] = fmin(v[k], mn);
            v[3 + k] = fmax(v[3 + k], mx);
        }
    }
    block_minmax6(v, red);
    if (tid == 0)
        for (int k = 0; k < 6; k++) bhull[6 * b + k] = v[k];
}

__device__ __forceinline__ void center_half(const double* mn, const double* mx, double* c, double* h) {
    for (int k = 0; k < 3; k++) {
        const double hh = (mx[k] - mn[k]) / 2;
        h[k] = hh;
        c[k] = mn[k] + hh;
    }
}

// Element boxes of one included body per block.
__global__ __launch_bounds__(BP_THREADS) void element_aabb_kernel(const int* bodies, const int* bvs, const int* bes,
                                                                 const int* bfs, const int* E, const int* F,
                                                                 const int* inc, const double* vmin,
                                                                 const double* vmax, double* vc, double* vh,
                                                                 double* ec, double* eh, double* fc, double* fh) {
    const int b = bodies[blockIdx.x];
    if (!inc[b]) return;
    for (int i = bvs[b] + threadIdx.x; i < bvs[b + 1]; i += BP_THREADS)
        center_half(vmin + 3 * i, vmax + 3 * i, vc + 3 * i, vh + 3 * i);
    for (int e = bes[b] + threadIdx.x; e < bes[b + 1]; e += BP_THREADS) {
        const int a = E[2 * e], c = E[2 * e + 1];
        double mn[3], mx[3];
        for (int k = 0; k < 3; k++) {
            mn[k] = fmin(vmin[3 * a + k], vmin[3 * c + k]);
            mx[k] = fmax(vmax[3 * a + k], vmax[3 * c + k]);
        }
        center_half(mn, mx, ec + 3 * e, eh + 3 * e);
    }
    for (int f = bfs[b] + threadIdx.x; f < bfs[b + 1]; f += BP_THREADS) {
        const int a = F[3 * f], c = F[3 * f + 1], d = F[3 * f + 2];
        double mn[3], mx[3];
        for (int k = 0; k < 3; k++) {
            mn[k] = fmin(fmin(vmin[3 * a + k], vmin[3 * c + k]), vmin[3 * d + k]);
            mx[k] = fmax(fmax(vmax[3 * a + k], vmax[3 * c + k]), vmax[3 * d + k]);
        }
        center_half(mn, mx, fc + 3 * f, fh + 3 * f);
    }
}

__device__ __forceinline__ double hull_eps(const double* h) {
    double m = 1.0;
    for (int k = 0; k < 6; k++) m = fmax(m, fabs(h[k]));
    return 1e-12 * m;
}

// Close body pairs (a < b) of one scene: one block per upper-triangular
// tile; hits appended with one atomic per wavefront.
__global__ __launch_bounds__(TILE_THREADS) void body_pair_kernel(const int4* tiles, const int* inc,
                                                                const int* group, const double* bhull, int* pairs,
                                                                long long cap, unsigned long long* count) {
    __shared__ double th[PAIR_TILE][6];
    __shared__ int tg[PAIR_TILE], ti[PAIR_TILE];
    const int4 T = tiles[blockIdx.x];
    const int end = T.z;
    const int lane = threadIdx.x & 63, quarter = threadIdx.x >> 6;
    const int a = T.x + lane;
    const int b0 = T.y;
    if (threadIdx.x < PAIR_TILE) {
        const int bb = b0 + threadIdx.x;
        if (bb < end) {
            ti[threadIdx.x] = inc[bb];
            tg[threadIdx.x] = group[bb];
            for (int k = 0; k < 6; k++) th[threadIdx.x][k] = bhull[6 * bb + k];
        }
    }
    __syncthreads();
    double ha[6] = { 0, 0, 0, 0, 0, 0 };
    int ga = 0;
    bool ia = false;
    double ea = 0;
    if (a < end) {
        ia = inc[a] != 0;
        ga = group[a];
        for (int k = 0; k < 6; k++) ha[k] = bhull[6 * a + k];
        if (ia) ea = hull_eps(ha);
    }
    const int m = min(PAIR_TILE, end - b0);
    for (int j = quarter * TILE_COLS; j < min(m, (quarter + 1) * TILE_COLS); j++) {
        const int b = b0 + j;
        const double* th_j = th[j];
        double hb[6];
        for (int k = 0; k < 6; k++) hb[k] = th_j[k];
        const double eps = fmax(ea, hull_eps(hb));
        // bitwise &: the bounds are read together, not one after another
        const bool hit = ia & (b > a) & (ti[j] != 0) & (tg[j] != ga) & (ha[0] <= hb[3] + eps) &
                         (hb[0] <= ha[3] + eps) & (ha[1] <= hb[4] + eps) & (hb[1] <= ha[4] + eps) &
                         (ha[2] <= hb[5] + eps) & (hb[2] <= ha[5] + eps);
        const unsigned long long mask = __ballot(hit);
        if (mask) {
            const int leader = __ffsll((long long)mask) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(count, (unsigned long long)__popcll(mask));
            base = __shfl(base, leader);
            if (hit) {
                const unsigned long long k = base + __popcll(mask & ((1ULL << lane) - 1ULL));
                if ((long long)k < cap) {
                    pairs[2 * k] = a;
                    pairs[2 * k + 1] = b;
                }
            }
        }
    }
}

// Bitwise & in these tests: every operand is loaded up front instead of
// each load waiting on the previous comparison.
__device__ __forceinline__ bool aabb_overlap(const double* ca, const double* ha, const double* cb, const double* hb) {
    return (fabs(ca[0] - cb[0]) <= ha[0] + hb[0]) & (fabs(ca[1] - cb[1]) <= ha[1] + hb[1]) &
           (fabs(ca[2] - cb[2]) <= ha[2] + hb[2]);
}

__device__ __forceinline__ bool reaches(const double* c, const double* h, const double* hull, double eps) {
    return (c[0] + h[0] >= hull[0] - eps) & (c[0] - h[0] <= hull[3] + eps) & (c[1] + h[1] >= hull[1] - eps) &
           (c[1] - h[1] <= hull[4] + eps) & (c[2] + h[2] >= hull[2] - eps) & (c[2] - h[2] <= hull[5] + eps);
}

// Filtered element list of one body held in LDS, or the body's whole range
// when the filtered list does not fit.
struct ElemList {
    int* ids;
    int n;
    int start, count;
    bool full;
    __device__ int get(int k) const { return full ? start + k : ids[k]; }
};

__device__ void build_list(int* lds, int* cnt, int start, int end, const double* c, const double* h,
                           const double* hull, double eps) {
    for (int i = start + threadIdx.x; i < end; i += BP_THREADS)
        if (reaches(c + 3 * i, h + 3 * i, hull, eps)) {
            const int k = atomicAdd(cnt, 1);
            if (k < LIST_CAP) lds[k] = i;
        }
}

__device__ __forceinline__ void emit(bool hit, unsigned long long key, unsigned long long* out, long long cap,
                                     unsigned long long* count) {
    const int lane = threadIdx.x & 63;
    const unsigned long long mask = __ballot(hit);
    if (!mask) return;
    const int leader = __ffsll((long long)mask) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(count, (unsigned long long)__popcll(mask));
    base = __shfl(base, leader);
    if (hit) {
        const unsigned long long k = base + __popcll(mask & ((1ULL << lane) - 1ULL));
        if ((long long)k < cap) out[k] = key;
    }
}

__global__ __launch_bounds__(BP_THREADS) void element_pair_kernel(
    const int* pairs, const int* bvs, const int* bes, const int* bfs, const int* group, const double* bhull,
    const double* vc, const double* vh, const double* ec, const double* eh, const double* fc, const double* fh,
    long long nv, long long ne, unsigned long long* out_ee, long long cap_ee, unsigned long long* out_fv,
    long long cap_fv, unsigned long long* counters, const unsigned long long* n_pairs, long long cap_pairs) {
    __shared__ int l_ae[LIST_CAP], l_af[LIST_CAP], l_av[LIST_CAP], l_be[LIST_CAP], l_bf[LIST_CAP], l_bv[LIST_CAP];
    __shared__ int n_ae, n_af, n_av, n_be, n_bf, n_bv;
    // body pairs in grid stride; their number is read on the device, so no
    // host round trip separates the two pair kernels
    const long long np = (long long)min(*n_pairs, (unsigned long long)cap_pairs);
    for (long long pi = blockIdx.x; pi < np; pi += gridDim.x) {
        __syncthreads();  // the previous pair's lists are no longer read
        const int A = pairs[2 * pi], B = pairs[2 * pi + 1];
        const double* hA = bhull + 6 * A;
        const double* hB = bhull + 6 * B;
        const double eps = fmax(hull_eps(hA), hull_eps(hB));
        if (threadIdx.x == 0) n_ae = n_af = n_av = n_be = n_bf = n_bv = 0;
        __syncthreads();
        build_list(l_ae, &n_ae, bes[A], bes[A + 1], ec, eh, hB, eps);
        build_list(l_af, &n_af, bfs[A], bfs[A + 1], fc, fh, hB, eps);
        build_list(l_av, &n_av, bvs[A], bvs[A + 1], vc, vh, hB, eps);
        build_list(l_be, &n_be, bes[B], bes[B + 1], ec, eh, hA, eps);
        build_list(l_bf, &n_bf, bfs[B], bfs[B + 1], fc, fh, hA, eps);
        build_list(l_bv, &n_bv, bvs[B], bvs[B + 1], vc, vh, hA, eps);
        __syncthreads();
        const ElemList ae{ l_ae, n_ae, bes[A], bes[A + 1] - bes[A], n_ae > LIST_CAP };
        const ElemList af{ l_af, n_af, bfs[A], bfs[A + 1] - bfs[A], n_af > LIST_CAP };
        const ElemList av{ l_av, n_av, bvs[A], bvs[A + 1] - bvs[A], n_av > LIST_CAP };
        const ElemList be{ l_be, n_be, bes[B], bes[B + 1] - bes[B], n_be > LIST_CAP };
        const ElemList bf{ l_bf, n_bf, bfs[B], bfs[B + 1] - bfs[B], n_bf > LIST_CAP };
        const ElemList bv{ l_bv, n_bv, bvs[B], bvs[B + 1] - bvs[B], n_bv > LIST_CAP };
        auto len = [](const ElemList& l) { return l.full ? l.count : l.n; };

        // edge-edge: edges of A (lower ids) against edges of B
        {
            const long long na = len(ae), nbb = len(be), tot = na * nbb;
            for (long long k0 = 0; k0 < tot; k0 += BP_THREADS) {
                const long long k = k0 + threadIdx.x;
                bool hit = false;
                unsigned long long key = 0;
                if (k < tot) {
                    const int ea = ae.get((int)(k / nbb)), eb = be.get((int)(k % nbb));
                    hit = aabb_overlap(ec + 3 * ea, eh + 3 * ea, ec + 3 * eb, eh + 3 * eb);
                    key = (unsigned long long)ea * (unsigned long long)ne + (unsigned long long)eb;
                }
                emit(hit, key, out_ee, cap_ee, counters + 0);
            }
        }
        // face-vertex: faces of A with vertices of B, and faces of B with vertices of A
        for (int dir = 0; dir < 2; dir++) {
            const ElemList& lf = dir == 0 ? af : bf;
            const ElemList& lv = dir == 0 ? bv : av;
            const long long nfl = len(lf), nvl = len(lv), tot = nfl * nvl;
            for (long long k0 = 0; k0 < tot; k0 += BP_THREADS) {
                const long long k = k0 + threadIdx.x;
                bool hit = false;
                unsigned long long key = 0;
                if (k < tot) {
                    const int f = lf.get((int)(k / nvl)), v = lv.get((int)(k % nvl));
                    hit = aabb_overlap(fc + 3 * f, fh + 3 * f, vc + 3 * v, vh + 3 * v);
                    key = (unsigned long long)f * (unsigned long long)nv + (unsigned long long)v;
                }
                emit(hit, key, out_fv, cap_fv, counters + 1);
            }
        }
    }
}

__global__ void unpack_kernel(const unsigned long long* keys, long long n, unsigned long long m, int* out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const unsigned long long k = keys[i];
        out[2 * i] = (int)(k / m);
        out[2 * i + 1] = (int)(k % m);
    }
}

int bit_length(unsigned long long x) {
    int b = 0;
    while (x) { b++; x >>= 1; }
    return b;
}

}  // namespace

#define BP_CHECK(x)                      \
    do {                                 \
        hipError_t e_ = (x);             \
        if (e_ != hipSuccess) return e_; \
    } while (0)

namespace {

// Per-body stages (boxes, inclusion, domain, vertex and element boxes).
hipError_t queue_element_boxes(BroadPhase& bp, const DevScene& sc, const SceneSel& sel, const double* d_p0,
                               const double* d_p1, double r, bool continuous, hipStream_t st,
                               bool element_boxes = true) {
    const int nb = sc.nb;
    BP_CHECK(bp.box_raw.ensure(6 * sizeof(double) * nb));
    BP_CHECK(bp.box_inf.ensure(6 * sizeof(double) * nb));
    BP_CHECK(bp.inc.ensure(sizeof(int) * nb));
    BP_CHECK(bp.dom.ensure(8 * sizeof(double) * sc.n_scenes));
    BP_CHECK(bp.bhull.ensure(6 * sizeof(double) * nb));
    BP_CHECK(bp.vmin.ensure(3 * sizeof(double) * std::max(sc.nv, 1)));
    BP_CHECK(bp.vmax.ensure(3 * sizeof(double) * std::max(sc.nv, 1)));
    BP_CHECK(bp.vc.ensure(3 * sizeof(double) * std::max(sc.nv, 1)));
    BP_CHECK(bp.vh.ensure(3 * sizeof(double) * std::max(sc.nv, 1)));
    BP_CHECK(bp.ec.ensure(3 * sizeof(double) * std::max(sc.ne, 1)));
    BP_CHECK(bp.eh.ensure(3 * sizeof(double) * std::max(sc.ne, 1)));
    BP_CHECK(bp.fc.ensure(3 * sizeof(double) * std::max(sc.nf, 1)));
    BP_CHECK(bp.fh.ensure(3 * sizeof(double) * std::max(sc.nf, 1)));
    BP_CHECK(bp.counters.ensure(4 * sizeof(unsigned long long)));
    unsigned long long* cnt = bp.counters.as<unsigned long long>();

    const double* p1 = continuous ? d_p1 : d_p0;
    const unsigned nbs = (unsigned)sel.n_bodies;
    const unsigned ntiles = (unsigned)sel.n_tiles;
    hipLaunchKernelGGL(body_box_kernel, dim3(nbs), dim3(BP_THREADS), 0, st, sel.bodies, sc.V, sc.body_vstart,
                       sc.is_static, sc.r_max, d_p0, p1, r, bp.box_raw.as<double>(), bp.box_inf.as<double>(),
                       bp.inc.as<int>(), cnt);
    hipLaunchKernelGGL(included_kernel, dim3(ntiles), dim3(TILE_THREADS), 0, st, sel.tiles, sc.group,
                       bp.box_inf.as<double>(), bp.inc.as<int>());
    hipLaunchKernelGGL(domain_kernel, dim3((unsigned)sel.n_scenes), dim3(BP_THREADS), 0, st, sel.scenes,
                       sc.scene_body, bp.inc.as<int>(), bp.box_raw.as<double>(), r, bp.dom.as<double>());
    if (continuous)
        hipLaunchKernelGGL(vertex_interval_kernel, dim3(nbs), dim3(BP_THREADS), 0, st, sel.bodies, sc.body_scene,
                           sc.V, sc.body_vstart, bp.inc.as<int>(), d_p0, d_p1, bp.dom.as<double>(), r,
                           bp.vmin.as<double>(), bp.vmax.as<double>(), bp.bhull.as<double>());
    else
        hipLaunchKernelGGL(vertex_static_kernel, dim3(nbs), dim3(BP_THREADS), 0, st, sel.bodies, sc.V,
                           sc.body_vstart, bp.inc.as<int>(), d_p0, r, bp.vmin.as<double>(), bp.vmax.as<double>(),
                           bp.bhull.as<double>());
    if (!element_boxes) return hipGetLastError();
    hipLaunchKernelGGL(element_aabb_kernel, dim3(nbs), dim3(BP_THREADS), 0, st, sel.bodies, sc.body_vstart,
                       sc.body_estart, sc.body_fstart, sc.E, sc.F, bp.inc.as<int>(), bp.vmin.as<double>(),
                       bp.vmax.as<double>(), bp.vc.as<double>(), bp.vh.as<double>(), bp.ec.as<double>(),
                       bp.eh.as<double>(), bp.fc.as<double>(), bp.fh.as<double>());
    return hipGetLastError();
}

// Body pairs at the capacity in bp (count in counters[2]).
hipError_t queue_body_pair_pass(BroadPhase& bp, const DevScene& sc, const SceneSel& sel, bool clear,
                                hipStream_t st) {
    unsigned long long* cnt = bp.counters.as<unsigned long long>();
    BP_CHECK(bp.pairs.ensure(2 * sizeof(int) * bp.cap_pairs));
    if (clear) BP_CHECK(hipMemsetAsync(cnt, 0, 4 * sizeof(unsigned long long), st));
    hipLaunchKernelGGL(body_pair_kernel, dim3((unsigned)sel.n_tiles), dim3(TILE_THREADS), 0, st, sel.tiles,
                       bp.inc.as<int>(), sc.group, bp.bhull.as<double>(), bp.pairs.as<int>(), bp.cap_pairs, cnt + 2);
    return hipGetLastError();
}

// Body pairs, then element pairs of every body pair, at the capacities in
// bp (the counters hold the true counts afterwards, also past a capacity).
hipError_t queue_pairs(BroadPhase& bp, const DevScene& sc, const SceneSel& sel, bool clear, hipStream_t st) {
    unsigned long long* cnt = bp.counters.as<unsigned long long>();
    BP_CHECK(bp.keys_ee.ensure(8 * bp.cap_ee));
    BP_CHECK(bp.keys_fv.ensure(8 * bp.cap_fv));
    BP_CHECK(queue_body_pair_pass(bp, sc, sel, clear, st));
    const unsigned grid = (unsigned)std::min<long long>(bp.cap_pairs, 2048);
    hipLaunchKernelGGL(element_pair_kernel, dim3(grid), dim3(BP_THREADS), 0, st, bp.pairs.as<int>(), sc.body_vstart,
                       sc.body_estart, sc.body_fstart, sc.group, bp.bhull.as<double>(), bp.vc.as<double>(),
                       bp.vh.as<double>(), bp.ec.as<double>(), bp.eh.as<double>(), bp.fc.as<double>(),
                       bp.fh.as<double>(), (long long)sc.nv, (long long)sc.ne, bp.keys_ee.as<unsigned long long>(),
                       bp.cap_ee, bp.keys_fv.as<unsigned long long>(), bp.cap_fv, cnt, cnt + 2, bp.cap_pairs);
    return hipGetLastError();
}

}  // namespace

hipError_t queue_body_boxes(BroadPhase& bp, const DevScene& sc, const SceneSel& sel, const double* d_p0,
                            const double* d_p1, double r, hipStream_t st) {
    const int nb = sc.nb;
    BP_CHECK(bp.box_raw.ensure(6 * sizeof(double) * nb));
    BP_CHECK(bp.box_inf.ensure(6 * sizeof(double) * nb));
    BP_CHECK(bp.inc.ensure(sizeof(int) * nb));
    BP_CHECK(bp.counters.ensure(4 * sizeof(unsigned long long)));
    hipLaunchKernelGGL(body_box_kernel, dim3((unsigned)sel.n_bodies), dim3(BP_THREADS), 0, st, sel.bodies, sc.V,
                       sc.body_vstart, sc.is_static, sc.r_max, d_p0, d_p1, r, bp.box_raw.as<double>(),
                       bp.box_inf.as<double>(), bp.inc.as<int>(), bp.counters.as<unsigned long long>());
    return hipGetLastError();
}

static void broad_phase_capacities(BroadPhase& bp, const SceneSel& sel) {
    const long long held_pairs = (long long)(bp.pairs.bytes / (2 * sizeof(int)));
    if (bp.min_cap > 0) {
        bp.cap_pairs = std::max<long long>(bp.min_cap, held_pairs);
        bp.cap_ee = std::max<long long>(bp.min_cap, (long long)(bp.keys_ee.bytes / 8));
        bp.cap_fv = std::max<long long>(bp.min_cap, (long long)(bp.keys_fv.bytes / 8));
        return;
    }
    bp.cap_pairs = std::max<long long>(std::max<long long>(1024, 32LL * sel.n_bodies), held_pairs);
    bp.cap_ee = std::max<long long>(bp.keys_ee.bytes / 8, 1 << 16);
    bp.cap_fv = std::max<long long>(bp.keys_fv.bytes / 8, 1 << 16);
}

bool broad_phase_fits(BroadPhase& bp, const unsigned long long* got) {
    const bool fits = (long long)got[2] <= bp.cap_pairs && (long long)got[0] <= bp.cap_ee &&
                      (long long)got[1] <= bp.cap_fv;
    bp.cap_pairs = std::max<long long>(bp.cap_pairs, (long long)got[2]);
    bp.cap_ee = std::max<long long>(bp.cap_ee, (long long)got[0]);
    bp.cap_fv = std::max<long long>(bp.cap_fv, (long long)got[1]);
    return fits;
}

hipError_t queue_broad_phase(BroadPhase& bp, const DevScene& sc, const SceneSel& sel, const double* d_p0,
                             const double* d_p1, double r, bool continuous, hipStream_t st) {
    BP_CHECK(queue_element_boxes(bp, sc, sel, d_p0, d_p1, r, continuous, st));
    broad_phase_capacities(bp, sel);
    return queue_pairs(bp, sc, sel, false, st);
}

hipError_t requeue_pairs(BroadPhase& bp, const DevScene& sc, const SceneSel& sel, hipStream_t st) {
    return queue_pairs(bp, sc, sel, true, st);
}

hipError_t queue_body_pairs(BroadPhase& bp, const DevScene& sc, const SceneSel& sel, const double* d_poses, double r,
                            hipStream_t st) {
    BP_CHECK(queue_element_boxes(bp, sc, sel, d_poses, d_poses, r, false, st, false));
    broad_phase_capacities(bp, sel);
    return queue_body_pair_pass(bp, sc, sel, false, st);
}

hipError_t requeue_body_pairs(BroadPhase& bp, const DevScene& sc, const SceneSel& sel, hipStream_t st) {
    return queue_body_pair_pass(bp, sc, sel, true, st);
}

hipError_t run_broad_phase(BroadPhase& bp, const DevScene& sc, const SceneSel& sel, const double* d_p0,
                           const double* d_p1, double r, bool continuous, bool sorted, hipStream_t st) {
    bp.n_ee = bp.n_fv = bp.n_body_pairs = 0;
    if (sc.nb <= 1 || sel.n_bodies <= 1 || sel.n_tiles == 0) return hipSuccess;
    // Pairs with one read-back of the three counts ([0] edge-edge, [1]
    // face-vertex, [2] body pairs). A pass whose body pairs overflow sees
    // only part of them, so up to two more passes, each with capacities of
    // at least the counts seen.
    BP_CHECK(bp.h_counters.ensure(4 * sizeof(unsigned long long)));
    const unsigned long long* got = bp.h_counters.as<unsigned long long>();
    BP_CHECK(queue_broad_phase(bp, sc, sel, d_p0, d_p1, r, continuous, st));
    bool fits = false;
    for (int attempt = 0; attempt < 3 && !fits; attempt++) {
        if (attempt > 0) BP_CHECK(requeue_pairs(bp, sc, sel, st));
        BP_CHECK(hipMemcpyAsync(bp.h_counters.p, bp.counters.p, 3 * sizeof(unsigned long long),
                                hipMemcpyDeviceToHost, st));
        BP_CHECK(hipStreamSynchronize(st));
        fits = broad_phase_fits(bp, got);
    }
    if (!fits) return hipErrorUnknown;
    bp.n_body_pairs = (long long)got[2];
    bp.n_ee = (long long)got[0];
    bp.n_fv = (long long)got[1];
    return finish_broad_phase(bp, sc, sorted, st);
}

// Unpacks (and, when `sorted`, first sorts) the element-pair keys into int32 pairs.
hipError_t finish_broad_phase(BroadPhase& bp, const DevScene& sc, bool sorted, hipStream_t st) {
    const long long nmax = std::max(bp.n_ee, bp.n_fv);
    BP_CHECK(bp.tmp.ensure(8 * std::max<long long>(nmax, 1)));
    BP_CHECK(bp.hist.ensure(sizeof(unsigned int) * radix_sort_hist_entries(std::max<long long>(nmax, 1))));
    BP_CHECK(bp.out_ee.ensure(2 * sizeof(int) * std::max<long long>(bp.n_ee, 1)));
    BP_CHECK(bp.out_fv.ensure(2 * sizeof(int) * std::max<long long>(bp.n_fv, 1)));
    const unsigned long long me = (unsigned long long)sc.ne, mv = (unsigned long long)sc.nv;
    if (bp.n_ee) {
        if (sorted)
            BP_CHECK(radix_sort_u64(bp.keys_ee.as<unsigned long long>(), bp.tmp.as<unsigned long long>(),
                                    bp.hist.as<unsigned int>(), bp.n_ee, bit_length(me * me), st));
        hipLaunchKernelGGL(unpack_kernel, dim3((unsigned)((bp.n_ee + 255) / 256)), dim3(256), 0, st,
                           bp.keys_ee.as<unsigned long long>(), bp.n_ee, me, bp.out_ee.as<int>());
    }
    if (bp.n_fv) {
        if (sorted)
            BP_CHECK(radix_sort_u64(bp.keys_fv.as<unsigned long long>(), bp.tmp.as<unsigned long long>(),
                                    bp.hist.as<unsigned int>(), bp.n_fv,
                                    bit_length((unsigned long long)sc.nf * mv), st));
        hipLaunchKernelGGL(unpack_kernel, dim3((unsigned)((bp.n_fv + 255) / 256)), dim3(256), 0, st,
                           bp.keys_fv.as<unsigned long long>(), bp.n_fv, mv, bp.out_fv.as<int>());
    }
    return hipGetLastError();
}

}  // namespace rigc
