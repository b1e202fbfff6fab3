// BVH broad phase (DetectionMethod::BVH, the reference's default
// detection_method, src/SimState.cpp:178) on the device.
//
// Restates detect_collision_candidates_rigid_bvh (src/ccd/rigid/broad_phase.cpp:92-121
// discrete, :205-230 continuous) for meshes without codimensional elements:
//   * close body pairs: close_bodies_bvh (src/physics/rigid_body_assembler.cpp:307-356),
//     inflated body boxes overlapping (inclusive), i < j, different groups;
//   * per pair, sort_body_pair and detect_body_pair_collision_candidates_bvh
//     (src/ccd/rigid/rigid_body_bvh.hpp:14-94): the body A with fewer faces
//     has its vertices carried into body B's frame with interval poses,
//     ((V_A R_A^T) + (p_A - p_B)^T) R_B (the poses as point intervals, or
//     interpolated over t in [0, 1] when continuous), boxes grown by r;
//   * the query (f, *) of detect_body_pair_collision_candidates_from_aabbs
//     (src/ccd/rigid/rigid_body_bvh.cpp:84-190): B's faces whose box (from
//     B's local vertices) meets A's face box grown by r again; then FV
//     candidates through the face owning the vertex, EE candidates through
//     the faces owning both edges (MeshSelector: lowest face index).
// The reference finds B's faces with a BVH; the candidate set is fixed by
// the box tests alone, so here one workgroup per body pair tests A's faces
// against B's faces directly, after culling each side against the other's
// hull. EE candidates keep the reference's orientation (A's edge first).
#include "rigc_internal.hpp"
#include "rigc_interval.hpp"
#include "rigc_rigid_iv.hpp"

namespace rigc {

namespace {

constexpr int BVH_THREADS = 256;
constexpr int BVH_TILE = 64;  // body tiles of the pair grid (as the hash grid's)

// ipc-toolkit AABB: min / max, center and half extent; overlap in the
// center form.
struct Box {
    double mn[3], mx[3], c[3], h[3];
};
RIGC_DI void box_finish(Box& b) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
        b.h[k] = (b.mx[k] - b.mn[k]) / 2;
        b.c[k] = b.mn[k] + b.h[k];
    }
}
RIGC_DI bool box_overlap(const Box& a, const Box& b) {
    return (fabs(a.c[0] - b.c[0]) <= a.h[0] + b.h[0]) & (fabs(a.c[1] - b.c[1]) <= a.h[1] + b.h[1]) &
           (fabs(a.c[2] - b.c[2]) <= a.h[2] + b.h[2]);
}
// box of A's vertex i (min / max in LDS)
RIGC_DI void a_vertex_box(const double* av, int i, double mn[3], double mx[3]) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
        mn[k] = av[6 * i + k];
        mx[k] = av[6 * i + 3 + k];
    }
}
// box of B's local vertex v grown by r (doubles)
RIGC_DI void b_vertex_box(const double* V, int v, double r, double mn[3], double mx[3]) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
        mn[k] = V[3 * v + k] - r;
        mx[k] = V[3 * v + k] + r;
    }
}
RIGC_DI Box merge2(const double amn[3], const double amx[3], const double bmn[3], const double bmx[3]) {
    Box r;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        r.mn[k] = fmin(amn[k], bmn[k]);
        r.mx[k] = fmax(amx[k], bmx[k]);
    }
    box_finish(r);
    return r;
}

// close_bodies_bvh over the selection's 64 x 64 body tiles.
__global__ __launch_bounds__(BVH_THREADS) void bvh_body_pair_kernel(const int4* tiles, const int* group,
                                                                    const double* box_inf, int* pairs, long long cap,
                                                                    unsigned long long* count) {
    __shared__ double tb[BVH_TILE][6];
    __shared__ int tg[BVH_TILE];
    const int4 T = tiles[blockIdx.x];
    const int end = T.z;
    const int lane = threadIdx.x & 63, quarter = threadIdx.x >> 6;
    const int a = T.x + lane;
    const int b0 = T.y;
    if (threadIdx.x < BVH_TILE) {
        const int bb = b0 + threadIdx.x;
        if (bb < end) {
            tg[threadIdx.x] = group[bb];
            for (int k = 0; k < 6; k++) tb[threadIdx.x][k] = box_inf[6 * bb + k];
        }
    }
    __syncthreads();
    double ba[6] = { 0, 0, 0, 0, 0, 0 };
    int ga = 0;
    const bool va = a < end;
    if (va) {
        ga = group[a];
        for (int k = 0; k < 6; k++) ba[k] = box_inf[6 * a + k];
    }
    const int m = min(BVH_TILE, end - b0);
    const int per = BVH_TILE / 4;
    for (int j = quarter * per; j < min(m, (quarter + 1) * per); j++) {
        const int b = b0 + j;
        const double* o = tb[j];
        const bool hit = va & (b > a) & (tg[j] != ga) & (ba[0] <= o[3]) & (o[0] <= ba[3]) & (ba[1] <= o[4]) &
                         (o[1] <= ba[4]) & (ba[2] <= o[5]) & (o[2] <= ba[5]);
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

struct BvhArgs {
    const double* V;
    const int* E;
    const int* F;
    const int* FE;
    const int* bvs;
    const int* bfs;
    const int* v2f;
    const int* e2f;
    const double* p0;
    const double* p1;
    double r;
    int continuous;
    const int* pairs;
    const unsigned long long* n_pairs;
    long long cap_pairs;
    unsigned long long* keys_ee;
    long long cap_ee;
    unsigned long long* keys_fv;
    long long cap_fv;
    unsigned long long* counts;  // [0] EE, [1] FV
    long long ne, nv;
    int max_faces;               // capacity of the face lists in LDS
};

RIGC_DI void emit(unsigned long long* keys, long long cap, unsigned long long* count, unsigned long long key) {
    const unsigned long long k = atomicAdd(count, 1ULL);
    if ((long long)k < cap) keys[k] = key;
}

// One workgroup per close body pair (grid-stride over the pairs). LDS: A's
// vertex boxes in B's frame, then the lists of A's and B's faces that meet
// the other side's hull (a face whose box misses the other body's hull
// meets none of its faces).
__global__ __launch_bounds__(BVH_THREADS) void bvh_pair_kernel(BvhArgs a) {
    extern __shared__ double lds[];
    __shared__ Iv RA[9], RB[9], D[3];
    __shared__ double part[BVH_THREADS / 64][12];  // per-wavefront hulls
    __shared__ double hull_q[6], hull_b[6];        // A's query boxes, B's vertices
    __shared__ int n_la, n_lb;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long long np = min((long long)*a.n_pairs, a.cap_pairs);
    for (long long pi = blockIdx.x; pi < np; pi += gridDim.x) {
        int A = a.pairs[2 * pi], B = a.pairs[2 * pi + 1];
        // sort_body_pair: the body with fewer BVH primitives (faces) first
        if (a.bfs[A + 1] - a.bfs[A] > a.bfs[B + 1] - a.bfs[B]) {
            const int t = A;
            A = B;
            B = t;
        }
        const int va0 = a.bvs[A], nva = a.bvs[A + 1] - va0;
        const int fa0 = a.bfs[A], nfa = a.bfs[A + 1] - fa0;
        const int vb0 = a.bvs[B], vb1 = a.bvs[B + 1];
        const int fb0 = a.bfs[B], nfb = a.bfs[B + 1] - fb0;
        double* av = lds;                                 // [nva][6]
        int* la = reinterpret_cast<int*>(lds + 6 * nva);  // [max_faces] A's faces
        int* lb = la + a.max_faces;                       // [max_faces] B's faces
        if (tid == 0) {
            Iv ra[3], rb[3], pa[3], pb[3];
            if (a.continuous) {
                const Iv t = Iv{ 0.0, 1.0 };
                pose_pos_iv(a.p0 + 6 * A, a.p1 + 6 * A, t, pa);
                pose_rot_iv(a.p0 + 6 * A, a.p1 + 6 * A, t, ra);
                pose_pos_iv(a.p0 + 6 * B, a.p1 + 6 * B, t, pb);
                pose_rot_iv(a.p0 + 6 * B, a.p1 + 6 * B, t, rb);
            } else {
                for (int k = 0; k < 3; k++) {
                    pa[k] = iv(a.p0[6 * A + k]);
                    ra[k] = iv(a.p0[6 * A + 3 + k]);
                    pb[k] = iv(a.p0[6 * B + k]);
                    rb[k] = iv(a.p0[6 * B + 3 + k]);
                }
            }
            Iv R[3][3];
            rotation_iv(ra, R);
            for (int i = 0; i < 9; i++) RA[i] = R[i / 3][i % 3];
            rotation_iv(rb, R);
            for (int i = 0; i < 9; i++) RB[i] = R[i / 3][i % 3];
            for (int k = 0; k < 3; k++) D[k] = pa[k] - pb[k];
            n_la = 0;
            n_lb = 0;
        }
        __syncthreads();
        // A's vertices in B's frame, boxes grown by r; the hulls of A's
        // query boxes (grown by r once more) and of B's vertices
        double q[6] = { INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY };
        double h[6] = { INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY };
        for (int i = tid; i < nva; i += BVH_THREADS) {
            const double* v = a.V + 3 * (va0 + i);
            Iv t[3];
#pragma unroll
            for (int j = 0; j < 3; j++) t[j] = world_coordinate(RA[3 * j], RA[3 * j + 1], RA[3 * j + 2], D[j], v);
#pragma unroll
            for (int m = 0; m < 3; m++) {
                const Iv w = (t[0] * RB[m] + t[1] * RB[3 + m]) + t[2] * RB[6 + m];
                const double lo = (w - a.r).lo, hi = (w + a.r).hi;
                av[6 * i + m] = lo;
                av[6 * i + 3 + m] = hi;
                q[m] = fmin(q[m], lo - a.r);
                q[3 + m] = fmax(q[3 + m], hi + a.r);
            }
        }
        for (int v = vb0 + tid; v < vb1; v += BVH_THREADS)
#pragma unroll
            for (int k = 0; k < 3; k++) {
                h[k] = fmin(h[k], a.V[3 * v + k]);
                h[3 + k] = fmax(h[3 + k], a.V[3 * v + k]);
            }
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1)
#pragma unroll
            for (int k = 0; k < 3; k++) {
                q[k] = fmin(q[k], __shfl_xor(q[k], o));
                q[3 + k] = fmax(q[3 + k], __shfl_xor(q[3 + k], o));
                h[k] = fmin(h[k], __shfl_xor(h[k], o));
                h[3 + k] = fmax(h[3 + k], __shfl_xor(h[3 + k], o));
            }
        if (lane == 0)
            for (int k = 0; k < 6; k++) {
                part[wave][k] = q[k];
                part[wave][6 + k] = h[k];
            }
        __syncthreads();
        if (tid < 12) {
            double x = part[0][tid];
            for (int w = 1; w < BVH_THREADS / 64; w++)
                x = (tid % 6) < 3 ? fmin(x, part[w][tid]) : fmax(x, part[w][tid]);
            if (tid < 6) hull_q[tid] = x;
            else hull_b[tid - 6] = x;
        }
        __syncthreads();
        // A's faces whose query box meets B's hull; B's faces whose box
        // meets the hull of A's query boxes
        for (int k = tid; k < nfa + nfb; k += BVH_THREADS) {
            if (k < nfa) {
                const int* FA = a.F + 3 * (fa0 + k);
                bool hit = true;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const double mn = fmin(fmin(av[6 * (FA[0] - va0) + c], av[6 * (FA[1] - va0) + c]),
                                           av[6 * (FA[2] - va0) + c]) - a.r;
                    const double mx = fmax(fmax(av[6 * (FA[0] - va0) + 3 + c], av[6 * (FA[1] - va0) + 3 + c]),
                                           av[6 * (FA[2] - va0) + 3 + c]) + a.r;
                    hit = hit & (mn <= hull_b[3 + c]) & (hull_b[c] <= mx);
                }
                if (hit) la[atomicAdd(&n_la, 1)] = fa0 + k;
            } else {
                const int* FB = a.F + 3 * (fb0 + k - nfa);
                bool hit = true;
#pragma unroll
                for (int c = 0; c < 3; c++) {
                    const double x0 = a.V[3 * FB[0] + c], x1 = a.V[3 * FB[1] + c], x2 = a.V[3 * FB[2] + c];
                    const double mn = fmin(fmin(x0, x1), x2), mx = fmax(fmax(x0, x1), x2);
                    hit = hit & (hull_q[c] <= mx) & (mn <= hull_q[3 + c]);
                }
                if (hit) lb[atomicAdd(&n_lb, 1)] = fb0 + k - nfa;
            }
        }
        __syncthreads();
        const int nla = n_la, nlb = n_lb;
        for (long long k = tid; k < (long long)nla * nlb; k += BVH_THREADS) {
            const int fa = la[k / nlb], fb = lb[k % nlb];
            const int* FA = a.F + 3 * fa;
            const int* FB = a.F + 3 * fb;
            Box fa_box;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                fa_box.mn[c] = fmin(fmin(av[6 * (FA[0] - va0) + c], av[6 * (FA[1] - va0) + c]), av[6 * (FA[2] - va0) + c]);
                fa_box.mx[c] = fmax(fmax(av[6 * (FA[0] - va0) + 3 + c], av[6 * (FA[1] - va0) + 3 + c]),
                                    av[6 * (FA[2] - va0) + 3 + c]);
            }
            // the BVH's test: A's face box grown by r against B's face box
            bool hit = true;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double x0 = a.V[3 * FB[0] + c], x1 = a.V[3 * FB[1] + c], x2 = a.V[3 * FB[2] + c];
                const double mn = fmin(fmin(x0, x1), x2), mx = fmax(fmax(x0, x1), x2);
                hit = hit & (fa_box.mn[c] - a.r <= mx) & (mn <= fa_box.mx[c] + a.r);
            }
            if (!hit) continue;
            box_finish(fa_box);
            Box fb_box;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                const double x0 = a.V[3 * FB[0] + c], x1 = a.V[3 * FB[1] + c], x2 = a.V[3 * FB[2] + c];
                fb_box.mn[c] = fmin(fmin(x0 - a.r, x1 - a.r), x2 - a.r);
                fb_box.mx[c] = fmax(fmax(x0 + a.r, x1 + a.r), x2 + a.r);
            }
            box_finish(fb_box);
            for (int j = 0; j < 3; j++) {
                // (f_v, f): A's vertex owned by fa against B's face
                const int va = FA[j];
                if (a.v2f[va] == fa) {
                    double mn[3], mx[3];
                    a_vertex_box(av, va - va0, mn, mx);
                    Box vb;
                    for (int c = 0; c < 3; c++) {
                        vb.mn[c] = mn[c];
                        vb.mx[c] = mx[c];
                    }
                    box_finish(vb);
                    if (box_overlap(vb, fb_box)) emit(a.keys_fv, a.cap_fv, a.counts + 1, (unsigned long long)fb * a.nv + va);
                }
                // (f, f_v): B's vertex owned by fb against A's face
                const int vbv = FB[j];
                if (a.v2f[vbv] == fb) {
                    double mn[3], mx[3];
                    b_vertex_box(a.V, vbv, a.r, mn, mx);
                    Box vb;
                    for (int c = 0; c < 3; c++) {
                        vb.mn[c] = mn[c];
                        vb.mx[c] = mx[c];
                    }
                    box_finish(vb);
                    if (box_overlap(fa_box, vb)) emit(a.keys_fv, a.cap_fv, a.counts + 1, (unsigned long long)fa * a.nv + vbv);
                }
            }
            // (f_e, f_e): edges owned by fa and by fb
            for (int ja = 0; ja < 3; ja++) {
                const int ea = a.FE[3 * fa + ja];
                if (a.e2f[ea] != fa) continue;
                double m0[3], x0[3], m1[3], x1[3];
                a_vertex_box(av, a.E[2 * ea] - va0, m0, x0);
                a_vertex_box(av, a.E[2 * ea + 1] - va0, m1, x1);
                const Box ea_box = merge2(m0, x0, m1, x1);
                for (int jb = 0; jb < 3; jb++) {
                    const int eb = a.FE[3 * fb + jb];
                    if (a.e2f[eb] != fb) continue;
                    double n0[3], y0[3], n1[3], y1[3];
                    b_vertex_box(a.V, a.E[2 * eb], a.r, n0, y0);
                    b_vertex_box(a.V, a.E[2 * eb + 1], a.r, n1, y1);
                    const Box eb_box = merge2(n0, y0, n1, y1);
                    if (box_overlap(ea_box, eb_box)) emit(a.keys_ee, a.cap_ee, a.counts, (unsigned long long)ea * a.ne + eb);
                }
            }
        }
        __syncthreads();  // the next pair overwrites the LDS lists
    }
}

}  // namespace

hipError_t run_broad_phase_bvh(BroadPhase& bp, const DevScene& sc, const SceneSel& sel, const double* d_p0,
                               const double* d_p1, double r, bool continuous, bool sorted, int num_cus,
                               hipStream_t st) {
    bp.n_ee = bp.n_fv = bp.n_body_pairs = 0;
    if (sc.nb <= 1 || sel.n_bodies <= 1 || sel.n_tiles == 0) return hipSuccess;
    hipError_t e;
    const int nb = sc.nb;
    if ((e = bp.box_raw.ensure(6 * sizeof(double) * nb)) != hipSuccess) return e;
    if ((e = bp.box_inf.ensure(6 * sizeof(double) * nb)) != hipSuccess) return e;
    if ((e = bp.inc.ensure(sizeof(int) * nb)) != hipSuccess) return e;
    if ((e = bp.counters.ensure(4 * sizeof(unsigned long long))) != hipSuccess) return e;
    if ((e = bp.h_counters.ensure(4 * sizeof(unsigned long long))) != hipSuccess) return e;
    const size_t lds = bvh_lds_bytes(sc);
    if (lds > 65536) return hipErrorInvalidValue;  // a body too large for the LDS staging (checked by the caller)
    const double* p1 = continuous ? d_p1 : d_p0;
    const long long held = (long long)(bp.pairs.bytes / (2 * sizeof(int)));
    long long cap_pairs = std::max<long long>(std::max<long long>(1024, 32LL * sel.n_bodies), held);
    long long cap_ee = std::max<long long>((long long)(bp.keys_ee.bytes / 8), 1 << 16);
    long long cap_fv = std::max<long long>((long long)(bp.keys_fv.bytes / 8), 1 << 16);
    if (bp.min_cap > 0) {
        cap_pairs = std::max<long long>(bp.min_cap, held);
        cap_ee = std::max<long long>(bp.min_cap, (long long)(bp.keys_ee.bytes / 8));
        cap_fv = std::max<long long>(bp.min_cap, (long long)(bp.keys_fv.bytes / 8));
    }
    unsigned long long* cnt = bp.counters.as<unsigned long long>();
    const unsigned long long* got = bp.h_counters.as<unsigned long long>();
    bool fits = false;
    for (int attempt = 0; attempt < 3 && !fits; attempt++) {
        if ((e = bp.pairs.ensure(2 * sizeof(int) * cap_pairs)) != hipSuccess) return e;
        if ((e = bp.keys_ee.ensure(8 * cap_ee)) != hipSuccess) return e;
        if ((e = bp.keys_fv.ensure(8 * cap_fv)) != hipSuccess) return e;
        // body boxes (and the counters cleared)
        if ((e = queue_body_boxes(bp, sc, sel, d_p0, p1, r, st)) != hipSuccess) return e;
        hipLaunchKernelGGL(bvh_body_pair_kernel, dim3((unsigned)sel.n_tiles), dim3(BVH_THREADS), 0, st, sel.tiles,
                           sc.group, bp.box_inf.as<double>(), bp.pairs.as<int>(), cap_pairs, cnt + 2);
        BvhArgs a{};
        a.V = sc.V; a.E = sc.E; a.F = sc.F; a.FE = sc.FE;
        a.bvs = sc.body_vstart; a.bfs = sc.body_fstart; a.v2f = sc.v2f; a.e2f = sc.e2f;
        a.p0 = d_p0; a.p1 = p1; a.r = r; a.continuous = continuous ? 1 : 0;
        a.pairs = bp.pairs.as<int>(); a.n_pairs = cnt + 2; a.cap_pairs = cap_pairs;
        a.keys_ee = bp.keys_ee.as<unsigned long long>(); a.cap_ee = cap_ee;
        a.keys_fv = bp.keys_fv.as<unsigned long long>(); a.cap_fv = cap_fv;
        a.counts = cnt; a.ne = sc.ne; a.nv = sc.nv; a.max_faces = sc.max_body_faces;
        const unsigned grid = (unsigned)std::min<long long>(cap_pairs, (long long)num_cus * 8);
        hipLaunchKernelGGL(bvh_pair_kernel, dim3(grid), dim3(BVH_THREADS), lds, st, a);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(bp.h_counters.p, cnt, 3 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st)) !=
            hipSuccess)
            return e;
        if ((e = hipStreamSynchronize(st)) != hipSuccess) return e;
        fits = (long long)got[2] <= cap_pairs && (long long)got[0] <= cap_ee && (long long)got[1] <= cap_fv;
        cap_pairs = std::max<long long>(cap_pairs, (long long)got[2]);
        cap_ee = std::max<long long>(cap_ee, (long long)got[0]);
        cap_fv = std::max<long long>(cap_fv, (long long)got[1]);
    }
    if (!fits) return hipErrorUnknown;
    bp.n_body_pairs = (long long)got[2];
    bp.n_ee = (long long)got[0];
    bp.n_fv = (long long)got[1];
    return finish_broad_phase(bp, sc, sorted, st);
}

}  // namespace rigc
