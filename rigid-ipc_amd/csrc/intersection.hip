// Edge-triangle intersection check between close bodies, for gfx950.
//
// Restates the 3D branch of RigidBodyProblem::detect_intersections
// (src/physics/rigid_body_problem.cpp:289-360): close body pairs at the
// poses without inflation (close_bodies, src/physics/rigid_body_assembler.cpp:268-356),
// edge-face candidates of each pair (detect_intersection_candidates_rigid_bvh,
// src/ccd/rigid/broad_phase.cpp:265-305, and
// detect_body_pair_intersection_candidates_from_aabbs,
// src/ccd/rigid/rigid_body_bvh.cpp:315-444), and is_edge_intersecting_triangle on
// the world vertices of each candidate. The reference stops at the first
// intersecting pair; here every intersecting (edge, face) pair is found,
// so the answer (the sorted list, its length) does not depend on the
// candidate set as long as that set holds every pair whose boxes overlap:
// an intersecting edge and triangle always have overlapping boxes, and the
// boxes are taken from the same world vertices as the predicate.
//
// One workgroup per close body pair (grid-stride over the device-side pair
// count): the hulls of both bodies' world vertices, then for each direction
// (edges of one body against faces of the other) the faces whose box meets
// the edge body's hull are staged in LDS, up to 256 per pass, and every
// edge whose box meets the face body's hull tests them: box overlap, then
// the edge-triangle rule with floating-point-filtered orientations. Pairs
// the filter cannot decide (exact zeros, near-degenerate configurations)
// are queued and settled by a second kernel with exact orientations.
#include <algorithm>

#include "rigc_exact.hpp"
#include "rigc_internal.hpp"

namespace rigc {
namespace {

#define IX_CHECK(x)                      \
    do {                                 \
        hipError_t e_ = (x);             \
        if (e_ != hipSuccess) return e_; \
    } while (0)

constexpr int IX_THREADS = 256;
constexpr int IX_WAVES = IX_THREADS / 64;
constexpr int IX_TILE = IX_THREADS;  // faces staged per pass, one per thread
constexpr int EXACT_THREADS = 64;
constexpr int EXACT_BLOCKS = 64;

struct IsectArgs {
    const int* pairs;                   // [n][2] close body pairs
    const unsigned long long* n_pairs;  // device count (may exceed cap_pairs)
    long long cap_pairs;
    const int* bvs;
    const int* bes;
    const int* bfs;
    const int* E;
    const int* F;
    const double* W;                    // [nv][3] world vertices
    unsigned long long nf;
    unsigned long long* keys;           // intersecting pairs, edge * nf + face
    long long cap_keys;
    unsigned long long* unsure;         // pairs left open by the filter
    long long cap_unsure;
    unsigned long long* counts;         // [0] intersecting, [1] open (true counts, also past a capacity)
};

__device__ __forceinline__ void append(unsigned long long key, unsigned long long* buf, long long cap,
                                       unsigned long long* count) {
    const unsigned long long k = atomicAdd(count, 1ULL);
    if ((long long)k < cap) buf[k] = key;
}

// Inclusive overlap of [min xyz, max xyz] boxes (bitwise &: all bounds read together).
__device__ __forceinline__ bool boxes_meet(const double* a, const double* b) {
    return (a[0] <= b[3]) & (b[0] <= a[3]) & (a[1] <= b[4]) & (b[1] <= a[4]) & (a[2] <= b[5]) & (b[2] <= a[5]);
}

__global__ __launch_bounds__(IX_THREADS) void edge_face_kernel(IsectArgs a) {
    __shared__ double tri[IX_TILE][9];
    __shared__ double tbox[IX_TILE][6];
    __shared__ int tface[IX_TILE];
    __shared__ double wred[IX_WAVES][6];
    __shared__ double hull[2][6];
    __shared__ int n_list;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const long long np = (long long)min(*a.n_pairs, (unsigned long long)a.cap_pairs);
    for (long long pi = blockIdx.x; pi < np; pi += gridDim.x) {
        const int body0 = a.pairs[2 * pi], body1 = a.pairs[2 * pi + 1];
#pragma unroll
        for (int s = 0; s < 2; s++) {
            const int b = s ? body1 : body0;
            double v[6] = { INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY };
            for (int i = a.bvs[b] + tid; i < a.bvs[b + 1]; i += IX_THREADS)
                for (int k = 0; k < 3; k++) {
                    const double w = a.W[3 * i + k];
                    v[k] = fmin(v[k], w);
                    v[3 + k] = fmax(v[3 + k], w);
                }
            for (int k = 0; k < 6; k++)
                for (int o = 32; o > 0; o >>= 1) {
                    const double y = __shfl_xor(v[k], o);
                    v[k] = k < 3 ? fmin(v[k], y) : fmax(v[k], y);
                }
            if (lane == 0)
                for (int k = 0; k < 6; k++) wred[wave][k] = v[k];
            __syncthreads();
            if (tid < 6) {
                double h = wred[0][tid];
                for (int w = 1; w < IX_WAVES; w++) h = tid < 3 ? fmin(h, wred[w][tid]) : fmax(h, wred[w][tid]);
                hull[s][tid] = h;
            }
            __syncthreads();
        }
        for (int dir = 0; dir < 2; dir++) {
            const int X = dir ? body1 : body0, Y = dir ? body0 : body1;  // edges of X, faces of Y
            const double* hx = hull[dir];
            const double* hy = hull[1 - dir];
            const int e0 = a.bes[X], e1 = a.bes[X + 1], f0 = a.bfs[Y], f1 = a.bfs[Y + 1];
            if (e0 == e1) continue;
            for (int fb = f0; fb < f1; fb += IX_TILE) {
                if (tid == 0) n_list = 0;
                __syncthreads();
                const int f = fb + tid;
                if (f < f1) {
                    double t[9], bx[6];
                    for (int c = 0; c < 3; c++) {
                        const int vi = a.F[3 * f + c];
                        for (int k = 0; k < 3; k++) t[3 * c + k] = a.W[3 * vi + k];
                    }
                    for (int k = 0; k < 3; k++) {
                        bx[k] = fmin(fmin(t[k], t[3 + k]), t[6 + k]);
                        bx[3 + k] = fmax(fmax(t[k], t[3 + k]), t[6 + k]);
                    }
                    if (boxes_meet(bx, hx)) {
                        const int slot = atomicAdd(&n_list, 1);
                        for (int k = 0; k < 9; k++) tri[slot][k] = t[k];
                        for (int k = 0; k < 6; k++) tbox[slot][k] = bx[k];
                        tface[slot] = f;
                    }
                }
                __syncthreads();
                const int n = n_list;
                if (n > 0) {
                    for (int e = e0 + tid; e < e1; e += IX_THREADS) {
                        const int i0 = a.E[2 * e], i1 = a.E[2 * e + 1];
                        double p[3], q[3], eb[6];
                        for (int k = 0; k < 3; k++) {
                            p[k] = a.W[3 * i0 + k];
                            q[k] = a.W[3 * i1 + k];
                            eb[k] = fmin(p[k], q[k]);
                            eb[3 + k] = fmax(p[k], q[k]);
                        }
                        if (!boxes_meet(eb, hy)) continue;
                        for (int j = 0; j < n; j++) {
                            if (!boxes_meet(eb, tbox[j])) continue;
                            double t[9];
                            for (int k = 0; k < 9; k++) t[k] = tri[j][k];
                            const int r = exact::edge_triangle_filtered(p, q, t, t + 3, t + 6);
                            if (r == 0) continue;
                            const unsigned long long key = (unsigned long long)e * a.nf + (unsigned long long)tface[j];
                            if (r > 0) append(key, a.keys, a.cap_keys, a.counts);
                            else append(key, a.unsure, a.cap_unsure, a.counts + 1);
                        }
                    }
                }
                __syncthreads();  // the list is read before the next pass rewrites it
            }
        }
        __syncthreads();  // the hulls are read before the next pair rewrites them
    }
}

// Pairs the filter left open, settled with exact orientations.
__global__ __launch_bounds__(EXACT_THREADS) void edge_face_exact_kernel(IsectArgs a) {
    const long long n = (long long)min(a.counts[1], (unsigned long long)a.cap_unsure);
    for (long long i = (long long)blockIdx.x * EXACT_THREADS + threadIdx.x; i < n;
         i += (long long)gridDim.x * EXACT_THREADS) {
        const unsigned long long key = a.unsure[i];
        const int e = (int)(key / a.nf), f = (int)(key % a.nf);
        double p[3], q[3], t[9];
        for (int k = 0; k < 3; k++) {
            p[k] = a.W[3 * a.E[2 * e] + k];
            q[k] = a.W[3 * a.E[2 * e + 1] + k];
            for (int c = 0; c < 3; c++) t[3 * c + k] = a.W[3 * a.F[3 * f + c] + k];
        }
        if (exact::edge_triangle(p, q, t, t + 3, t + 6)) append(key, a.keys, a.cap_keys, a.counts);
    }
}

__global__ void unpack_edge_face_kernel(const unsigned long long* keys, long long n, unsigned long long nf, int* out) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        out[2 * i] = (int)(keys[i] / nf);
        out[2 * i + 1] = (int)(keys[i] % nf);
    }
}

int key_bits(unsigned long long maxval) {
    int b = 0;
    while (b < 64 && (maxval >> b) != 0ULL) b++;
    return b;
}

}  // namespace

hipError_t detect_intersections(Assembly& as, BroadPhase& bp, Intersections& ix, const DevScene& sc,
                                const SceneSel& sel, const double* d_poses, hipStream_t st) {
    ix.n = ix.n_unsure = 0;
    if (sc.nb <= 1 || sel.n_bodies <= 1 || sel.n_tiles == 0 || sc.ne == 0 || sc.nf == 0) return hipSuccess;
    IX_CHECK(current_world_vertices(as, sc, sel, d_poses, st));
    IX_CHECK(queue_body_pairs(bp, sc, sel, d_poses, 0.0, st));
    IX_CHECK(ix.counts.ensure(2 * sizeof(unsigned long long)));
    IX_CHECK(ix.h_counts.ensure(4 * sizeof(unsigned long long)));
    if (bp.min_cap > 0) {
        ix.cap_keys = ix.cap_unsure = bp.min_cap;
    } else {
        ix.cap_keys = std::max<long long>(4096, (long long)(ix.keys.bytes / 8));
        ix.cap_unsure = std::max<long long>(4096, (long long)(ix.unsure.bytes / 8));
    }
    // Counts come back with one read-back per pass ([0] body pairs, [1]
    // intersecting, [2] open). A pass that overflowed a capacity is run
    // again at the counts it saw: body pairs first, then the open pairs
    // (the exact kernel saw only part of them), then the results.
    const unsigned long long* got = ix.h_counts.as<unsigned long long>();
    bool fits = false, pairs_short = false;
    for (int attempt = 0; attempt < 4 && !fits; attempt++) {
        if (pairs_short) IX_CHECK(requeue_body_pairs(bp, sc, sel, st));
        IX_CHECK(ix.keys.ensure(8 * ix.cap_keys));
        IX_CHECK(ix.unsure.ensure(8 * ix.cap_unsure));
        IX_CHECK(hipMemsetAsync(ix.counts.p, 0, 2 * sizeof(unsigned long long), st));
        IsectArgs x;
        x.pairs = bp.pairs.as<int>();
        x.n_pairs = bp.counters.as<unsigned long long>() + 2;
        x.cap_pairs = bp.cap_pairs;
        x.bvs = sc.body_vstart;
        x.bes = sc.body_estart;
        x.bfs = sc.body_fstart;
        x.E = sc.E;
        x.F = sc.F;
        x.W = as.W.as<double>();
        x.nf = (unsigned long long)sc.nf;
        x.keys = ix.keys.as<unsigned long long>();
        x.cap_keys = ix.cap_keys;
        x.unsure = ix.unsure.as<unsigned long long>();
        x.cap_unsure = ix.cap_unsure;
        x.counts = ix.counts.as<unsigned long long>();
        const unsigned grid = (unsigned)std::min<long long>(std::max<long long>(bp.cap_pairs, 1), 2048);
        hipLaunchKernelGGL(edge_face_kernel, dim3(grid), dim3(IX_THREADS), 0, st, x);
        hipLaunchKernelGGL(edge_face_exact_kernel, dim3(EXACT_BLOCKS), dim3(EXACT_THREADS), 0, st, x);
        IX_CHECK(hipGetLastError());
        unsigned long long* h = ix.h_counts.as<unsigned long long>();
        IX_CHECK(hipMemcpyAsync(h, bp.counters.as<unsigned long long>() + 2, sizeof(unsigned long long),
                                hipMemcpyDeviceToHost, st));
        IX_CHECK(hipMemcpyAsync(h + 1, ix.counts.p, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        IX_CHECK(hipStreamSynchronize(st));
        pairs_short = (long long)got[0] > bp.cap_pairs;
        fits = !pairs_short && (long long)got[1] <= ix.cap_keys && (long long)got[2] <= ix.cap_unsure;
        bp.cap_pairs = std::max<long long>(bp.cap_pairs, (long long)got[0]);
        ix.cap_keys = std::max<long long>(ix.cap_keys, (long long)got[1]);
        ix.cap_unsure = std::max<long long>(ix.cap_unsure, (long long)got[2]);
    }
    if (!fits) return hipErrorUnknown;
    ix.n = (long long)got[1];
    ix.n_unsure = (long long)got[2];
    IX_CHECK(ix.out.ensure(2 * sizeof(int) * std::max<long long>(ix.n, 1)));
    if (ix.n == 0) return hipSuccess;
    if (ix.n > 1) {
        IX_CHECK(ix.tmp.ensure(8 * ix.n));
        IX_CHECK(ix.hist.ensure(sizeof(unsigned int) * radix_sort_hist_entries(ix.n)));
        IX_CHECK(radix_sort_u64(ix.keys.as<unsigned long long>(), ix.tmp.as<unsigned long long>(),
                                ix.hist.as<unsigned int>(), ix.n,
                                key_bits((unsigned long long)sc.ne * (unsigned long long)sc.nf), st));
    }
    hipLaunchKernelGGL(unpack_edge_face_kernel, dim3((unsigned)((ix.n + 255) / 256)), dim3(256), 0, st,
                       ix.keys.as<unsigned long long>(), ix.n, (unsigned long long)sc.nf, ix.out.as<int>());
    return hipGetLastError();
}

}  // namespace rigc
