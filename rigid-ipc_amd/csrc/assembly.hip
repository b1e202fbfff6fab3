// Constraint set, barrier and lagged-friction assembly on the device.
//
//   construct_constraint_set   src/opt/distance_barrier_constraint.cpp:267-306
//                              (-> ipc::construct_constraint_set)
//   compute_barrier_term       src/problems/distance_barrier_rb_problem.cpp:1299-1399
//   apply_chain_rule           :1182-1248 (world_vertices_diff, project_to_psd)
//   local_*_to_global          :1143-1179, merge_derivative_storage :1260-1297
//   update_friction_constraints :168-188, compute_friction_term :1410-1572
//   compute_minimum_distance   src/opt/distance_barrier_constraint.cpp:308-323
//
// Pipeline (one HIP stream):
//   body_rotation_kernel   per body: R(r)
//   body_frame_kernel      per body: dR/dr, d2R/dr2 (closed form), nine threads each
//   world_vertex_kernel    per vertex: R x + p (round-to-nearest, as the host)
//   static broad phase     (broadphase.hip) with inflation (dhat + dmin) / 2,
//                          queued without a read-back
//   classify_kernel        per packed candidate key (counts read on the
//                          device): closest-feature type, activity
//                          d < (dhat + dmin)^2, edge-edge mollifier test;
//                          appends a sort key to the VV / EV / EE / FV list;
//                          one read-back returns candidate and class counts
//   radix sort + run-length encoding: VV / EV duplicates -> multiplicity
//   record_kernel          ConRec list, order VV, EV, EE, FV
//   local_eg_kernel        thread per constraint: energy (+ 12-gradient)
//   local_hess_kernel      one wavefront per constraint: vertex-space
//                          Hessian in LDS, chain rule to the two bodies'
//                          12 dofs, parallel-order Jacobi PSD projection
//   body / pair reduction  sorted (body, constraint) keys -> per-body
//                          gradient and diagonal 6x6 blocks, sorted body
//                          pairs -> off-diagonal 6x6 blocks; fixed order,
//                          so results are deterministic.
#include <algorithm>
#include <cstring>
#include <vector>

#include "rigc_barrier.hpp"
#include "rigc_internal.hpp"

namespace rigc {
namespace {

#define AS_CHECK(x)                           \
    do {                                      \
        hipError_t e_ = (x);                  \
        if (e_ != hipSuccess) return e_;      \
    } while (0)

constexpr int THREADS = 256;

int bits_for(unsigned long long maxval) {
    int b = 0;
    while (b < 64 && (maxval >> b) != 0ULL) b++;
    return b;
}

// --------------------------------------------------------------- frames
// Rotations of all bodies (one thread each) and the frames: nine threads
// per body, each computing the shared terms and the entries q, q + 9, ...
// of the 81 (q = blockIdx.y, the same for a whole block, so each entry is
// taken with constant indices and no branch diverges).
__global__ __launch_bounds__(THREADS) void body_rotation_kernel(int nb, const double* poses, double* R) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b < nb) rotation_d(poses + 6 * b + 3, R + 9 * b);
}

template <int Q>
__device__ __forceinline__ void frame_entries(const FrameTerms& T, double* out) {
#pragma unroll
    for (int k = 0; k < 9; k++) out[Q + 9 * k] = frame_entry(T, Q + 9 * k);
}

__global__ __launch_bounds__(THREADS) void body_frame_kernel(int nb, const double* poses, BodyFrame* F) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    FrameTerms T;
    frame_terms(poses + 6 * b + 3, T);
    double* out = reinterpret_cast<double*>(F + b);  // dR[3][9] then d2R[6][9]
    switch (blockIdx.y) {
        case 0: frame_entries<0>(T, out); break;
        case 1: frame_entries<1>(T, out); break;
        case 2: frame_entries<2>(T, out); break;
        case 3: frame_entries<3>(T, out); break;
        case 4: frame_entries<4>(T, out); break;
        case 5: frame_entries<5>(T, out); break;
        case 6: frame_entries<6>(T, out); break;
        case 7: frame_entries<7>(T, out); break;
        default: frame_entries<8>(T, out); break;
    }
}

__global__ __launch_bounds__(THREADS) void world_vertex_kernel(int nv, const double* V, const int* v_body,
                                                              const double* poses, const double* R, double* W) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    const int b = v_body[i];
    world_vertex_d(R + 9 * b, poses + 6 * b, V + 3 * i, W + 3 * i);
}

// Frames and world vertices of listed bodies only (a block per body).
__global__ __launch_bounds__(THREADS) void body_list_vertex_kernel(const int* bodies, const int* bvs, const double* V,
                                                                  const double* poses, double* R, BodyFrame* F,
                                                                  double* W) {
    __shared__ double Rs[9];
    const int b = bodies[blockIdx.x];
    if (threadIdx.x == 0) {
        rotation_d(poses + 6 * b + 3, Rs);
        for (int k = 0; k < 9; k++) R[9 * b + k] = Rs[k];
    }
    if (F && threadIdx.x == 64) body_frame(poses + 6 * b + 3, F[b]);
    __syncthreads();
    for (int i = bvs[b] + threadIdx.x; i < bvs[b + 1]; i += THREADS) world_vertex_d(Rs, poses + 6 * b, V + 3 * i, W + 3 * i);
}

// --------------------------------------------------------------- constraint set
struct ClassifyArgs {
    // the broad phase's packed keys (e0 * ne + e1, f * nv + v) and their
    // counts on the device; the grid covers both capacities
    const unsigned long long* kee;
    const unsigned long long* kfv;
    const unsigned long long* n_pairs;  // [0] edge-edge, [1] face-vertex
    long long cap_ee, cap_fv;
    unsigned long long ne;
    const int* E;
    const int* F;
    const int* FE;
    const double* W;
    int nv;
    double offset_sqr;
    unsigned long long* keys[4];
    unsigned int* count;
};

__device__ inline void load3(const double* W, int v, double* x) {
    x[0] = W[3 * v];
    x[1] = W[3 * v + 1];
    x[2] = W[3 * v + 2];
}

// ipc::construct_constraint_set per candidate: the closest-feature pair
// decides the constraint kind; a nearly parallel edge pair stays an
// edge-edge constraint (mollified).
// Classes are keyed by elements: VV (v0, v1), EV (e, v), EE and FV by
// the candidate's own packed key, so each class sorts into the order of
// the sorted candidate list.
__global__ __launch_bounds__(THREADS) void classify_kernel(ClassifyArgs a) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.cap_ee + a.cap_fv) return;
    double X[12];
    int cls;
    unsigned long long key;
    const unsigned long long NV = (unsigned long long)a.nv;
    if (i < a.cap_ee) {
        if ((unsigned long long)i >= min(a.n_pairs[0], (unsigned long long)a.cap_ee)) return;
        const unsigned long long kp = a.kee[i];
        const int ei = (int)(kp / a.ne), ej = (int)(kp % a.ne);
        const int i0 = a.E[2 * ei], i1 = a.E[2 * ei + 1], j0 = a.E[2 * ej], j1 = a.E[2 * ej + 1];
        load3(a.W, i0, X);
        load3(a.W, i1, X + 3);
        load3(a.W, j0, X + 6);
        load3(a.W, j1, X + 9);
        int t = edge_edge_type(X, X + 3, X + 6, X + 9);
        DistForm f;
        constraint_form(CON_EE, X, f);
        if (!(f.d < a.offset_sqr)) return;
        double ua[3], ub[3], n[3];
        v3_sub(X + 3, X, ua);
        v3_sub(X + 9, X + 6, ub);
        v3_cross(ua, ub, n);
        const double eps_x = 1e-3 * v3_dot(ua, ua) * v3_dot(ub, ub);
        if (v3_dot(n, n) < eps_x) t = EE_AB;
        auto vv = [&](int p, int q) {
            cls = CON_VV;
            key = (unsigned long long)min(p, q) * NV + (unsigned long long)max(p, q);
        };
        auto ev = [&](int e, int v) {
            cls = CON_EV;
            key = (unsigned long long)e * NV + (unsigned long long)v;
        };
        switch (t) {
        case EE_A0B0: vv(i0, j0); break;
        case EE_A0B1: vv(i0, j1); break;
        case EE_A1B0: vv(i1, j0); break;
        case EE_A1B1: vv(i1, j1); break;
        case EE_AB0: ev(ei, j0); break;
        case EE_AB1: ev(ei, j1); break;
        case EE_A0B: ev(ej, i0); break;
        case EE_A1B: ev(ej, i1); break;
        default: cls = CON_EE; key = kp; break;
        }
    } else {
        const long long j = i - a.cap_ee;
        if ((unsigned long long)j >= min(a.n_pairs[1], (unsigned long long)a.cap_fv)) return;
        const unsigned long long kp = a.kfv[j];
        const int fc = (int)(kp / NV), v = (int)(kp % NV);
        const int f0 = a.F[3 * fc], f1 = a.F[3 * fc + 1], f2 = a.F[3 * fc + 2];
        load3(a.W, v, X);
        load3(a.W, f0, X + 3);
        load3(a.W, f1, X + 6);
        load3(a.W, f2, X + 9);
        DistForm f;
        constraint_form(CON_FV, X, f);
        if (!(f.d < a.offset_sqr)) return;
        auto vv = [&](int p, int q) {
            cls = CON_VV;
            key = (unsigned long long)min(p, q) * NV + (unsigned long long)max(p, q);
        };
        auto ev = [&](int e, int w) {
            cls = CON_EV;
            key = (unsigned long long)e * NV + (unsigned long long)w;
        };
        switch (point_triangle_type(X, X + 3, X + 6, X + 9)) {
        case PT_T0: vv(v, f0); break;
        case PT_T1: vv(v, f1); break;
        case PT_T2: vv(v, f2); break;
        case PT_E0: ev(a.FE[3 * fc], v); break;
        case PT_E1: ev(a.FE[3 * fc + 1], v); break;
        case PT_E2: ev(a.FE[3 * fc + 2], v); break;
        default: cls = CON_FV; key = kp; break;
        }
    }
    const unsigned int slot = atomicAdd(&a.count[cls], 1u);
    a.keys[cls][slot] = key;
}

__global__ __launch_bounds__(THREADS) void head_flags_kernel(const unsigned long long* keys, long long n,
                                                            unsigned int* flags) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// pos = exclusive scan of the head flags: start[pos[i]] = i at heads.
__global__ __launch_bounds__(THREADS) void head_starts_kernel(const unsigned int* flags, const unsigned int* pos,
                                                             long long n, unsigned int* start) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (flags[i]) start[pos[i]] = (unsigned int)i;
}

struct RecordArgs {
    int cls;
    const unsigned long long* keys;  // sorted keys of this class
    const unsigned int* start;       // VV / EV: run starts (nullptr: one record per key)
    long long n_keys;                // keys of this class
    long long n_out;                 // records to write
    ConRec* out;
    unsigned long long ne;
    const int* E;
    const int* F;
    const int* v_body;
    const double* W;
    int nv;
};

__global__ __launch_bounds__(THREADS) void record_kernel(RecordArgs a) {
    const long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= a.n_out) return;
    ConRec r;
    r.eps_x = -1.0;
    r.v[0] = r.v[1] = r.v[2] = r.v[3] = -1;
    long long k = u;
    int mult = 1;
    if (a.start) {
        k = a.start[u];
        const long long end = (u + 1 < a.n_out) ? (long long)a.start[u + 1] : a.n_keys;
        mult = (int)(end - k);
    }
    const unsigned long long key = a.keys[k];
    const unsigned long long NV = (unsigned long long)a.nv;
    int nvx = 4;
    switch (a.cls) {
    case CON_VV:
        r.a = (int)(key / NV);
        r.b = (int)(key % NV);
        r.v[0] = r.a;
        r.v[1] = r.b;
        nvx = 2;
        break;
    case CON_EV:
        r.a = (int)(key / NV);
        r.b = (int)(key % NV);
        r.v[0] = r.b;
        r.v[1] = a.E[2 * r.a];
        r.v[2] = a.E[2 * r.a + 1];
        nvx = 3;
        break;
    case CON_EE: {
        r.a = (int)(key / a.ne);
        r.b = (int)(key % a.ne);
        r.v[0] = a.E[2 * r.a];
        r.v[1] = a.E[2 * r.a + 1];
        r.v[2] = a.E[2 * r.b];
        r.v[3] = a.E[2 * r.b + 1];
        // eps_x from V_rest = V (distance_barrier_constraint.cpp:297-299)
        double X[12], ua[3], ub[3];
        for (int s = 0; s < 4; s++) load3(a.W, r.v[s], X + 3 * s);
        v3_sub(X + 3, X, ua);
        v3_sub(X + 9, X + 6, ub);
        r.eps_x = 1e-3 * v3_dot(ua, ua) * v3_dot(ub, ub);
        break;
    }
    default:
        r.a = (int)(key / NV);
        r.b = (int)(key % NV);
        r.v[0] = r.b;
        r.v[1] = a.F[3 * r.a];
        r.v[2] = a.F[3 * r.a + 1];
        r.v[3] = a.F[3 * r.a + 2];
        break;
    }
    r.type = a.cls;
    r.mult = mult;
    r.body[0] = a.v_body[r.v[0]];
    r.body[1] = a.v_body[r.v[nvx - 1]];
    a.out[u] = r;
}

// --------------------------------------------------------------- local terms
__device__ inline int con_nv(int type) { return type == CON_VV ? 2 : (type == CON_EV ? 3 : 4); }
// RigidBody*Constraint::vertex_local_body_ids
// (src/problems/rigid_body_collision_constraint.hpp:12-240)
__device__ inline int slot_side(int type, int s) { return type == CON_EE ? (s >> 1) : (s == 0 ? 0 : 1); }

struct LocalArgs {
    const ConRec* cons;
    const FricRec* fric;
    long long n;
    const double* V;
    const int* v_body;
    const double* W;   // world vertices at the evaluation poses
    const double* W0;  // friction: at the start of the step
    const BodyFrame* frames;
    double dhat, dmin, epsv_h;
    const double* epsv_h_scene;  // friction: per-scene epsv_h (nullptr: epsv_h)
    const int* body_scene;
    int want_grad;
    int project;       // project_to_psd the 12 x 12 blocks (0: exact Hessian)
    double* energy;    // [n]
    double* lgrad;     // [n][12]
    double* lhess;     // [n][144]
};

template <bool FRIC>
__device__ inline void local_setup(const LocalArgs& a, long long c, ConRec& rec, BarrierLocal& B,
                                   FrictionLocal& Fl) {
    double X[12];
    if (FRIC) {
        const FricRec& f = a.fric[c];
        rec = f.c;
        const int nv = con_nv(rec.type);
        for (int s = 0; s < 4; s++)
            for (int k = 0; k < 3; k++)
                X[3 * s + k] = s < nv ? a.W[3 * rec.v[s] + k] - a.W0[3 * rec.v[s] + k] : 0.0;
        double basis[3][2];
        for (int k = 0; k < 3; k++) {
            basis[k][0] = f.basis[2 * k];
            basis[k][1] = f.basis[2 * k + 1];
        }
        const double eh = a.epsv_h_scene ? a.epsv_h_scene[a.body_scene[rec.body[0]]] : a.epsv_h;
        friction_local_setup(Fl, rec.type, nv, f.coords, basis, f.normal_force, f.mu, rec.mult, X, eh);
    } else {
        rec = a.cons[c];
        const int nv = con_nv(rec.type);
        for (int s = 0; s < 4; s++)
            for (int k = 0; k < 3; k++) X[3 * s + k] = s < nv ? a.W[3 * rec.v[s] + k] : 0.0;
        barrier_local_setup(B, rec.type, X, rec.eps_x, rec.mult, a.dhat, a.dmin);
    }
}

// Vertex Jacobian column c of slot vertex: (dR/dr_c) x, component k.
__device__ inline double jac_rot(const BodyFrame& F, const double* x, int c, int k) {
    return (F.dR[c][3 * k] * x[0] + F.dR[c][3 * k + 1] * x[1]) + F.dR[c][3 * k + 2] * x[2];
}

// Thread per constraint: energy and (optionally) the 12-gradient in the
// two bodies' dofs.
template <bool FRIC>
__global__ __launch_bounds__(THREADS) void local_eg_kernel(LocalArgs a) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.n) return;
    ConRec rec;
    BarrierLocal B;
    FrictionLocal Fl;
    local_setup<FRIC>(a, c, rec, B, Fl);
    a.energy[c] = FRIC ? Fl.energy : barrier_local_energy(B);
    if (!a.want_grad) return;
    const int nv = con_nv(rec.type);
    double lg[12];
    for (int r = 0; r < 12; r++) lg[r] = 0.0;
    for (int s = 0; s < nv; s++) {
        double g[3];
        for (int k = 0; k < 3; k++) g[k] = FRIC ? friction_local_grad(Fl, 3 * s + k) : barrier_local_grad(B, s, k);
        const int side = slot_side(rec.type, s);
        const int v = rec.v[s];
        const BodyFrame& F = a.frames[a.v_body[v]];
        const double* x = a.V + 3 * v;
        double t[6];
        for (int d = 0; d < 3; d++) t[d] = g[d];
        for (int d = 0; d < 3; d++) {
            double sum = 0.0;
            for (int k = 0; k < 3; k++) sum += jac_rot(F, x, d, k) * g[k];
            t[3 + d] = sum;
        }
        for (int d = 0; d < 6; d++) lg[6 * side + d] += t[d];
    }
    for (int r = 0; r < 12; r++) a.lgrad[12 * c + r] = lg[r];
}

// Pair k (0..5) of round r (0..10) of the parallel Jacobi ordering: the
// circle method with index 0 fixed and 1..11 turning one place per round,
// so the 11 rounds of 6 disjoint pairs cover the 66 pairs of 12 indices
// once. Round 0 is (0,11) (1,10) (2,9) (3,8) (4,7) (5,6).
__device__ inline void jacobi_pair(int r, int k, int& p, int& q) {
    if (k == 0) {
        p = 0;
        q = 1 + (10 - r);
    } else {
        int a = k - 1 - r, b = 10 - k - r;
        if (a < 0) a += 11;
        if (b < 0) b += 11;
        p = 1 + a;
        q = 1 + b;
    }
}

__device__ inline double wave_sum_uniform(double x) {
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o);
    return __shfl(x, 0);  // lane 0's value, so every lane takes the same branch
}

// Jacobi rotation (c, s) that zeroes A[p][q] of the symmetric 12 x 12 A.
__device__ inline void jacobi_rotation(const double* A, int p, int q, double& c, double& s) {
    const double apq = A[12 * p + q];
    c = 1.0;
    s = 0.0;
    if (apq != 0.0) {
        const double theta = (A[13 * q] - A[13 * p]) / (2 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1));
        c = 1 / sqrt(t * t + 1);
        s = t * c;
    }
}

// One wavefront (the whole block) per constraint.
template <bool FRIC>
__global__ __launch_bounds__(64) void local_hess_kernel(LocalArgs a) {
    __shared__ double sHv[144];  // vertex-space Hessian
    __shared__ double sHJ[144];  // Hv * J
    __shared__ double sB[144];   // body-space Hessian before projection
    __shared__ double sA[144];   // Jacobi iterate
    __shared__ double sV[144];   // eigenvectors
    __shared__ double sg[12];
    __shared__ double sG[4][9];  // slot Jacobian (rotation columns): [k][c]
    __shared__ double sT[4][6];  // slot rotational Hessian term, (c,d) packed
    __shared__ double sCS[6][2];  // the round's rotations (c, s)
    const int lane = threadIdx.x;
    for (long long c = blockIdx.x; c < a.n; c += gridDim.x) {
        ConRec rec;
        BarrierLocal B;
        FrictionLocal Fl;
        local_setup<FRIC>(a, c, rec, B, Fl);
        const int nv = con_nv(rec.type);
        if (lane == 0) a.energy[c] = FRIC ? Fl.energy : barrier_local_energy(B);
        if (lane < 12) sg[lane] = FRIC ? friction_local_grad(Fl, lane) : barrier_local_grad(B, lane / 3, lane % 3);
        for (int e = lane; e < 144; e += 64) {
            const int r = e / 12, q = e % 12;
            sHv[e] = FRIC ? friction_local_hess(Fl, r, q) : barrier_local_hess(B, r / 3, r % 3, q / 3, q % 3);
        }
        if (lane < 36) {
            const int s = lane / 9, kc = lane % 9;
            double val = 0.0;
            if (s < nv) {
                const int v = rec.v[s];
                val = jac_rot(a.frames[a.v_body[v]], a.V + 3 * v, kc % 3, kc / 3);
            }
            sG[s][kc] = val;  // [k][c] with k = kc / 3, c = kc % 3
        }
        __syncthreads();
        if (lane < 24) {
            const int s = lane / 6, cd = lane % 6;
            double val = 0.0;
            if (s < nv) {
                const int v = rec.v[s];
                const BodyFrame& F = a.frames[a.v_body[v]];
                const double* x = a.V + 3 * v;
                for (int k = 0; k < 3; k++) {
                    const double h = (F.d2R[cd][3 * k] * x[0] + F.d2R[cd][3 * k + 1] * x[1]) + F.d2R[cd][3 * k + 2] * x[2];
                    val += h * sg[3 * s + k];
                }
            }
            sT[s][cd] = val;
        }
        // HJ = Hv J (vertex rows x body columns)
        for (int e = lane; e < 144; e += 64) {
            const int r = e / 12, col = e % 12, side = col / 6, q = col % 6;
            double sum = 0.0;
            for (int s = 0; s < nv; s++) {
                if (slot_side(rec.type, s) != side) continue;
                if (q < 3) {
                    sum += sHv[12 * r + 3 * s + q];
                } else {
                    for (int j = 0; j < 3; j++) sum += sHv[12 * r + 3 * s + j] * sG[s][3 * j + (q - 3)];
                }
            }
            sHJ[e] = sum;
        }
        __syncthreads();
        // B = J^T HJ + rotational terms; body gradient J^T g
        for (int e = lane; e < 144; e += 64) {
            const int rr = e / 12, col = e % 12, side = rr / 6, p = rr % 6;
            double sum = 0.0;
            for (int s = 0; s < nv; s++) {
                if (slot_side(rec.type, s) != side) continue;
                if (p < 3) {
                    sum += sHJ[12 * (3 * s + p) + col];
                } else {
                    for (int j = 0; j < 3; j++) sum += sG[s][3 * j + (p - 3)] * sHJ[12 * (3 * s + j) + col];
                }
            }
            if (col / 6 == side && p >= 3 && col % 6 >= 3)
                for (int s = 0; s < nv; s++)
                    if (slot_side(rec.type, s) == side) sum += sT[s][sym_index(p - 3, col % 6 - 3)];
            sB[e] = sum;
            sA[e] = sum;
            sV[e] = (rr == col) ? 1.0 : 0.0;
        }
        if (lane < 12) {
            const int side = lane / 6, p = lane % 6;
            double sum = 0.0;
            for (int s = 0; s < nv; s++) {
                if (slot_side(rec.type, s) != side) continue;
                if (p < 3) {
                    sum += sg[3 * s + p];
                } else {
                    for (int j = 0; j < 3; j++) sum += sG[s][3 * j + (p - 3)] * sg[3 * s + j];
                }
            }
            a.lgrad[12 * c + lane] = sum;
        }
        __syncthreads();
        // project_to_psd: Jacobi in parallel order; unchanged when PSD
        double prev_off = INFINITY;
        for (int sweep = 0; a.project && sweep < 100; sweep++) {
            double off = 0.0, tot = 0.0;
            for (int e = lane; e < 144; e += 64) {
                const double v = sA[e];
                tot += v * v;
                if (e / 12 != e % 12) off += v * v;
            }
            off = wave_sum_uniform(off);
            tot = wave_sum_uniform(tot);
            if (off <= 1e-30 * tot || off == 0) break;
            // Rank-deficient blocks settle at a rounding floor just above
            // that bound (their near-zero rows keep rounding noise): stop
            // once a sweep no longer cuts the off-diagonal mass by 4 and it
            // is already below 1e-24 of the total.
            if (off > 0.25 * prev_off && off <= 1e-24 * tot) break;
            prev_off = off;
            // Round r rotates its six pairs (jacobi_pair) at once. Lanes
            // 0..5 compute the rotations, which the other lanes read from
            // LDS; lane l < 36 owns the 2 x 2 block of A at (pair l / 6,
            // pair l % 6), and the 72 (row, pair) entries of V go to lanes l
            // and l + 64. One barrier publishes the rotations and separates
            // the round's reads from its writes, one follows the writes.
            for (int round = 0; round < 11; round++) {
                const int ia = lane < 36 ? lane / 6 : 0, ib = lane % 6;
                const int k1 = (lane + 64) % 6;
                int pa, qa, pb, qb, vp1, vq1;
                jacobi_pair(round, ia, pa, qa);
                jacobi_pair(round, ib, pb, qb);
                jacobi_pair(round, k1, vp1, vq1);
                if (lane < 6) {
                    double rc, rs;
                    jacobi_rotation(sA, pb, qb, rc, rs);
                    sCS[lane][0] = rc;
                    sCS[lane][1] = rs;
                }
                double x00 = 0.0, x01 = 0.0, x10 = 0.0, x11 = 0.0;
                if (lane < 36) {
                    x00 = sA[12 * pa + pb]; x01 = sA[12 * pa + qb];
                    x10 = sA[12 * qa + pb]; x11 = sA[12 * qa + qb];
                }
                // V entries: row t / 6 with pair t % 6, t = lane, lane + 64
                const int vr0 = lane / 6, vr1 = (lane + 64) / 6;
                const double u0 = sV[12 * vr0 + pb], w0 = sV[12 * vr0 + qb];
                double u1 = 0.0, w1 = 0.0;
                if (lane + 64 < 72) {
                    u1 = sV[12 * vr1 + vp1];
                    w1 = sV[12 * vr1 + vq1];
                }
                __syncthreads();
                const double ca = sCS[ia][0], sa = sCS[ia][1];
                const double cb = sCS[ib][0], sb = sCS[ib][1];
                const double c1 = sCS[k1][0], s1 = sCS[k1][1];
                if (lane < 36) {
                    // columns (pair b), then rows (pair a)
                    const double y00 = cb * x00 - sb * x01, y01 = sb * x00 + cb * x01;
                    const double y10 = cb * x10 - sb * x11, y11 = sb * x10 + cb * x11;
                    sA[12 * pa + pb] = ca * y00 - sa * y10;
                    sA[12 * qa + pb] = sa * y00 + ca * y10;
                    sA[12 * pa + qb] = ca * y01 - sa * y11;
                    sA[12 * qa + qb] = sa * y01 + ca * y11;
                }
                sV[12 * vr0 + pb] = cb * u0 - sb * w0;
                sV[12 * vr0 + qb] = sb * u0 + cb * w0;
                if (lane + 64 < 72) {
                    sV[12 * vr1 + vp1] = c1 * u1 - s1 * w1;
                    sV[12 * vr1 + vq1] = s1 * u1 + c1 * w1;
                }
                __syncthreads();
            }
        }
        bool psd = true;
        for (int i = 0; i < 12; i++) psd = psd && sA[13 * i] >= 0.0;
        psd = psd || !a.project;
        for (int e = lane; e < 144; e += 64) {
            double out;
            if (psd) {
                out = sB[e];
            } else {
                const int i = e / 12, j = e % 12;
                out = 0.0;
                for (int k = 0; k < 12; k++) {
                    const double lam = sA[13 * k] > 0.0 ? sA[13 * k] : 0.0;
                    out += sV[12 * i + k] * lam * sV[12 * j + k];
                }
            }
            a.lhess[144 * c + e] = out;
        }
        __syncthreads();
    }
}

// --------------------------------------------------------------- reduction
// Per-scene sums of per-constraint values, one block per selected scene.
// The records are grouped by class; within a class the records of one scene
// are contiguous and scenes appear in ascending order. Each scene's range of
// each class is summed with the same thread partition relative to the range
// start, and the class sums are added in class order, so a scene's sum does
// not depend on which other scenes share the call.
__device__ inline int rec_scene(const ConRec* cons, const FricRec* fric, const int* body_scene, long long i) {
    return body_scene[fric ? fric[i].c.body[0] : cons[i].body[0]];
}

__global__ __launch_bounds__(1024) void scene_sum_kernel(const double* x, const ConRec* cons, const FricRec* fric,
                                                        ClassRanges cr, const int* body_scene, const int* scenes,
                                                        int whole, double* out) {
    __shared__ double part[1024];
    __shared__ long long rng[2];
    const int sc = scenes[blockIdx.x];
    double acc = 0.0;
    for (int k = 0; k < cr.n_class; k++) {
        if (threadIdx.x == 0) {
            long long lo = cr.off[k], hi = cr.off[k + 1];
            if (!whole) {
                long long a = lo, b = hi;
                while (a < b) {
                    const long long m = (a + b) >> 1;
                    if (rec_scene(cons, fric, body_scene, m) < sc) a = m + 1;
                    else b = m;
                }
                long long c = a, d = hi;
                while (c < d) {
                    const long long m = (c + d) >> 1;
                    if (rec_scene(cons, fric, body_scene, m) <= sc) c = m + 1;
                    else d = m;
                }
                lo = a;
                hi = c;
            }
            rng[0] = lo;
            rng[1] = hi;
        }
        __syncthreads();
        double v = 0.0;
        for (long long i = rng[0] + threadIdx.x; i < rng[1]; i += 1024) v += x[i];
        part[threadIdx.x] = v;
        __syncthreads();
        for (int w = 512; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
            __syncthreads();
        }
        if (threadIdx.x == 0) acc += part[0];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = acc;
}

// Per selected scene, the record range of each class ([lo, hi) pairs).
__global__ void scene_class_ranges_kernel(const ConRec* cons, ClassRanges cr, const int* body_scene,
                                          const int* scenes, int n_sel, long long* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_sel) return;
    const int sc = scenes[i];
    for (int k = 0; k < 4; k++) {
        long long lo = 0, hi = 0;
        if (k < cr.n_class) {
            long long a = cr.off[k], b = cr.off[k + 1];
            while (a < b) {
                const long long m = (a + b) >> 1;
                if (body_scene[cons[m].body[0]] < sc) a = m + 1;
                else b = m;
            }
            long long c = a, d = cr.off[k + 1];
            while (c < d) {
                const long long m = (c + d) >> 1;
                if (body_scene[cons[m].body[0]] <= sc) c = m + 1;
                else d = m;
            }
            lo = a;
            hi = c;
        }
        out[8 * i + 2 * k] = lo;
        out[8 * i + 2 * k + 1] = hi;
    }
}

// key = body << 32 | (2 c + side)
__global__ __launch_bounds__(THREADS) void diag_keys_kernel(const ConRec* cons, const FricRec* fric, long long n,
                                                           unsigned long long* keys) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const ConRec& r = fric ? fric[c].c : cons[c];
    keys[2 * c] = ((unsigned long long)r.body[0] << 32) | (unsigned long long)(2 * c);
    keys[2 * c + 1] = ((unsigned long long)r.body[1] << 32) | (unsigned long long)(2 * c + 1);
}

// key = (min body * nb + max body) << idx_bits | c
__global__ __launch_bounds__(THREADS) void off_keys_kernel(const ConRec* cons, const FricRec* fric, long long n,
                                                          int nb, int idx_bits, unsigned long long* keys) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const ConRec& r = fric ? fric[c].c : cons[c];
    const unsigned long long lo = (unsigned long long)min(r.body[0], r.body[1]);
    const unsigned long long hi = (unsigned long long)max(r.body[0], r.body[1]);
    keys[c] = ((lo * (unsigned long long)nb + hi) << idx_bits) | (unsigned long long)c;
}

__device__ long long lower_bound_u64(const unsigned long long* k, long long n, unsigned long long x) {
    long long lo = 0, hi = n;
    while (lo < hi) {
        const long long mid = (lo + hi) >> 1;
        if (k[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Block per body: lanes 0..35 the diagonal 6x6 block, lanes 36..41 the
// gradient, summed in sorted (constraint) order. The terms are fetched
// BODY_BATCH at a time (independent loads in flight; a body on the plane
// has hundreds of constraints) and added in that same order.
constexpr int BODY_BATCH = 8;
__global__ __launch_bounds__(64) void body_reduce_kernel(const int* bodies, const unsigned long long* keys,
                                                        long long nk, const double* lgrad, const double* lhess,
                                                        int want_hess, double* grad, double* diag) {
    const int b = bodies[blockIdx.x], lane = threadIdx.x;
    const long long lo = lower_bound_u64(keys, nk, (unsigned long long)b << 32);
    const long long hi = lower_bound_u64(keys, nk, (unsigned long long)(b + 1) << 32);
    if (lane >= 42 || (lane < 36 && !want_hess)) return;
    // this lane's entry: offset within a constraint's 12 x 12 block (or 12
    // gradient entries) for side 0; side 1 adds 78 (or 6)
    const bool hess = lane < 36;
    const double* src = hess ? lhess : lgrad;
    const long long stride = hess ? 144 : 12;
    const int off0 = hess ? 12 * (lane / 6) + lane % 6 : lane - 36;
    const int side_step = hess ? 78 : 6;
    double s = 0.0;
    long long i = lo;
    for (; i + BODY_BATCH <= hi; i += BODY_BATCH) {
        double v[BODY_BATCH];
#pragma unroll
        for (int k = 0; k < BODY_BATCH; k++) {
            const unsigned long long t = keys[i + k] & 0xffffffffULL;
            v[k] = src[stride * (long long)(t >> 1) + off0 + side_step * (int)(t & 1ULL)];
        }
#pragma unroll
        for (int k = 0; k < BODY_BATCH; k++) s += v[k];
    }
    for (; i < hi; i++) {
        const unsigned long long t = keys[i] & 0xffffffffULL;
        s += src[stride * (long long)(t >> 1) + off0 + side_step * (int)(t & 1ULL)];
    }
    if (hess) diag[36 * b + lane] = s;
    else grad[6 * b + lane - 36] = s;
}

__global__ __launch_bounds__(THREADS) void off_heads_kernel(const unsigned long long* keys, long long n, int idx_bits,
                                                           unsigned int* flags) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    flags[i] = (i == 0 || (keys[i] >> idx_bits) != (keys[i - 1] >> idx_bits)) ? 1u : 0u;
}

// Block per sorted key; heads sum their run into block pos[i].
__global__ __launch_bounds__(64) void pair_reduce_kernel(const unsigned long long* keys, long long n, int idx_bits,
                                                        int nb, const unsigned int* flags, const unsigned int* pos,
                                                        const ConRec* cons, const FricRec* fric, const double* lhess,
                                                        int* pairs, double* blocks) {
    const long long i = blockIdx.x;
    if (!flags[i]) return;
    const int lane = threadIdx.x;
    const unsigned long long pk = keys[i] >> idx_bits;
    const int bi = (int)(pk / (unsigned long long)nb), bj = (int)(pk % (unsigned long long)nb);
    const unsigned long long mask = (idx_bits >= 64) ? ~0ULL : ((1ULL << idx_bits) - 1ULL);
    const unsigned int u = pos[i];
    if (lane < 36) {
        const int p = lane / 6, q = lane % 6;
        double s = 0.0;
        for (long long j = i; j < n && (keys[j] >> idx_bits) == pk; j++) {
            const long long c = (long long)(keys[j] & mask);
            const ConRec& r = fric ? fric[c].c : cons[c];
            // block (bi, bj) = local (side of bi, side of bj)
            if (r.body[0] == bi) s += lhess[144 * c + 12 * p + 6 + q];
            else s += lhess[144 * c + 12 * (6 + p) + q];
        }
        blocks[36 * (long long)u + lane] = s;
    } else if (lane == 36) {
        pairs[2 * (long long)u] = bi;
        pairs[2 * (long long)u + 1] = bj;
    }
}

// --------------------------------------------------------------- friction set
__global__ __launch_bounds__(THREADS) void friction_record_kernel(const ConRec* cons, long long n, const double* W,
                                                                 double dhat, double dmin, const double* kappa_scene,
                                                                 const double* mu_scene, const int* body_scene,
                                                                 FricRec* out) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const ConRec r = cons[c];
    const int sc = body_scene[r.body[0]];
    const double kappa = kappa_scene[sc], mu = mu_scene[sc];
    const int nv = con_nv(r.type);
    double X[12];
    for (int s = 0; s < 4; s++)
        for (int k = 0; k < 3; k++) X[3 * s + k] = s < nv ? W[3 * r.v[s] + k] : 0.0;
    DistForm f;
    constraint_form(r.type, X, f);
    FricRec o;
    o.c = r;
    double basis[3][2];
    friction_frame(r.type, X, o.coords, basis);
    for (int k = 0; k < 3; k++) {
        o.basis[2 * k] = basis[k][0];
        o.basis[2 * k + 1] = basis[k][1];
    }
    const double dh = 2 * dmin * dhat + dhat * dhat, dm2 = dmin * dmin;
    // compute_normal_force_magnitude: -kappa b'(d) 2 sqrt(d)
    o.normal_force = -kappa * (ipc_barrier_gradient(f.d - dm2, dh) * (2 * sqrt(f.d)));
    o.mu = mu;
    out[c] = o;
}

// Friction adjoint, one thread per lagged friction record: with
// y = J(x) lambda (the record's vertex-space image of lambda at the
// evaluation poses), the record's share of (d grad_x D / d x0)^T lambda in
// the vertex space of x0 is
//   -H_v y                       (D depends on x0 through U = V(x) - V(x0);
//                                 H_v = m mu lambda_N T M T^T, the friction
//                                 Hessian in U)
//   + (g1 . y) m mu dlambda_N/dd grad_V d(x0)
//                                (through lambda_N = -kappa b'(d) 2 sqrt(d);
//                                 g1 = f1(|T^T U|) T T^T U / |T^T U|)
// and J(x0)^T maps it to the two bodies' dofs at x0. The frozen tangent
// basis T and closest-point weights are held fixed.
__global__ __launch_bounds__(THREADS) void friction_adjoint_kernel(const FricRec* fric, long long n, const double* V,
                                                                  const int* v_body, const double* W,
                                                                  const double* W0, const BodyFrame* F1,
                                                                  const BodyFrame* F0, double epsv_h, double dhat,
                                                                  double dmin, double kappa, const double* lambda,
                                                                  double* out) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const FricRec f = fric[c];
    const ConRec& r = f.c;
    const int nv = con_nv(r.type);
    double U[12], X0[12];
    for (int s = 0; s < 4; s++)
        for (int k = 0; k < 3; k++) {
            U[3 * s + k] = s < nv ? W[3 * r.v[s] + k] - W0[3 * r.v[s] + k] : 0.0;
            X0[3 * s + k] = s < nv ? W0[3 * r.v[s] + k] : 0.0;
        }
    double basis[3][2];
    for (int k = 0; k < 3; k++) {
        basis[k][0] = f.basis[2 * k];
        basis[k][1] = f.basis[2 * k + 1];
    }
    FrictionLocal L;
    friction_local_setup(L, r.type, nv, f.coords, basis, f.normal_force, f.mu, r.mult, U, epsv_h);
    double y[12];
    for (int e = 0; e < 12; e++) y[e] = 0.0;
    for (int s = 0; s < nv; s++) {
        const int v = r.v[s];
        const int b = v_body[v];
        const double* lam = lambda + 6 * b;
        const double* x = V + 3 * v;
        for (int k = 0; k < 3; k++) {
            double sum = lam[k];
            for (int d = 0; d < 3; d++) sum += jac_rot(F1[b], x, d, k) * lam[3 + d];
            y[3 * s + k] = sum;
        }
    }
    double ty[2] = { 0.0, 0.0 };
    for (int e = 0; e < 12; e++) {
        ty[0] += L.T[e][0] * y[e];
        ty[1] += L.T[e][1] * y[e];
    }
    const double mt0 = L.M[0][0] * ty[0] + L.M[0][1] * ty[1];
    const double mt1 = L.M[1][0] * ty[0] + L.M[1][1] * ty[1];
    double z[12];
    for (int e = 0; e < 12; e++) z[e] = -L.scale * (L.T[e][0] * mt0 + L.T[e][1] * mt1);
    DistForm fm;
    constraint_form(r.type, X0, fm);
    const double dh = 2 * dmin * dhat + dhat * dhat, dm2 = dmin * dmin;
    const double sd = sqrt(fm.d);
    const double b1 = ipc_barrier_gradient(fm.d - dm2, dh), b2 = ipc_barrier_hessian(fm.d - dm2, dh);
    const double dlam = sd > 0.0 ? -kappa * (b2 * (2 * sd) + b1 / sd) : 0.0;
    const double g1y = L.f1x * (L.u[0] * ty[0] + L.u[1] * ty[1]);
    const double coeff = g1y * ((double)r.mult * f.mu * dlam);
    for (int s = 0; s < nv; s++)
        for (int k = 0; k < 3; k++) z[3 * s + k] += coeff * form_grad_vertex(fm, s, k);
    double o[12];
    for (int e = 0; e < 12; e++) o[e] = 0.0;
    for (int s = 0; s < nv; s++) {
        const int side = slot_side(r.type, s);
        const int v = r.v[s];
        const int b = v_body[v];
        const double* x = V + 3 * v;
        for (int k = 0; k < 3; k++) o[6 * side + k] += z[3 * s + k];
        for (int d = 0; d < 3; d++) {
            double sum = 0.0;
            for (int k = 0; k < 3; k++) sum += jac_rot(F0[b], x, d, k) * z[3 * s + k];
            o[6 * side + 3 + d] += sum;
        }
    }
    for (int e = 0; e < 12; e++) out[12 * c + e] = o[e];
}

// Minimum squared distance over the set (non-negative doubles order as
// their bit patterns).
__global__ __launch_bounds__(THREADS) void min_distance_kernel(const ConRec* cons, long long n, const double* W,
                                                              const int* body_scene, unsigned long long* out_scene) {
    const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const ConRec r = cons[c];
    const int nv = con_nv(r.type);
    double X[12];
    for (int s = 0; s < 4; s++)
        for (int k = 0; k < 3; k++) X[3 * s + k] = s < nv ? W[3 * r.v[s] + k] : 0.0;
    DistForm f;
    constraint_form(r.type, X, f);
    unsigned long long bits;
    memcpy(&bits, &f.d, sizeof(bits));
    atomicMin(out_scene + body_scene[r.body[0]], bits);
}

unsigned grid_for(long long n, int threads = THREADS) { return (unsigned)((n + threads - 1) / threads); }

// World vertices into W (as.W or as.W0) and, with `frames`, the body frames
// at d_poses; nothing is recomputed when the pose array's content tag shows
// that W (and the frames) already hold its values for every body.
hipError_t world_vertices(Assembly& as, const DevScene& sc, const SceneSel& sel, const double* d_poses, DevBuf& W,
                          bool frames, hipStream_t st) {
    // W over every body brings the frames along: a gradient at the same
    // poses (the usual next call) then needs no second launch
    frames = frames || (sel.all && &W == &as.W);
    unsigned long long t = 0;
    for (int k = 0; k < 3; k++)
        if (as.tag_ptr[k] == d_poses) t = as.tag[k];
    unsigned long long& w_tag = &W == &as.W0 ? as.W0_tag : as.W_tag;
    if (sel.all && t && w_tag == t && (!frames || as.F_tag == t)) return hipSuccess;
    w_tag = 0;
    if (frames) as.F_tag = 0;
    AS_CHECK(as.R.ensure(9 * sizeof(double) * sc.nb));
    AS_CHECK(as.frames.ensure(sizeof(BodyFrame) * sc.nb));
    AS_CHECK(W.ensure(3 * sizeof(double) * std::max(sc.nv, 1)));
    if (sel.all) {
        hipLaunchKernelGGL(body_rotation_kernel, dim3(grid_for(sc.nb)), dim3(THREADS), 0, st, sc.nb, d_poses,
                           as.R.as<double>());
        if (frames)
            hipLaunchKernelGGL(body_frame_kernel, dim3(grid_for(sc.nb), 9), dim3(THREADS), 0, st, sc.nb, d_poses,
                               as.frames.as<BodyFrame>());
        if (sc.nv)
            hipLaunchKernelGGL(world_vertex_kernel, dim3(grid_for(sc.nv)), dim3(THREADS), 0, st, sc.nv, sc.V,
                               sc.v_body, d_poses, as.R.as<double>(), W.as<double>());
    } else if (sel.n_bodies) {
        hipLaunchKernelGGL(body_list_vertex_kernel, dim3((unsigned)sel.n_bodies), dim3(THREADS), 0, st, sel.bodies,
                           sc.body_vstart, sc.V, d_poses, as.R.as<double>(),
                           frames ? as.frames.as<BodyFrame>() : nullptr, W.as<double>());
    }
    AS_CHECK(hipGetLastError());
    if (sel.all) {
        w_tag = t;
        if (frames) as.F_tag = t;
    }
    return hipSuccess;
}

__global__ void gather_rows_kernel(const int* rows, int n, int width, const double* src, double* dst) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)n * width) return;
    const long long r = i / width, k = i % width;
    dst[i] = src[(long long)rows[r] * width + k];
}

__global__ void scatter_rows_kernel(const int* rows, int n, int width, const double* src, double* dst) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long long)n * width) return;
    const long long r = i / width, k = i % width;
    dst[(long long)rows[r] * width + k] = src[i];
}

// Byte-range copies: entry e = (source 0 / 1, source offset, destination
// offset, bytes), all multiples of 16; one block per entry.
__global__ __launch_bounds__(THREADS) void copy_ranges_kernel(const long long* ent, const char* src0, const char* src1,
                                                             char* dst) {
    const long long* e = ent + 4 * blockIdx.x;
    const uint4* s = reinterpret_cast<const uint4*>((e[0] ? src1 : src0) + e[1]);
    uint4* d = reinterpret_cast<uint4*>(dst + e[2]);
    const long long n = e[3] / 16;
    for (long long i = threadIdx.x; i < n; i += THREADS) d[i] = s[i];
}

ClassRanges cons_ranges(const Assembly& as) {
    ClassRanges cr{};
    cr.n_class = 4;
    cr.off[0] = 0;
    for (int k = 0; k < 4; k++) cr.off[k + 1] = cr.off[k] + as.counts[k];
    return cr;
}

// Per selected scene, [lo, hi) of each class of the current constraint set
// (host copy, [n_sel][8]).
hipError_t scene_ranges(Assembly& as, const DevScene& sc, const SceneSel& sel, std::vector<long long>& h,
                        hipStream_t st) {
    h.assign(8 * (size_t)sel.n_scenes, 0);
    if (sel.n_scenes == 0 || as.n_con == 0) return hipSuccess;
    AS_CHECK(as.ranges.ensure(8 * sizeof(long long) * sel.n_scenes));
    hipLaunchKernelGGL(scene_class_ranges_kernel, dim3(grid_for(sel.n_scenes)), dim3(THREADS), 0, st,
                       as.cons.as<ConRec>(), cons_ranges(as), sc.body_scene, sel.scenes, sel.n_scenes,
                       as.ranges.as<long long>());
    AS_CHECK(hipGetLastError());
    AS_CHECK(as.h_ranges.ensure(sizeof(long long) * h.size()));
    AS_CHECK(hipMemcpyAsync(as.h_ranges.p, as.ranges.p, sizeof(long long) * h.size(), hipMemcpyDeviceToHost, st));
    AS_CHECK(hipStreamSynchronize(st));
    std::memcpy(h.data(), as.h_ranges.p, sizeof(long long) * h.size());
    return hipSuccess;
}

hipError_t run_copies(Assembly& as, const std::vector<long long>& ent, const void* src0, const void* src1, void* dst,
                      hipStream_t st) {
    const long long n = (long long)ent.size() / 4;
    if (n == 0) return hipSuccess;
    // the table is staged in page-locked memory: the caller synchronises the
    // stream before the next table is staged
    AS_CHECK(as.h_entries.ensure(sizeof(long long) * ent.size()));
    std::memcpy(as.h_entries.p, ent.data(), sizeof(long long) * ent.size());
    AS_CHECK(as.entries.ensure(sizeof(long long) * ent.size()));
    AS_CHECK(hipMemcpyAsync(as.entries.p, as.h_entries.p, sizeof(long long) * ent.size(), hipMemcpyHostToDevice,
                            st));
    hipLaunchKernelGGL(copy_ranges_kernel, dim3((unsigned)n), dim3(THREADS), 0, st, as.entries.as<long long>(),
                       static_cast<const char*>(src0), static_cast<const char*>(src1), static_cast<char*>(dst));
    return hipGetLastError();
}

}  // namespace

// World vertices at d_poses in as.W (kept while the pose tag is unchanged).
hipError_t current_world_vertices(Assembly& as, const DevScene& sc, const SceneSel& sel, const double* d_poses,
                                  hipStream_t st) {
    return world_vertices(as, sc, sel, d_poses, as.W, false, st);
}

void gather_rows(const int* rows, int n, int width, const double* src, double* dst, hipStream_t st) {
    if (n > 0)
        hipLaunchKernelGGL(gather_rows_kernel, dim3(grid_for((long long)n * width)), dim3(THREADS), 0, st, rows, n,
                           width, src, dst);
}

void scatter_rows(const int* rows, int n, int width, const double* src, double* dst, hipStream_t st) {
    if (n > 0)
        hipLaunchKernelGGL(scatter_rows_kernel, dim3(grid_for((long long)n * width)), dim3(THREADS), 0, st, rows, n,
                           width, src, dst);
}

hipError_t build_constraint_set(Assembly& as, BroadPhase& bp, const DevScene& sc, const SceneSel& sel,
                                const double* d_poses, double dhat, double dmin, int num_cus, hipStream_t st) {
    (void)num_cus;
    as.have_cons = false;  // the lagged friction sets are independent and survive
    as.n_con = 0;
    for (int k = 0; k < 4; k++) as.counts[k] = 0;
    if (sc.nb <= 1 || sel.n_bodies <= 1) {
        as.have_cons = true;
        return hipSuccess;
    }
    if (sel.n_tiles == 0) {
        as.have_cons = true;
        return hipSuccess;
    }
    // The broad phase is queued without a read-back: classification reads
    // its packed keys and counts on the device, and one synchronisation
    // returns both the candidate and the class counts (a pass whose
    // candidates overflowed the capacities is run again at the counts).
    AS_CHECK(queue_broad_phase(bp, sc, sel, d_poses, d_poses, (dhat + dmin) / 2.0, false, st));
    AS_CHECK(world_vertices(as, sc, sel, d_poses, as.W, false, st));
    AS_CHECK(as.ccount.ensure(8 * sizeof(unsigned int)));
    AS_CHECK(as.h_mail.ensure(sizeof(unsigned int) * MAIL_SLOTS));
    AS_CHECK(bp.h_counters.ensure(4 * sizeof(unsigned long long)));
    unsigned int* mail = as.h_mail.as<unsigned int>();
    const unsigned long long* got = bp.h_counters.as<unsigned long long>();
    bool fits = false;
    for (int attempt = 0; attempt < 3 && !fits; attempt++) {
        if (attempt > 0) AS_CHECK(requeue_pairs(bp, sc, sel, st));
        const long long cap = bp.cap_ee + bp.cap_fv;
        for (int k = 0; k < 4; k++) AS_CHECK(as.ckeys[k].ensure(sizeof(unsigned long long) * cap));
        AS_CHECK(hipMemsetAsync(as.ccount.p, 0, 8 * sizeof(unsigned int), st));
        ClassifyArgs ca;
        ca.kee = bp.keys_ee.as<unsigned long long>();
        ca.kfv = bp.keys_fv.as<unsigned long long>();
        ca.n_pairs = bp.counters.as<unsigned long long>();
        ca.cap_ee = bp.cap_ee;
        ca.cap_fv = bp.cap_fv;
        ca.ne = (unsigned long long)sc.ne;
        ca.E = sc.E;
        ca.F = sc.F;
        ca.FE = sc.FE;
        ca.W = as.W.as<double>();
        ca.nv = sc.nv;
        ca.offset_sqr = (dmin + dhat) * (dmin + dhat);
        for (int k = 0; k < 4; k++) ca.keys[k] = as.ckeys[k].as<unsigned long long>();
        ca.count = as.ccount.as<unsigned int>();
        hipLaunchKernelGGL(classify_kernel, dim3(grid_for(cap)), dim3(THREADS), 0, st, ca);
        AS_CHECK(hipGetLastError());
        AS_CHECK(hipMemcpyAsync(bp.h_counters.p, bp.counters.p, 3 * sizeof(unsigned long long),
                                hipMemcpyDeviceToHost, st));
        AS_CHECK(hipMemcpyAsync(mail + MAIL_CLASS, as.ccount.p, 4 * sizeof(unsigned int), hipMemcpyDeviceToHost,
                                st));
        AS_CHECK(hipStreamSynchronize(st));
        fits = broad_phase_fits(bp, got);
    }
    if (!fits) return hipErrorUnknown;
    bp.n_body_pairs = (long long)got[2];
    bp.n_ee = (long long)got[0];
    bp.n_fv = (long long)got[1];
    if (bp.n_ee + bp.n_fv == 0) {
        as.have_cons = true;
        return hipSuccess;
    }
    unsigned int cnt[4];
    for (int k = 0; k < 4; k++) cnt[k] = mail[MAIL_CLASS + k];

    const unsigned long long key_max[4] = {
        (unsigned long long)sc.nv * (unsigned long long)sc.nv, (unsigned long long)sc.ne * (unsigned long long)sc.nv,
        (unsigned long long)sc.ne * (unsigned long long)sc.ne, (unsigned long long)sc.nf * (unsigned long long)sc.nv };
    long long nmax = 1;
    for (int k = 0; k < 4; k++) nmax = std::max<long long>(nmax, cnt[k]);
    AS_CHECK(as.ctmp.ensure(sizeof(unsigned long long) * nmax));
    AS_CHECK(as.chist.ensure(sizeof(unsigned int) * radix_sort_hist_entries(nmax)));
    AS_CHECK(as.cflags.ensure(sizeof(unsigned int) * nmax));
    AS_CHECK(as.cpos.ensure(sizeof(unsigned int) * (2 * nmax + 2)));
    AS_CHECK(as.cwork.ensure(sizeof(unsigned int) * (scan_workspace_entries(nmax) + 4)));
    for (int k = 0; k < 4; k++)
        if (cnt[k] > 1)
            AS_CHECK(radix_sort_u64(as.ckeys[k].as<unsigned long long>(), as.ctmp.as<unsigned long long>(),
                                    as.chist.as<unsigned int>(), cnt[k], bits_for(key_max[k]), st));

    // run-length encode VV and EV (multiplicities); run starts kept per
    // class; both run counts come back with one synchronisation
    long long nout[4] = { cnt[0], cnt[1], cnt[2], cnt[3] };
    DevBuf* starts[2] = { &as.cstart[0], &as.cstart[1] };
    bool runs = false;
    for (int k = 0; k < 2; k++) {
        if (cnt[k] == 0) continue;
        unsigned int* flags = as.cflags.as<unsigned int>();
        unsigned int* pos = as.cpos.as<unsigned int>();
        unsigned int* total = as.cwork.as<unsigned int>() + scan_workspace_entries(nmax) + k;
        AS_CHECK(starts[k]->ensure(sizeof(unsigned int) * cnt[k]));
        hipLaunchKernelGGL(head_flags_kernel, dim3(grid_for(cnt[k])), dim3(THREADS), 0, st,
                           as.ckeys[k].as<unsigned long long>(), (long long)cnt[k], flags);
        exclusive_scan_u32(flags, pos, cnt[k], as.cwork.as<unsigned int>(), total, st);
        hipLaunchKernelGGL(head_starts_kernel, dim3(grid_for(cnt[k])), dim3(THREADS), 0, st, flags, pos,
                           (long long)cnt[k], starts[k]->as<unsigned int>());
        AS_CHECK(hipMemcpyAsync(mail + MAIL_RUNS + k, total, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
        runs = true;
    }
    if (runs) {
        AS_CHECK(hipStreamSynchronize(st));
        for (int k = 0; k < 2; k++)
            if (cnt[k]) nout[k] = mail[MAIL_RUNS + k];
    }
    const long long n = nout[0] + nout[1] + nout[2] + nout[3];
    AS_CHECK(as.cons.ensure(sizeof(ConRec) * std::max<long long>(n, 1)));
    long long off = 0;
    for (int k = 0; k < 4; k++) {
        if (nout[k] == 0) continue;
        RecordArgs ra;
        ra.cls = k;
        ra.keys = as.ckeys[k].as<unsigned long long>();
        ra.start = k < 2 ? starts[k]->as<unsigned int>() : nullptr;
        ra.n_keys = cnt[k];
        ra.n_out = nout[k];
        ra.out = as.cons.as<ConRec>() + off;
        ra.ne = (unsigned long long)sc.ne;
        ra.E = sc.E;
        ra.F = sc.F;
        ra.v_body = sc.v_body;
        ra.W = as.W.as<double>();
        ra.nv = sc.nv;
        hipLaunchKernelGGL(record_kernel, dim3(grid_for(nout[k])), dim3(THREADS), 0, st, ra);
        off += nout[k];
    }
    AS_CHECK(hipGetLastError());
    for (int k = 0; k < 4; k++) as.counts[k] = nout[k];
    as.n_con = n;
    as.have_cons = true;
    return hipSuccess;
}

hipError_t constraint_counts(Assembly& as, const DevScene& sc, const SceneSel& sel, std::vector<long long>& counts,
                             hipStream_t st) {
    counts.assign(4 * (size_t)sel.n_scenes, 0);
    if (sc.n_scenes == 1) {
        for (int k = 0; k < 4 && sel.n_scenes == 1; k++) counts[k] = as.counts[k];
        return hipSuccess;
    }
    std::vector<long long> h;
    AS_CHECK(scene_ranges(as, sc, sel, h, st));
    for (int i = 0; i < sel.n_scenes; i++)
        for (int k = 0; k < 4; k++) counts[4 * i + k] = h[8 * i + 2 * k + 1] - h[8 * i + 2 * k];
    return hipSuccess;
}

hipError_t build_friction_set(Assembly& as, const DevScene& sc, const SceneSel& sel, const double* d_poses,
                              double dhat, double dmin, const double* d_kappa_scene, const double* d_mu_scene,
                              hipStream_t st) {
    const int ns = sc.n_scenes;
    if ((int)as.fric_off.size() != ns + 1) {
        as.fric_off.assign(ns + 1, 0);
        as.fric_valid.assign(ns, 0);
    }
    AS_CHECK(as.fric_new.ensure(sizeof(FricRec) * std::max<long long>(as.n_con, 1)));
    if (as.n_con) {
        AS_CHECK(world_vertices(as, sc, sel, d_poses, as.W, false, st));
        hipLaunchKernelGGL(friction_record_kernel, dim3(grid_for(as.n_con)), dim3(THREADS), 0, st,
                           as.cons.as<ConRec>(), as.n_con, as.W.as<double>(), dhat, dmin, d_kappa_scene, d_mu_scene,
                           sc.body_scene, as.fric_new.as<FricRec>());
        AS_CHECK(hipGetLastError());
    }
    // Merge: the selected scenes' new records (scene-major, class order kept
    // within a scene) and the other scenes' current records.
    std::vector<long long> rng;
    if (ns == 1) {
        rng.assign(8, 0);
        for (int k = 0, o = 0; k < 4; k++) {
            rng[2 * k] = o;
            rng[2 * k + 1] = o + as.counts[k];
            o += (int)as.counts[k];
        }
    } else {
        AS_CHECK(scene_ranges(as, sc, sel, rng, st));
    }
    std::vector<long long> new_off(ns + 1, 0), ent;
    std::vector<int> pos(ns, -1);
    for (int i = 0; i < sel.n_scenes; i++) pos[sel.h_scenes[i]] = i;
    const long long R = (long long)sizeof(FricRec);
    for (int sidx = 0; sidx < ns; sidx++) {
        long long cnt = 0;
        if (pos[sidx] >= 0) {
            const long long* r = &rng[8 * pos[sidx]];
            for (int k = 0; k < 4; k++) {
                const long long m = r[2 * k + 1] - r[2 * k];
                if (m > 0) {
                    ent.insert(ent.end(), { 0, r[2 * k] * R, (new_off[sidx] + cnt) * R, m * R });
                    cnt += m;
                }
            }
        } else if (as.fric_valid[sidx]) {
            cnt = as.fric_off[sidx + 1] - as.fric_off[sidx];
            if (cnt > 0) ent.insert(ent.end(), { 1, as.fric_off[sidx] * R, new_off[sidx] * R, cnt * R });
        }
        new_off[sidx + 1] = new_off[sidx] + cnt;
    }
    const long long total = new_off[ns];
    AS_CHECK(as.fric_tmp.ensure(sizeof(FricRec) * std::max<long long>(total, 1)));
    AS_CHECK(run_copies(as, ent, as.fric_new.p, as.fric.p, as.fric_tmp.p, st));
    std::swap(as.fric, as.fric_tmp);
    as.fric_off = new_off;
    for (int i = 0; i < sel.n_scenes; i++) as.fric_valid[sel.h_scenes[i]] = 1;
    as.n_fric = total;
    as.have_fric = true;
    for (int sidx = 0; sidx < ns; sidx++) as.have_fric = as.have_fric && as.fric_valid[sidx];
    return hipStreamSynchronize(st);
}

hipError_t minimum_distance(Assembly& as, const DevScene& sc, const SceneSel& sel, const double* d_poses,
                            std::vector<double>& min_sq, hipStream_t st) {
    min_sq.assign(sel.n_scenes, HUGE_VAL);
    if (as.n_con == 0) return hipSuccess;
    AS_CHECK(world_vertices(as, sc, sel, d_poses, as.W, false, st));
    AS_CHECK(as.mind.ensure(sizeof(unsigned long long) * sc.n_scenes));
    // staging: [0, n_scenes) the +inf start values, [n_scenes, 2 n_scenes) the result
    AS_CHECK(as.h_mind.ensure(2 * sizeof(unsigned long long) * sc.n_scenes));
    unsigned long long* bits = as.h_mind.as<unsigned long long>();
    for (int i = 0; i < sc.n_scenes; i++) bits[i] = 0x7ff0000000000000ULL;
    AS_CHECK(hipMemcpyAsync(as.mind.p, bits, sizeof(unsigned long long) * sc.n_scenes, hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(min_distance_kernel, dim3(grid_for(as.n_con)), dim3(THREADS), 0, st, as.cons.as<ConRec>(),
                       as.n_con, as.W.as<double>(), sc.body_scene, as.mind.as<unsigned long long>());
    AS_CHECK(hipGetLastError());
    AS_CHECK(hipMemcpyAsync(bits + sc.n_scenes, as.mind.p, sizeof(unsigned long long) * sc.n_scenes,
                            hipMemcpyDeviceToHost, st));
    AS_CHECK(hipStreamSynchronize(st));
    for (int i = 0; i < sel.n_scenes; i++)
        std::memcpy(&min_sq[i], &bits[sc.n_scenes + sel.h_scenes[i]], sizeof(double));
    return hipSuccess;
}

hipError_t assemble_potential(Assembly& as, const DevScene& sc, const SceneSel& sel, bool friction,
                              const double* d_poses, const double* d_poses_t0, double dhat, double dmin,
                              double epsv_h, const double* d_epsv_h_scene, bool want_grad, bool want_hess,
                              int num_cus, AssemblyOut& out, hipStream_t st, bool project) {
    out = AssemblyOut{};
    out.energy.assign(sel.n_scenes, 0.0);
    out.hess = want_hess;
    as.n_off = 0;
    const int nb = sc.nb;
    want_grad = want_grad || want_hess;
    AS_CHECK(as.grad.ensure(6 * sizeof(double) * nb));
    AS_CHECK(as.diag.ensure(36 * sizeof(double) * nb));

    // the records: the constraint set, or the selected scenes' friction sets
    const FricRec* fric = nullptr;
    long long n = as.n_con;
    if (friction) {
        if (sel.all) {
            fric = as.fric.as<FricRec>();
            n = as.n_fric;
        } else {
            std::vector<long long> ent;
            n = 0;
            const long long R = (long long)sizeof(FricRec);
            for (int i = 0; i < sel.n_scenes; i++) {
                const int sidx = sel.h_scenes[i];
                const long long m = as.fric_off[sidx + 1] - as.fric_off[sidx];
                if (m > 0) ent.insert(ent.end(), { 0, as.fric_off[sidx] * R, n * R, m * R });
                n += m;
            }
            AS_CHECK(as.fric_stage.ensure(sizeof(FricRec) * std::max<long long>(n, 1)));
            AS_CHECK(run_copies(as, ent, as.fric.p, as.fric.p, as.fric_stage.p, st));
            fric = as.fric_stage.as<FricRec>();
        }
    }
    if (n == 0) {
        if (want_grad) AS_CHECK(hipMemsetAsync(as.grad.p, 0, 6 * sizeof(double) * nb, st));
        if (want_hess) AS_CHECK(hipMemsetAsync(as.diag.p, 0, 36 * sizeof(double) * nb, st));
        return hipGetLastError();
    }
    out.n_rec = n;
    AS_CHECK(as.h_mail.ensure(sizeof(unsigned int) * MAIL_SLOTS));
    AS_CHECK(as.h_esum.ensure(sizeof(double) * std::max(sel.n_scenes, 1)));
    AS_CHECK(world_vertices(as, sc, sel, d_poses, as.W, want_grad, st));
    if (friction) AS_CHECK(world_vertices(as, sc, sel, d_poses_t0, as.W0, false, st));  // frames stay at d_poses
    AS_CHECK(as.energy.ensure(sizeof(double) * n));
    AS_CHECK(as.lgrad.ensure(12 * sizeof(double) * n));
    if (want_hess) AS_CHECK(as.lhess.ensure(144 * sizeof(double) * n));
    AS_CHECK(as.esum.ensure(sizeof(double) * std::max(sel.n_scenes, 1)));
    LocalArgs la;
    la.cons = as.cons.as<ConRec>();
    la.fric = fric;
    la.n = n;
    la.V = sc.V;
    la.v_body = sc.v_body;
    la.W = as.W.as<double>();
    la.W0 = friction ? as.W0.as<double>() : nullptr;
    la.frames = as.frames.as<BodyFrame>();
    la.dhat = dhat;
    la.dmin = dmin;
    la.epsv_h = epsv_h;
    la.epsv_h_scene = d_epsv_h_scene;
    la.body_scene = sc.body_scene;
    la.want_grad = want_grad ? 1 : 0;
    la.project = project ? 1 : 0;
    la.energy = as.energy.as<double>();
    la.lgrad = as.lgrad.as<double>();
    la.lhess = want_hess ? as.lhess.as<double>() : nullptr;
    if (want_hess) {
        const unsigned grid = (unsigned)std::min<long long>(n, (long long)num_cus * 32);
        if (friction) hipLaunchKernelGGL(local_hess_kernel<true>, dim3(grid), dim3(64), 0, st, la);
        else hipLaunchKernelGGL(local_hess_kernel<false>, dim3(grid), dim3(64), 0, st, la);
    } else {
        if (friction) hipLaunchKernelGGL(local_eg_kernel<true>, dim3(grid_for(n)), dim3(THREADS), 0, st, la);
        else hipLaunchKernelGGL(local_eg_kernel<false>, dim3(grid_for(n)), dim3(THREADS), 0, st, la);
    }
    AS_CHECK(hipGetLastError());
    ClassRanges cr{};
    if (friction) {
        cr.n_class = 1;
        cr.off[0] = 0;
        cr.off[1] = n;
    } else {
        cr = cons_ranges(as);
    }
    hipLaunchKernelGGL(scene_sum_kernel, dim3((unsigned)sel.n_scenes), dim3(1024), 0, st, as.energy.as<double>(),
                       friction ? nullptr : as.cons.as<ConRec>(), fric, cr, sc.body_scene, sel.scenes,
                       sc.n_scenes == 1 ? 1 : 0, as.esum.as<double>());

    const ConRec* cons = friction ? nullptr : as.cons.as<ConRec>();
    if (want_grad) {
        const long long nk = 2 * n;
        AS_CHECK(as.dkeys.ensure(sizeof(unsigned long long) * nk));
        AS_CHECK(as.ktmp.ensure(sizeof(unsigned long long) * nk));
        AS_CHECK(as.khist.ensure(sizeof(unsigned int) * radix_sort_hist_entries(nk)));
        hipLaunchKernelGGL(diag_keys_kernel, dim3(grid_for(n)), dim3(THREADS), 0, st, cons, fric, n,
                           as.dkeys.as<unsigned long long>());
        const int bits = 32 + bits_for((unsigned long long)nb);
        AS_CHECK(radix_sort_u64(as.dkeys.as<unsigned long long>(), as.ktmp.as<unsigned long long>(),
                                as.khist.as<unsigned int>(), nk, bits, st));
        hipLaunchKernelGGL(body_reduce_kernel, dim3((unsigned)sel.n_bodies), dim3(64), 0, st, sel.bodies,
                           as.dkeys.as<unsigned long long>(), nk, as.lgrad.as<double>(),
                           want_hess ? as.lhess.as<double>() : nullptr, want_hess ? 1 : 0, as.grad.as<double>(),
                           as.diag.as<double>());
        AS_CHECK(hipGetLastError());
    }
    if (want_hess) {
        const int idx_bits = std::max(1, bits_for((unsigned long long)n));
        const int pair_bits = bits_for((unsigned long long)nb * (unsigned long long)nb);
        if (idx_bits + pair_bits > 64) return hipErrorInvalidValue;
        AS_CHECK(as.okeys.ensure(sizeof(unsigned long long) * n));
        AS_CHECK(as.ktmp.ensure(sizeof(unsigned long long) * n));
        AS_CHECK(as.khist.ensure(sizeof(unsigned int) * radix_sort_hist_entries(n)));
        AS_CHECK(as.oflags.ensure(sizeof(unsigned int) * n));
        AS_CHECK(as.opos.ensure(sizeof(unsigned int) * n));
        AS_CHECK(as.owork.ensure(sizeof(unsigned int) * (scan_workspace_entries(n) + 4)));
        AS_CHECK(as.opairs.ensure(2 * sizeof(int) * n));
        AS_CHECK(as.oblocks.ensure(36 * sizeof(double) * n));
        hipLaunchKernelGGL(off_keys_kernel, dim3(grid_for(n)), dim3(THREADS), 0, st, cons, fric, n, nb, idx_bits,
                           as.okeys.as<unsigned long long>());
        AS_CHECK(radix_sort_u64(as.okeys.as<unsigned long long>(), as.ktmp.as<unsigned long long>(),
                                as.khist.as<unsigned int>(), n, idx_bits + pair_bits, st));
        hipLaunchKernelGGL(off_heads_kernel, dim3(grid_for(n)), dim3(THREADS), 0, st,
                           as.okeys.as<unsigned long long>(), n, idx_bits, as.oflags.as<unsigned int>());
        unsigned int* total = as.owork.as<unsigned int>() + scan_workspace_entries(n);
        exclusive_scan_u32(as.oflags.as<unsigned int>(), as.opos.as<unsigned int>(), n, as.owork.as<unsigned int>(),
                           total, st);
        hipLaunchKernelGGL(pair_reduce_kernel, dim3((unsigned)n), dim3(64), 0, st, as.okeys.as<unsigned long long>(),
                           n, idx_bits, nb, as.oflags.as<unsigned int>(), as.opos.as<unsigned int>(), cons, fric,
                           as.lhess.as<double>(), as.opairs.as<int>(), as.oblocks.as<double>());
        AS_CHECK(hipGetLastError());
        AS_CHECK(hipMemcpyAsync(as.h_mail.as<unsigned int>() + MAIL_OFF, total, sizeof(unsigned int),
                                hipMemcpyDeviceToHost, st));
    }
    AS_CHECK(hipMemcpyAsync(as.h_esum.p, as.esum.p, sizeof(double) * sel.n_scenes, hipMemcpyDeviceToHost, st));
    out.pending = true;
    return hipGetLastError();
}

hipError_t friction_adjoint(Assembly& as, const DevScene& sc, const SceneSel& sel, const double* d_poses,
                            const double* d_poses_t0, double epsv_h, double dhat, double dmin, double kappa,
                            const double* d_lambda, double* out, hipStream_t st) {
    const int nb = sc.nb;
    for (int i = 0; i < 6 * nb; i++) out[i] = 0.0;
    const long long n = as.n_fric;
    if (n == 0) return hipSuccess;
    AS_CHECK(world_vertices(as, sc, sel, d_poses_t0, as.W0, false, st));
    AS_CHECK(world_vertices(as, sc, sel, d_poses, as.W, true, st));
    AS_CHECK(as.frames0.ensure(sizeof(BodyFrame) * nb));
    hipLaunchKernelGGL(body_frame_kernel, dim3(grid_for(nb), 9), dim3(THREADS), 0, st, nb, d_poses_t0,
                       as.frames0.as<BodyFrame>());
    AS_CHECK(as.fadj.ensure(12 * sizeof(double) * n));
    hipLaunchKernelGGL(friction_adjoint_kernel, dim3(grid_for(n)), dim3(THREADS), 0, st, as.fric.as<FricRec>(), n,
                       sc.V, sc.v_body, as.W.as<double>(), as.W0.as<double>(), as.frames.as<BodyFrame>(),
                       as.frames0.as<BodyFrame>(), epsv_h, dhat, dmin, kappa, d_lambda, as.fadj.as<double>());
    AS_CHECK(hipGetLastError());
    std::vector<double> loc(12 * (size_t)n);
    std::vector<FricRec> recs((size_t)n);
    AS_CHECK(hipMemcpyAsync(loc.data(), as.fadj.p, 12 * sizeof(double) * n, hipMemcpyDeviceToHost, st));
    AS_CHECK(hipMemcpyAsync(recs.data(), as.fric.p, sizeof(FricRec) * n, hipMemcpyDeviceToHost, st));
    AS_CHECK(hipStreamSynchronize(st));
    for (long long c = 0; c < n; c++)
        for (int side = 0; side < 2; side++) {
            const int b = recs[(size_t)c].c.body[side];
            for (int k = 0; k < 6; k++) out[6 * b + k] += loc[12 * (size_t)c + 6 * side + k];
        }
    return hipSuccess;
}

void finish_potential(Assembly& as, AssemblyOut& out) {
    if (!out.pending) return;
    std::memcpy(out.energy.data(), as.h_esum.p, sizeof(double) * out.energy.size());
    if (out.hess) {
        out.n_off = as.h_mail.as<unsigned int>()[MAIL_OFF];
        as.n_off = out.n_off;
    }
    out.pending = false;
}

}  // namespace rigc
