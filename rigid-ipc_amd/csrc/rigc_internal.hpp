// Internal declarations shared by the HIP translation units of librigc.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/rigc.h"

namespace rigc {

// Device-resident scene (body-local geometry, global element ids).
struct DevScene {
    int nb = 0, nv = 0, ne = 0, nf = 0;
    double* V = nullptr;        // [nv][3] body-local vertices
    int* E = nullptr;           // [ne][2]
    int* F = nullptr;           // [nf][3]
    int* FE = nullptr;          // [nf][3]
    int* v_body = nullptr;      // [nv]
    int* e_body = nullptr;      // [ne]
    int* f_body = nullptr;      // [nf]
    int* body_vstart = nullptr; // [nb+1]
    int* body_estart = nullptr; // [nb+1]
    int* body_fstart = nullptr; // [nb+1]
    int* group = nullptr;       // [nb]
    uint8_t* is_static = nullptr;
    double* edge_len = nullptr; // [ne] body-local edge length
    double* r_max = nullptr;    // [nb]
};

// Grow-only device allocation.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n) {
        if (n <= bytes) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        const size_t want = n < 256 ? 256 : n;
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) bytes = want;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T> T* as() const { return static_cast<T*>(p); }
};

// Broad-phase working set (device buffers reused across calls).
struct BroadPhase {
    DevBuf box_raw, box_inf, inc, dom, vmin, vmax, bhull;
    DevBuf vc, vh, ec, eh, fc, fh;
    DevBuf pairs, keys_ee, keys_fv, tmp, hist, counters;
    DevBuf out_ee, out_fv;  // sorted (a, b) int32 pairs
    long long n_ee = 0, n_fv = 0, n_body_pairs = 0;
};

// Candidate detection (continuous when p1 != p0 is meant: `continuous`),
// leaving sorted int32 pairs in bp.out_ee / bp.out_fv.
hipError_t run_broad_phase(BroadPhase& bp, const DevScene& sc, const double* d_p0, const double* d_p1,
                           double inflation_radius, bool continuous, hipStream_t stream);

hipError_t radix_sort_u64(unsigned long long* keys, unsigned long long* tmp, unsigned int* hist, long long n,
                          int bits, hipStream_t stream);
size_t radix_sort_hist_entries(long long n);

// Narrow-phase candidate: kind 0 = edge-edge (a = edge0, b = edge1),
// kind 1 = face-vertex (a = face, b = vertex).
struct CcdArgs {
    const double* V;
    const int* E;
    const int* F;
    const int* FE;
    const int* v_body;
    const int* e_body;
    const int* f_body;
    const double* edge_len;
    const double* poses_t0;  // [nb][6]
    const double* poses_t1;
    const int* ee;           // [n_ee][2]
    const int* fv;           // [n_fv][2]
    long long n_ee, n_fv;
    double toi_tol;
    long long eval_budget;   // 0 = unbounded
    int use_global_best;     // prune with the running minimum over all queries
    unsigned long long* global_best;  // bits of a non-negative double
    unsigned long long* counter;      // work counter (query index)
    unsigned long long* total_evals;
    double* toi_out;         // per query (optional)
    signed char* hit_out;    // per query (optional): 1 hit, 0 none, -1 budget, -2 depth
    // per-query workspace, n_ee + n_fv entries each
    unsigned long long* q_best;  // least accepted t.lower (bits), +inf if none
    unsigned int* q_visits;      // box visits (only counted when eval_budget > 0)
    int* q_flag;                 // 0 ok, 1 budget exceeded, 2 depth exceeded
};

void launch_ccd(const CcdArgs& a, int grid, int block, hipStream_t stream);

}  // namespace rigc
