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
    // A context may hold several independent scenes concatenated (bodies,
    // vertices, edges and faces of scene s follow those of scene s - 1):
    // no pair of bodies of different scenes is ever a candidate, and every
    // reduction (domain, earliest TOI, energies, minimum distance) is per
    // scene.
    int n_scenes = 1;
    int* body_scene = nullptr;      // [nb]
    int* scene_body = nullptr;      // [n_scenes + 1] body ranges
    std::vector<int> h_scene_body;  // host copy of scene_body
    std::vector<int> h_scene_edge;  // [n_scenes + 1] edge ranges (host)
    // BVH broad phase (MeshSelector, src/utils/mesh_selector.cpp:47-120):
    // the lowest face index holding each vertex / edge; bvh_ok: every
    // vertex on an edge and every edge on a face (no codimensional elements)
    int* v2f = nullptr;             // [nv]
    int* e2f = nullptr;             // [ne]
    bool bvh_ok = false;
    int max_body_vertices = 0, max_body_faces = 0;
};

// Scenes taking part in one call: their ids (ascending), their bodies, and
// the 64 x 64 body tiles of each scene's upper-triangular body-pair grid
// (x = first body of the row tile, y = first body of the column tile,
// z = end of the scene's body range; y >= x).
struct SceneSel {
    int n_scenes = 0;
    const int* scenes = nullptr;
    int n_bodies = 0;
    const int* bodies = nullptr;
    long long n_tiles = 0;
    const int4* tiles = nullptr;
    std::vector<int> h_scenes;
    bool all = false;  // every scene of the context is selected
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

// Grow-only page-locked host allocation. Copies between it and the device
// are queued without blocking the host (a copy from pageable memory makes
// the runtime wait for it). Its contents may change or be read only once the
// stream has been synchronised after the last copy queued on it, and it is
// grown only then.
struct HostBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t n) {
        if (n <= bytes) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
        size_t want = 256;
        while (want < n) want *= 2;
        hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e == hipSuccess) bytes = want;
        return e;
    }
    void release() {
        if (p) (void)hipHostFree(p);
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
    DevBuf out_ee, out_fv;  // (a, b) int32 pairs
    HostBuf h_counters;     // element-pair counts read back
    long long n_ee = 0, n_fv = 0, n_body_pairs = 0;
    long long cap_pairs = 0, cap_ee = 0, cap_fv = 0;  // capacities of the last queued pass
    long long min_cap = 0;  // > 0: start every capacity from this floor (exercises the overflow passes)
};

// Body boxes of the selected bodies (RigidBody::compute_bounding_box over
// p0 -> p1, raw and grown by r, in bp.box_raw / bp.box_inf) with the pair
// counters cleared.
hipError_t queue_body_boxes(BroadPhase& bp, const DevScene& sc, const SceneSel& sel, const double* d_p0,
                            const double* d_p1, double r, hipStream_t stream);
// LDS bytes the BVH broad phase stages per workgroup for this scene (A's
// vertex boxes and two face lists); at most 64 KiB is supported.
inline size_t bvh_lds_bytes(const DevScene& sc) {
    return sizeof(double) * 6 * (size_t)sc.max_body_vertices + 2 * sizeof(int) * (size_t)sc.max_body_faces;
}
// DetectionMethod::BVH candidates (bvh.hip): close body pairs from the
// inflated body boxes, then per pair the owner-face rules of
// detect_body_pair_collision_candidates_from_aabbs with the smaller body's
// vertices carried into the other's frame with interval poses (over t in
// [0, 1] when continuous). EE keys keep the reference's (edge of the body
// with fewer faces, other edge) orientation. Output as run_broad_phase.
hipError_t run_broad_phase_bvh(BroadPhase& bp, const DevScene& sc, const SceneSel& sel, const double* d_p0,
                               const double* d_p1, double inflation_radius, bool continuous, bool sorted,
                               int num_cus, hipStream_t stream);

// Candidate detection over the selected scenes (continuous when p1 != p0
// is meant: `continuous`), leaving int32 pairs in bp.out_ee / bp.out_fv,
// sorted (the reference's order) when `sorted`.
hipError_t run_broad_phase(BroadPhase& bp, const DevScene& sc, const SceneSel& sel, const double* d_p0,
                           const double* d_p1, double inflation_radius, bool continuous, bool sorted,
                           hipStream_t stream);
// The same in parts, for callers that read the counts back with other
// results: queue_broad_phase queues every stage (packed keys e0 * ne + e1
// and f * nv + v in bp.keys_ee / keys_fv, counts [ee, fv, body pairs] in
// bp.counters) without waiting; after the read-back, broad_phase_fits
// checks the counts against the capacities (raising them to the counts);
// requeue_pairs runs the pair stages again at the raised capacities;
// finish_broad_phase unpacks (and sorts) the n_ee / n_fv keys.
hipError_t queue_broad_phase(BroadPhase& bp, const DevScene& sc, const SceneSel& sel, const double* d_p0,
                             const double* d_p1, double inflation_radius, bool continuous, hipStream_t stream);
bool broad_phase_fits(BroadPhase& bp, const unsigned long long* counts);
hipError_t requeue_pairs(BroadPhase& bp, const DevScene& sc, const SceneSel& sel, hipStream_t stream);
hipError_t finish_broad_phase(BroadPhase& bp, const DevScene& sc, bool sorted, hipStream_t stream);
// Close body pairs alone (discrete, inflation r) in bp.pairs, their count
// in bp.counters[2] (past bp.cap_pairs when the pass overflowed: raise
// bp.cap_pairs and requeue).
hipError_t queue_body_pairs(BroadPhase& bp, const DevScene& sc, const SceneSel& sel, const double* d_poses, double r,
                            hipStream_t stream);
hipError_t requeue_body_pairs(BroadPhase& bp, const DevScene& sc, const SceneSel& sel, hipStream_t stream);

// Edge-triangle intersections between close bodies (working set).
struct Intersections {
    DevBuf keys, unsure, counts, tmp, hist, out;  // keys edge * nf + face; out (edge, face) int32
    HostBuf h_counts;
    long long n = 0, n_unsure = 0;                // intersecting pairs; pairs the filter left open
    long long cap_keys = 0, cap_unsure = 0;
    void release() {
        for (DevBuf* b : { &keys, &unsure, &counts, &tmp, &hist, &out }) b->release();
        h_counts.release();
    }
};
struct Assembly;
// Every (edge of one body, face of another) pair of the close body pairs
// at d_poses (no inflation) whose closed segment and closed triangle meet,
// sorted by (edge, face), into ix.out ([ix.n][2]).
hipError_t detect_intersections(Assembly& as, BroadPhase& bp, Intersections& ix, const DevScene& sc,
                                const SceneSel& sel, const double* d_poses, hipStream_t stream);

hipError_t radix_sort_u64(unsigned long long* keys, unsigned long long* tmp, unsigned int* hist, long long n,
                          int bits, hipStream_t stream);
size_t radix_sort_hist_entries(long long n);

// A pending subtree of one query's bisection tree: the dyadic box
// (index, level per dimension t, alpha, beta) and the query's t tolerance
// after its start-box check.
struct CcdTask {
    long long qi;
    double tol0;
    unsigned long long idx[3];
    int lev;  // lev_t | lev_a << 8 | lev_b << 16
    int pad;
};

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
    int use_global_best;     // prune with the running minimum over the query's scene
    unsigned long long* global_best;  // [n_scenes] bits of non-negative doubles
    const int* body_scene;   // [nb]
    const unsigned char* scene_skip;  // optional [n_scenes]: queries of flagged scenes are skipped
    unsigned long long* counter;      // work counter (query index)
    unsigned long long* total_evals;
    double* toi_out;         // per query (optional)
    signed char* hit_out;    // per query (optional): 1 hit, 0 none, -1 budget, -2 depth
    // per-query workspace, n_ee + n_fv entries each
    unsigned long long* q_best;  // least accepted t.lower (bits), +inf if none
    unsigned int* q_visits;      // box visits (only counted when eval_budget > 0)
    int* q_flag;                 // 0 ok, 1 budget exceeded, 2 depth exceeded
    // pass structure (set by launch_ccd): the work pool is the queries
    // (tasks_in == nullptr) or the subtrees exported by the previous pass
    const CcdTask* tasks_in;
    long long n_tasks_in;
    CcdTask* tasks_out;
    unsigned long long* n_tasks_out;
    long long tasks_cap;
    int walk_budget;             // evaluations after which a lane exports its open subtrees
    unsigned long long* scene_evals;  // optional [n_scenes]: box evaluations per scene
    int spec_levels;             // speculative levels below an owner's box (0..5; 0 = none)
};

// Task buffers of the multi-pass narrow phase (grow-only).
struct CcdWork {
    DevBuf tasks[2], count;
    HostBuf h_count;         // export count read back after each pass
    int passes = 0;          // traversal launches of the last launch_ccd
    float kernel_ms = 0.0f;  // their summed duration (events around each launch)
    hipEvent_t k0 = nullptr, k1 = nullptr;
    void release() {
        tasks[0].release();
        tasks[1].release();
        count.release();
        h_count.release();
        if (k0) (void)hipEventDestroy(k0);
        if (k1) (void)hipEventDestroy(k1);
        k0 = k1 = nullptr;
    }
};

// Narrow phase: a first pass over the queries, then passes over exported
// subtrees until none is left (a host read of the export count between
// passes). a.counter is reset before every pass.
// Device-to-host copies a caller wants made when the narrow phase ends; they
// are queued after every pass, ahead of that pass's synchronisation.
struct CcdReadback {
    void* dst;
    const void* src;
    size_t bytes;
};
// Resident 256-lane blocks of the persistent CCD kernel per CU: the kernel
// runs one wavefront per SIMD so that its register demand (above 256 per
// lane) lives in VGPRs and AGPRs instead of scratch.
// The CCD kernel fits two 256-thread blocks per CU (two wavefronts per SIMD).
// A call with few queries is bound by its longest bisection, whose
// wavefront runs faster alone on its SIMD, so it launches one block per CU;
// a call with many queries is bound by throughput and launches two.
constexpr int CCD_BLOCKS_PER_CU = 2;
inline int ccd_grid(int num_cus, long long n_queries, long long wide_queries) {
    return num_cus * (n_queries >= wide_queries ? CCD_BLOCKS_PER_CU : 1);
}
hipError_t launch_ccd(CcdArgs a, int grid, int block, hipStream_t stream, CcdWork& work,
                      const CcdReadback* readback = nullptr, int n_readback = 0);
// The narrow phase in one pass (no export of long walks, so no task
// buffers): a.n_tasks_out must point to one device counter, a.tasks_out is
// ignored. Same results as launch_ccd.
hipError_t launch_ccd_single(CcdArgs a, int grid, int block, hipStream_t stream);
// The narrow phase of rigc_narrow_phase_enqueue on `stream`, kernels only.
hipError_t enqueue_narrow_phase(CcdArgs a, int grid, unsigned long long* scene_best, int n_scenes, double* out,
                                hipStream_t stream);

// ---- the narrow phase without allocation (noalloc.hip)
// Size of the static device array that holds every per-call buffer.
constexpr size_t NOALLOC_SCRATCH_BYTES = RIGC_NOALLOC_SCRATCH_BYTES;
// Byte offsets of one call's buffers in that array (256-aligned).
struct NoallocLayout {
    size_t poses0, poses1, ee, fv, toi, hit, counters, qbest, qvisits, qflag, scene_best, result, total;
};
// Device address of the array on `device` (the current device), looked up
// once per device.
hipError_t noalloc_scratch(int device, void** ptr);
NoallocLayout noalloc_layout(long long nb, long long n_ee, long long n_fv, int n_scenes);
// Sets the per-scene minima, runs the narrow phase over a (whose buffers
// point into base) and writes the result words (least TOI bits, unresolved
// flag) at base + L.result.
hipError_t noalloc_narrow_phase(CcdArgs a, int grid, unsigned char* base, const NoallocLayout& L, int n_scenes,
                                int per_query, hipStream_t stream);

// Exclusive prefix sum of n counts; work holds scan_workspace_entries(n)
// entries and *total (device) receives the sum.
size_t scan_workspace_entries(long long n);
void exclusive_scan_u32(const unsigned int* in, unsigned int* out, long long n, unsigned int* work,
                        unsigned int* total, hipStream_t stream);

// One collision constraint (ipc-toolkit vertex order):
//   VV (v0, v1), EV (v, e0, e1), EE (ea0, ea1, eb0, eb1), FV (v, f0, f1, f2)
// a, b: VV (v0, v1); EV (edge, vertex); EE (edge0, edge1); FV (face, vertex).
struct ConRec {
    int v[4];
    int type;
    int mult;
    int body[2];
    int a, b;
    double eps_x;  // EE mollifier threshold (-1 otherwise)
};

// Lagged friction constraint: the collision constraint plus its closest
// point, tangent basis (3x2 row-major), normal force and coefficient.
struct FricRec {
    ConRec c;
    double coords[2];
    double basis[6];
    double normal_force;
    double mu;
};

// Record ranges of up to four classes: class k is [off[k], off[k + 1]).
struct ClassRanges {
    long long off[5];
    int n_class;
};

// Assembly working set (device buffers reused across calls).
struct Assembly {
    DevBuf R, frames, W, W0;            // body rotations / frames, world vertices
    DevBuf frames0, fadj;               // friction adjoint: frames at the step start, per-record output
    DevBuf ckeys[4], cstart[2], ccount, ctmp, chist, cflags, cpos, cwork;
    DevBuf cons;                        // ConRec [n_con], classes VV, EV, EE, FV in turn
    long long n_con = 0, counts[4] = { 0, 0, 0, 0 };
    bool have_cons = false;
    // Lagged friction sets of all scenes, scene after scene: scene s owns
    // [fric_off[s], fric_off[s + 1]) when fric_valid[s].
    DevBuf fric, fric_tmp, fric_new, fric_stage, ranges, entries;
    long long n_fric = 0;
    bool have_fric = false;             // every scene has a friction set
    std::vector<long long> fric_off;
    std::vector<char> fric_valid;
    DevBuf energy, lgrad, lhess, esum;  // per-constraint outputs and the energy sum
    DevBuf dkeys, okeys, ktmp, khist, oflags, opos, owork;
    DevBuf diag, grad, opairs, oblocks, mind;
    long long n_off = 0;
    // Content tags of the pose arrays world vertices are computed from
    // (set by the caller: equal non-zero tags mean equal contents, 0 means
    // unknown), and the tags W, W0 and frames were last computed at over
    // every body (0: not reusable).
    const double* tag_ptr[3] = { nullptr, nullptr, nullptr };
    unsigned long long tag[3] = { 0, 0, 0 };
    unsigned long long W_tag = 0, W0_tag = 0, F_tag = 0;
    // page-locked staging: counts read back (h_mail), the range-copy table,
    // per-scene class ranges, energy sums and minimum distances
    HostBuf h_mail, h_entries, h_ranges, h_esum, h_mind;

    void release() {
        for (DevBuf* b : { &R, &frames, &W, &W0, &ckeys[0], &ckeys[1], &ckeys[2], &ckeys[3], &cstart[0], &cstart[1],
                           &ccount, &ctmp, &chist, &cflags, &cpos, &cwork, &cons, &fric, &fric_tmp, &fric_new,
                           &fric_stage, &ranges, &entries, &energy, &lgrad, &lhess, &esum, &dkeys, &okeys, &ktmp,
                           &khist, &oflags, &opos, &owork, &diag, &grad, &opairs, &oblocks, &mind })
            b->release();
        for (HostBuf* b : { &h_mail, &h_entries, &h_ranges, &h_esum, &h_mind }) b->release();
    }
};

// Slots of Assembly::h_mail (unsigned int).
enum { MAIL_CLASS = 0, MAIL_RUNS = 4, MAIL_OFF = 6, MAIL_SLOTS = 8 };

// Result of assemble_potential. Its read-backs are queued, not waited for:
// once the stream has been synchronised, finish_potential fills energy and
// n_off.
struct AssemblyOut {
    std::vector<double> energy;  // per selected scene
    long long n_off = 0;         // off-diagonal blocks produced
    long long n_rec = 0;         // records assembled (n_off <= n_rec)
    bool pending = false;
    bool hess = false;
};

// Constraint set of the selected scenes at poses, kept in as.cons.
// World vertices at d_poses into as.W ([nv][3]).
hipError_t current_world_vertices(Assembly& as, const DevScene& sc, const SceneSel& sel, const double* d_poses,
                                  hipStream_t stream);
hipError_t build_constraint_set(Assembly& as, BroadPhase& bp, const DevScene& sc, const SceneSel& sel,
                                const double* d_poses, double dhat, double dmin, int num_cus, hipStream_t stream);
// Constraints per selected scene and class ([n_sel][4]) of the current set.
hipError_t constraint_counts(Assembly& as, const DevScene& sc, const SceneSel& sel, std::vector<long long>& counts,
                             hipStream_t stream);
// Barrier (friction == false, over the current constraint set) or friction
// potential (over the selected scenes' friction sets). Outputs stay on the
// device: as.grad [nb][6], as.diag [nb][36] (rows of the selected bodies),
// as.opairs / as.oblocks [n_off]; energies per selected scene. Nothing is
// waited for: synchronise the stream, then call finish_potential.
hipError_t assemble_potential(Assembly& as, const DevScene& sc, const SceneSel& sel, bool friction,
                              const double* d_poses, const double* d_poses_t0, double dhat, double dmin,
                              double epsv_h, const double* d_epsv_h_scene, bool want_grad, bool want_hess,
                              int num_cus, AssemblyOut& out, hipStream_t stream, bool project = true);
void finish_potential(Assembly& as, AssemblyOut& out);
// Friction adjoint over the whole friction set (single-scene contexts):
// per record, the 12 body dofs of (d grad_x D / d x0)^T lambda; summed into
// out [nb][6] on the host in record order (deterministic).
hipError_t friction_adjoint(Assembly& as, const DevScene& sc, const SceneSel& sel, const double* d_poses,
                            const double* d_poses_t0, double epsv_h, double dhat, double dmin, double kappa,
                            const double* d_lambda, double* out, hipStream_t stream);
// Friction sets of the selected scenes from the current constraint set
// (which must be the selected scenes' set at d_poses); kappa / mu per scene
// (device arrays indexed by scene id).
hipError_t build_friction_set(Assembly& as, const DevScene& sc, const SceneSel& sel, const double* d_poses,
                              double dhat, double dmin, const double* d_kappa_scene, const double* d_mu_scene,
                              hipStream_t stream);
// Minimum squared distance over the current set, per selected scene.
hipError_t minimum_distance(Assembly& as, const DevScene& sc, const SceneSel& sel, const double* d_poses,
                            std::vector<double>& min_sq, hipStream_t stream);

// Row gathers / scatters between request order and scene order:
// dst[i][:] = src[rows[i]][:] (gather) or dst[rows[i]][:] = src[i][:].
void gather_rows(const int* rows, int n, int width, const double* src, double* dst, hipStream_t stream);
void scatter_rows(const int* rows, int n, int width, const double* src, double* dst, hipStream_t stream);

// Piecewise-linear CCD (plccd.hip): one lane per query, EE queries first.
// Each lane owns arena[lane * (tstack_cap + 7 * heap_cap)] doubles.
// hit_out: 1 impact (toi_out = TOI), 0 none (+inf), 2 arena slice too small.
struct PlccdArgs {
    const double* V;
    const int* E;
    const int* F;
    const int* v_body;
    const int* e_body;
    const int* f_body;
    const double* poses_t0;
    const double* poses_t1;
    const int* ee;               // [n_ee][2] edge pairs
    const int* fv;               // [n_fv][2] (face, vertex)
    long long n_ee;
    const int* query_ids;        // queries to run (nullptr: 0 .. n_queries-1)
    long long n_queries;
    double earliest_toi;
    double msd;                  // minimum separation distance
    double* toi_out;             // [n_ee + n_fv]
    signed char* hit_out;        // [n_ee + n_fv]
    double* arena;
    long long n_lanes;           // lanes that run queries (one arena slice each)
    int tstack_cap;
    long long heap_cap;
    unsigned long long* counters;  // [0] boxes visited, [1] trajectory pieces
};
hipError_t launch_plccd(const PlccdArgs& a, int grid, hipStream_t stream);
// One query per 64-lane block, stack (64) and heap (512 entries) in LDS; n_lanes unused.
hipError_t launch_plccd_wave(const PlccdArgs& a, int grid, hipStream_t stream);

}  // namespace rigc
