// C ABI of librigc: context, scene upload and the entry points declared in
// include/rigc.h. Host orchestration only; the kernels live in the other
// translation units.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "rigc_internal.hpp"

using namespace rigc;


struct rigc_ctx {
    int device = 0;
    int num_cus = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    rigc_stats stats{};

    bool has_scene = false;
    DevScene scene;
    std::vector<DevBuf> scene_bufs;
    double average_edge_length = 0;

    // per-call buffers
    DevBuf poses0, poses1, ee, fv, toi, hit, counters, qbest, qvisits, qflag;
    BroadPhase bp;
    Assembly as;
    CcdWork ccdw;
    int walk_budget = 4096;
    double last_toi = HUGE_VAL;  // answer of the previous rigc_compute_earliest_toi
    hipEvent_t ev2 = nullptr;

    rigc_status fail(rigc_status s, const std::string& msg) {
        err = msg;
        return s;
    }
    rigc_status hip_fail(hipError_t e, const char* where) {
        err = std::string(where) + ": " + hipGetErrorString(e);
        return RIGC_ERR_DEVICE;
    }
};

#define RIGC_HIP(call)                                   \
    do {                                                 \
        hipError_t e_ = (call);                          \
        if (e_ != hipSuccess) return ctx->hip_fail(e_, #call); \
    } while (0)

namespace {

template <typename T>
hipError_t upload(rigc_ctx* ctx, T** dst, const T* src, size_t n) {
    ctx->scene_bufs.emplace_back();
    DevBuf& b = ctx->scene_bufs.back();
    hipError_t e = b.ensure(std::max<size_t>(n, 1) * sizeof(T));
    if (e != hipSuccess) return e;
    *dst = b.as<T>();
    if (n) e = hipMemcpyAsync(*dst, src, n * sizeof(T), hipMemcpyHostToDevice, ctx->stream);
    return e;
}

void free_scene(rigc_ctx* ctx) {
    for (auto& b : ctx->scene_bufs) b.release();
    ctx->scene_bufs.clear();
    ctx->scene = DevScene{};
    ctx->has_scene = false;
}

}  // namespace

extern "C" {

rigc_status rigc_create(rigc_ctx** out, int32_t device) {
    if (!out) return RIGC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) return RIGC_ERR_DEVICE;
    rigc_ctx* ctx = new rigc_ctx();
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev0);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev1);
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev2);
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        delete ctx;
        return RIGC_ERR_DEVICE;
    }
    ctx->num_cus = prop.multiProcessorCount;
    // RIGC_CCD_WALK_BUDGET: box evaluations after which a lane hands its
    // pending subtrees to the next narrow-phase pass (0 = never)
    if (const char* wb = std::getenv("RIGC_CCD_WALK_BUDGET")) ctx->walk_budget = std::atoi(wb);
    *out = ctx;
    return RIGC_OK;
}

rigc_status rigc_destroy(rigc_ctx* ctx) {
    if (!ctx) return RIGC_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_scene(ctx);
    for (DevBuf* b : { &ctx->poses0, &ctx->poses1, &ctx->ee, &ctx->fv, &ctx->toi, &ctx->hit, &ctx->counters,
                       &ctx->qbest, &ctx->qvisits, &ctx->qflag })
        b->release();
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->ev2) (void)hipEventDestroy(ctx->ev2);
    for (DevBuf* b : { &ctx->bp.box_raw, &ctx->bp.box_inf, &ctx->bp.inc, &ctx->bp.dom, &ctx->bp.vmin, &ctx->bp.vmax,
                       &ctx->bp.bhull, &ctx->bp.vc, &ctx->bp.vh, &ctx->bp.ec, &ctx->bp.eh, &ctx->bp.fc, &ctx->bp.fh,
                       &ctx->bp.pairs, &ctx->bp.keys_ee, &ctx->bp.keys_fv, &ctx->bp.tmp, &ctx->bp.hist,
                       &ctx->bp.counters, &ctx->bp.out_ee, &ctx->bp.out_fv })
        b->release();
    ctx->as.release();
    ctx->ccdw.release();
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return RIGC_OK;
}

const char* rigc_last_error(const rigc_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

rigc_status rigc_get_stats(const rigc_ctx* ctx, rigc_stats* stats) {
    if (!ctx || !stats) return RIGC_ERR_INVALID_ARGUMENT;
    *stats = ctx->stats;
    return RIGC_OK;
}

rigc_status rigc_set_scene(rigc_ctx* ctx, const rigc_scene* s) {
    if (!ctx || !s) return RIGC_ERR_INVALID_ARGUMENT;
    if (s->num_bodies < 1 || s->num_vertices < 0 || s->num_edges < 0 || s->num_faces < 0)
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_set_scene: negative sizes");
    if (!s->vertices || !s->body_vertex_start || !s->body_edge_start || !s->body_face_start || !s->group_ids
        || !s->is_static || (s->num_edges && !s->edges) || (s->num_faces && (!s->faces || !s->face_edges)))
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_set_scene: missing arrays");
    const int nb = s->num_bodies, nv = s->num_vertices, ne = s->num_edges, nf = s->num_faces;
    if (s->body_vertex_start[0] != 0 || s->body_vertex_start[nb] != nv || s->body_edge_start[0] != 0
        || s->body_edge_start[nb] != ne || s->body_face_start[0] != 0 || s->body_face_start[nb] != nf)
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_set_scene: body ranges do not cover the arrays");
    for (int i = 0; i < 2 * ne; i++)
        if (s->edges[i] < 0 || s->edges[i] >= nv)
            return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_set_scene: edge vertex out of range");
    for (int i = 0; i < 3 * nf; i++)
        if (s->faces[i] < 0 || s->faces[i] >= nv || s->face_edges[i] < 0 || s->face_edges[i] >= ne)
            return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_set_scene: face index out of range");

    // Host-side derived data, following RigidBody's constructor
    // (src/physics/rigid_body.cpp:124-142) and RigidBody::edge_length
    // (src/physics/rigid_body.hpp:151-156).
    std::vector<int> v_body(nv), e_body(ne), f_body(nf);
    std::vector<double> r_max(nb, 0.0), elen(ne);
    double avg_sum = 0;
    for (int b = 0; b < nb; b++) {
        for (int i = s->body_vertex_start[b]; i < s->body_vertex_start[b + 1]; i++) {
            v_body[i] = b;
            const double* v = s->vertices + 3 * i;
            r_max[b] = std::max(r_max[b], std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]));
        }
        double body_avg = 0;
        const int e0 = s->body_edge_start[b], e1 = s->body_edge_start[b + 1];
        for (int i = e0; i < e1; i++) {
            e_body[i] = b;
            const double* a = s->vertices + 3 * s->edges[2 * i];
            const double* c = s->vertices + 3 * s->edges[2 * i + 1];
            const double d0 = c[0] - a[0], d1 = c[1] - a[1], d2 = c[2] - a[2];
            elen[i] = std::sqrt((d0 * d0 + d1 * d1) + d2 * d2);
            body_avg += elen[i];
        }
        if (e1 > e0) body_avg /= (e1 - e0);
        avg_sum += (e1 - e0) * body_avg;
        for (int i = s->body_face_start[b]; i < s->body_face_start[b + 1]; i++) f_body[i] = b;
    }
    ctx->average_edge_length = ne ? avg_sum / ne : 0.0;

    RIGC_HIP(hipSetDevice(ctx->device));
    free_scene(ctx);
    DevScene& d = ctx->scene;
    d.nb = nb; d.nv = nv; d.ne = ne; d.nf = nf;
    RIGC_HIP(upload(ctx, &d.V, s->vertices, 3 * (size_t)nv));
    RIGC_HIP(upload(ctx, &d.E, s->edges, 2 * (size_t)ne));
    RIGC_HIP(upload(ctx, &d.F, s->faces, 3 * (size_t)nf));
    RIGC_HIP(upload(ctx, &d.FE, s->face_edges, 3 * (size_t)nf));
    RIGC_HIP(upload(ctx, &d.v_body, v_body.data(), (size_t)nv));
    RIGC_HIP(upload(ctx, &d.e_body, e_body.data(), (size_t)ne));
    RIGC_HIP(upload(ctx, &d.f_body, f_body.data(), (size_t)nf));
    RIGC_HIP(upload(ctx, &d.body_vstart, s->body_vertex_start, (size_t)nb + 1));
    RIGC_HIP(upload(ctx, &d.body_estart, s->body_edge_start, (size_t)nb + 1));
    RIGC_HIP(upload(ctx, &d.body_fstart, s->body_face_start, (size_t)nb + 1));
    RIGC_HIP(upload(ctx, &d.group, s->group_ids, (size_t)nb));
    RIGC_HIP(upload(ctx, &d.is_static, s->is_static, (size_t)nb));
    RIGC_HIP(upload(ctx, &d.edge_len, elen.data(), (size_t)ne));
    RIGC_HIP(upload(ctx, &d.r_max, r_max.data(), (size_t)nb));
    RIGC_HIP(hipStreamSynchronize(ctx->stream));
    ctx->as.have_cons = ctx->as.have_fric = false;
    ctx->as.n_con = ctx->as.n_fric = 0;
    ctx->last_toi = HUGE_VAL;
    ctx->has_scene = true;
    return RIGC_OK;
}

rigc_status rigc_narrow_phase(rigc_ctx* ctx, const double* poses_t0, const double* poses_t1,
                              const int32_t* ee, int64_t n_ee, const int32_t* fv, int64_t n_fv,
                              double toi_tolerance, int64_t eval_budget, int32_t per_query,
                              double* earliest_toi, double* toi_out, int8_t* hit_out) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_narrow_phase: no scene");
    if (!poses_t0 || !poses_t1 || !earliest_toi || n_ee < 0 || n_fv < 0 || (n_ee && !ee) || (n_fv && !fv)
        || !(toi_tolerance > 0))
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_narrow_phase: bad arguments");
    const DevScene& d = ctx->scene;
    for (int64_t i = 0; i < n_ee; i++)
        if (ee[2 * i] < 0 || ee[2 * i] >= d.ne || ee[2 * i + 1] < 0 || ee[2 * i + 1] >= d.ne)
            return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_narrow_phase: edge id out of range");
    for (int64_t i = 0; i < n_fv; i++)
        if (fv[2 * i] < 0 || fv[2 * i] >= d.nf || fv[2 * i + 1] < 0 || fv[2 * i + 1] >= d.nv)
            return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_narrow_phase: face/vertex id out of range");

    RIGC_HIP(hipSetDevice(ctx->device));
    const size_t nq = (size_t)(n_ee + n_fv);
    const size_t pose_bytes = 6 * sizeof(double) * (size_t)d.nb;
    RIGC_HIP(ctx->poses0.ensure(pose_bytes));
    RIGC_HIP(ctx->poses1.ensure(pose_bytes));
    RIGC_HIP(ctx->ee.ensure(2 * sizeof(int32_t) * std::max<size_t>(n_ee, 1)));
    RIGC_HIP(ctx->fv.ensure(2 * sizeof(int32_t) * std::max<size_t>(n_fv, 1)));
    RIGC_HIP(ctx->toi.ensure(sizeof(double) * std::max<size_t>(nq, 1)));
    RIGC_HIP(ctx->hit.ensure(std::max<size_t>(nq, 1)));
    RIGC_HIP(ctx->counters.ensure(4 * sizeof(unsigned long long)));
    RIGC_HIP(ctx->qbest.ensure(sizeof(unsigned long long) * std::max<size_t>(nq, 1)));
    RIGC_HIP(ctx->qvisits.ensure(sizeof(unsigned int) * std::max<size_t>(nq, 1)));
    RIGC_HIP(ctx->qflag.ensure(sizeof(int) * std::max<size_t>(nq, 1)));
    hipStream_t st = ctx->stream;
    RIGC_HIP(hipMemcpyAsync(ctx->poses0.p, poses_t0, pose_bytes, hipMemcpyHostToDevice, st));
    RIGC_HIP(hipMemcpyAsync(ctx->poses1.p, poses_t1, pose_bytes, hipMemcpyHostToDevice, st));
    if (n_ee) RIGC_HIP(hipMemcpyAsync(ctx->ee.p, ee, 2 * sizeof(int32_t) * n_ee, hipMemcpyHostToDevice, st));
    if (n_fv) RIGC_HIP(hipMemcpyAsync(ctx->fv.p, fv, 2 * sizeof(int32_t) * n_fv, hipMemcpyHostToDevice, st));
    unsigned long long init[4] = { 0x7ff0000000000000ULL, 0, 0, 0 };
    RIGC_HIP(hipMemcpyAsync(ctx->counters.p, init, sizeof(init), hipMemcpyHostToDevice, st));

    CcdArgs a{};
    a.V = d.V; a.E = d.E; a.F = d.F; a.FE = d.FE;
    a.v_body = d.v_body; a.e_body = d.e_body; a.f_body = d.f_body; a.edge_len = d.edge_len;
    a.poses_t0 = ctx->poses0.as<double>();
    a.poses_t1 = ctx->poses1.as<double>();
    a.ee = ctx->ee.as<int>();
    a.fv = ctx->fv.as<int>();
    a.n_ee = n_ee;
    a.n_fv = n_fv;
    a.toi_tol = toi_tolerance;
    a.eval_budget = eval_budget;
    a.use_global_best = per_query ? 0 : 1;
    a.global_best = ctx->counters.as<unsigned long long>();
    a.counter = ctx->counters.as<unsigned long long>() + 1;
    a.total_evals = ctx->counters.as<unsigned long long>() + 2;
    a.toi_out = ctx->toi.as<double>();
    a.hit_out = ctx->hit.as<signed char>();
    a.q_best = ctx->qbest.as<unsigned long long>();
    a.q_visits = ctx->qvisits.as<unsigned int>();
    a.q_flag = ctx->qflag.as<int>();

    const int block = 256;
    const int grid = ctx->num_cus * 2;  // persistent: two 256-lane blocks per CU
    RIGC_HIP(hipEventRecord(ctx->ev0, st));
    a.walk_budget = ctx->walk_budget;
    if (nq) RIGC_HIP(launch_ccd(a, grid, block, st, ctx->ccdw));
    RIGC_HIP(hipEventRecord(ctx->ev1, st));

    std::vector<double> tois(nq);
    std::vector<signed char> hits(nq);
    if (nq) {
        RIGC_HIP(hipMemcpyAsync(tois.data(), ctx->toi.p, sizeof(double) * nq, hipMemcpyDeviceToHost, st));
        RIGC_HIP(hipMemcpyAsync(hits.data(), ctx->hit.p, nq, hipMemcpyDeviceToHost, st));
    }
    unsigned long long res[4];
    RIGC_HIP(hipMemcpyAsync(res, ctx->counters.p, sizeof(res), hipMemcpyDeviceToHost, st));
    RIGC_HIP(hipStreamSynchronize(st));
    float ms = 0;
    RIGC_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.narrow_phase_ms = ms;
    ctx->stats.box_evaluations = (int64_t)res[2];
    ctx->stats.num_candidates = (int64_t)nq;
    ctx->stats.ccd_kernel_ms = nq ? ctx->ccdw.kernel_ms : 0.0;
    ctx->stats.ccd_launches = nq ? ctx->ccdw.passes : 0;

    bool unresolved = false;
    double best = INFINITY;
    for (size_t i = 0; i < nq; i++) {
        if (hits[i] < 0) unresolved = true;
        if (hits[i] == 1) best = std::min(best, tois[i]);
    }
    if (per_query) {
        *earliest_toi = best;
    } else {
        double g;
        std::memcpy(&g, &res[0], sizeof(double));
        *earliest_toi = g;
    }
    if (toi_out) std::memcpy(toi_out, tois.data(), sizeof(double) * nq);
    if (hit_out) std::memcpy(hit_out, hits.data(), nq);
    if (unresolved)
        return ctx->fail(RIGC_ERR_UNRESOLVED, "rigc_narrow_phase: a query exceeded its evaluation budget");
    return RIGC_OK;
}

namespace {

rigc_status upload_poses(rigc_ctx* ctx, const double* p0, const double* p1) {
    const size_t pose_bytes = 6 * sizeof(double) * (size_t)ctx->scene.nb;
    RIGC_HIP(ctx->poses0.ensure(pose_bytes));
    RIGC_HIP(ctx->poses1.ensure(pose_bytes));
    RIGC_HIP(hipMemcpyAsync(ctx->poses0.p, p0, pose_bytes, hipMemcpyHostToDevice, ctx->stream));
    if (p1) RIGC_HIP(hipMemcpyAsync(ctx->poses1.p, p1, pose_bytes, hipMemcpyHostToDevice, ctx->stream));
    return RIGC_OK;
}

CcdArgs ccd_args(rigc_ctx* ctx) {
    const DevScene& d = ctx->scene;
    CcdArgs a{};
    a.V = d.V; a.E = d.E; a.F = d.F; a.FE = d.FE;
    a.v_body = d.v_body; a.e_body = d.e_body; a.f_body = d.f_body; a.edge_len = d.edge_len;
    a.poses_t0 = ctx->poses0.as<double>();
    a.poses_t1 = ctx->poses1.as<double>();
    a.global_best = ctx->counters.as<unsigned long long>();
    a.counter = ctx->counters.as<unsigned long long>() + 1;
    a.total_evals = ctx->counters.as<unsigned long long>() + 2;
    return a;
}

}  // namespace

rigc_status rigc_broad_phase(rigc_ctx* ctx, const double* poses_t0, const double* poses_t1,
                             double inflation_radius, int32_t* ee, int64_t ee_capacity, int64_t* n_ee,
                             int32_t* fv, int64_t fv_capacity, int64_t* n_fv) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_broad_phase: no scene");
    if (!poses_t0 || !n_ee || !n_fv || ee_capacity < 0 || fv_capacity < 0 || !(inflation_radius >= 0)
        || (ee_capacity && !ee) || (fv_capacity && !fv))
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_broad_phase: bad arguments");
    RIGC_HIP(hipSetDevice(ctx->device));
    rigc_status st = upload_poses(ctx, poses_t0, poses_t1);
    if (st != RIGC_OK) return st;
    RIGC_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    RIGC_HIP(run_broad_phase(ctx->bp, ctx->scene, ctx->poses0.as<double>(), ctx->poses1.as<double>(),
                             inflation_radius, poses_t1 != nullptr, ctx->stream));
    RIGC_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    *n_ee = ctx->bp.n_ee;
    *n_fv = ctx->bp.n_fv;
    const bool fits = ctx->bp.n_ee <= ee_capacity && ctx->bp.n_fv <= fv_capacity;
    if (fits) {
        if (ctx->bp.n_ee)
            RIGC_HIP(hipMemcpyAsync(ee, ctx->bp.out_ee.p, 2 * sizeof(int32_t) * ctx->bp.n_ee, hipMemcpyDeviceToHost,
                                    ctx->stream));
        if (ctx->bp.n_fv)
            RIGC_HIP(hipMemcpyAsync(fv, ctx->bp.out_fv.p, 2 * sizeof(int32_t) * ctx->bp.n_fv, hipMemcpyDeviceToHost,
                                    ctx->stream));
    }
    RIGC_HIP(hipStreamSynchronize(ctx->stream));
    float ms = 0;
    RIGC_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.broad_phase_ms = ms;
    ctx->stats.num_candidates = ctx->bp.n_ee + ctx->bp.n_fv;
    if (!fits) return ctx->fail(RIGC_ERR_CAPACITY, "rigc_broad_phase: output capacity too small");
    return RIGC_OK;
}

rigc_status rigc_compute_earliest_toi(rigc_ctx* ctx, const double* poses_t0, const double* poses_t1,
                                      double minimum_separation_distance, double* earliest_toi) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_compute_earliest_toi: no scene");
    if (!poses_t0 || !poses_t1 || !earliest_toi || !(minimum_separation_distance >= 0))
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_compute_earliest_toi: bad arguments");
    RIGC_HIP(hipSetDevice(ctx->device));
    rigc_status st = upload_poses(ctx, poses_t0, poses_t1);
    if (st != RIGC_OK) return st;
    hipStream_t s = ctx->stream;
    RIGC_HIP(hipEventRecord(ctx->ev0, s));
    RIGC_HIP(run_broad_phase(ctx->bp, ctx->scene, ctx->poses0.as<double>(), ctx->poses1.as<double>(),
                             minimum_separation_distance / 2.0, true, s));
    RIGC_HIP(hipEventRecord(ctx->ev1, s));
    const long long nq = ctx->bp.n_ee + ctx->bp.n_fv;
    RIGC_HIP(ctx->counters.ensure(4 * sizeof(unsigned long long)));
    RIGC_HIP(ctx->qbest.ensure(sizeof(unsigned long long) * std::max<long long>(nq, 1)));
    RIGC_HIP(ctx->qvisits.ensure(sizeof(unsigned int) * std::max<long long>(nq, 1)));
    RIGC_HIP(ctx->qflag.ensure(sizeof(int) * std::max<long long>(nq, 1)));
    CcdArgs a = ccd_args(ctx);
    a.ee = ctx->bp.out_ee.as<int>();
    a.fv = ctx->bp.out_fv.as<int>();
    a.n_ee = ctx->bp.n_ee;
    a.n_fv = ctx->bp.n_fv;
    a.toi_tol = 1e-4;
    a.eval_budget = 0;
    a.use_global_best = 1;
    a.toi_out = nullptr;
    a.hit_out = nullptr;
    a.q_best = ctx->qbest.as<unsigned long long>();
    a.q_visits = ctx->qvisits.as<unsigned int>();
    a.q_flag = ctx->qflag.as<int>();
    const int grid = ctx->num_cus * 2;  // persistent: two 256-lane blocks per CU
    a.walk_budget = ctx->walk_budget;
    // Windowed search: a first pass starts the running minimum at a bound T
    // (four times the previous call's answer; successive line-search calls
    // return similar times). It prunes every box with t.lower >= T, so if
    // it accepts a box the minimum below T is the answer; otherwise the
    // full search runs. Either way the result is the exact minimum.
    const double inf = HUGE_VAL;
    double window = inf;
    if (ctx->last_toi < 0.25) window = std::max(4.0 * ctx->last_toi, std::ldexp(1.0, -20));
    double g = inf;
    long long evals = 0, launches = 0;
    double kernel_ms = 0.0;
    auto run = [&](double bound) -> hipError_t {
        unsigned long long init[4] = { 0, 0, 0, 0 };
        std::memcpy(&init[0], &bound, sizeof(double));
        hipError_t e = hipMemcpyAsync(ctx->counters.p, init, sizeof(init), hipMemcpyHostToDevice, s);
        if (e == hipSuccess && nq) e = launch_ccd(a, grid, 256, s, ctx->ccdw);
        unsigned long long res[4] = { 0, 0, 0, 0 };
        if (e == hipSuccess) e = hipMemcpyAsync(res, ctx->counters.p, sizeof(res), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        std::memcpy(&g, &res[0], sizeof(double));
        evals += (long long)res[2];
        launches += ctx->ccdw.passes;
        kernel_ms += ctx->ccdw.kernel_ms;
        return e;
    };
    if (window < inf) {
        RIGC_HIP(run(window));
        if (!(g < window)) RIGC_HIP(run(inf));
    } else {
        RIGC_HIP(run(inf));
    }
    RIGC_HIP(hipEventRecord(ctx->ev2, s));
    RIGC_HIP(hipStreamSynchronize(s));
    float bms = 0, nms = 0;
    RIGC_HIP(hipEventElapsedTime(&bms, ctx->ev0, ctx->ev1));
    RIGC_HIP(hipEventElapsedTime(&nms, ctx->ev1, ctx->ev2));
    ctx->stats.broad_phase_ms = bms;
    ctx->stats.narrow_phase_ms = nms;
    ctx->stats.box_evaluations = evals;
    ctx->stats.num_candidates = nq;
    ctx->stats.ccd_kernel_ms = kernel_ms;
    ctx->stats.ccd_launches = launches;
    ctx->last_toi = g;
    *earliest_toi = g;  // +inf when no candidate collides
    return RIGC_OK;
}

namespace {

rigc_status copy_potential(rigc_ctx* ctx, const AssemblyOut& out, int32_t flags, double* energy, double* grad,
                           rigc_block_hessian* hess) {
    const DevScene& d = ctx->scene;
    hipStream_t st = ctx->stream;
    *energy = out.energy;
    const bool want_hess = (flags & RIGC_HESSIAN) != 0;
    const bool want_grad = want_hess || (flags & RIGC_GRADIENT) != 0;
    if (want_grad && grad)
        RIGC_HIP(hipMemcpyAsync(grad, ctx->as.grad.p, 6 * sizeof(double) * d.nb, hipMemcpyDeviceToHost, st));
    bool fits = true;
    if (want_hess && hess) {
        RIGC_HIP(hipMemcpyAsync(hess->diag, ctx->as.diag.p, 36 * sizeof(double) * d.nb, hipMemcpyDeviceToHost, st));
        hess->count = out.n_off;
        fits = out.n_off <= hess->capacity;
        if (fits && out.n_off) {
            RIGC_HIP(hipMemcpyAsync(hess->pairs, ctx->as.opairs.p, 2 * sizeof(int32_t) * out.n_off,
                                    hipMemcpyDeviceToHost, st));
            RIGC_HIP(hipMemcpyAsync(hess->off, ctx->as.oblocks.p, 36 * sizeof(double) * out.n_off,
                                    hipMemcpyDeviceToHost, st));
        }
    }
    RIGC_HIP(hipStreamSynchronize(st));
    if (!fits) return ctx->fail(RIGC_ERR_CAPACITY, "off-diagonal Hessian capacity too small");
    return RIGC_OK;
}

bool bad_hess(int32_t flags, double* grad, rigc_block_hessian* hess) {
    const bool want_hess = (flags & RIGC_HESSIAN) != 0;
    const bool want_grad = want_hess || (flags & RIGC_GRADIENT) != 0;
    if (want_grad && !grad) return true;
    if (want_hess && (!hess || !hess->diag || hess->capacity < 0 || (hess->capacity && (!hess->pairs || !hess->off))))
        return true;
    return false;
}

}  // namespace

rigc_status rigc_construct_constraint_set(rigc_ctx* ctx, const double* poses, double dhat, double dmin,
                                          int64_t* counts) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_construct_constraint_set: no scene");
    if (!poses || !(dhat > 0) || !(dmin >= 0))
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_construct_constraint_set: bad arguments");
    RIGC_HIP(hipSetDevice(ctx->device));
    rigc_status s = upload_poses(ctx, poses, nullptr);
    if (s != RIGC_OK) return s;
    RIGC_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    RIGC_HIP(build_constraint_set(ctx->as, ctx->bp, ctx->scene, ctx->poses0.as<double>(), dhat, dmin, ctx->num_cus,
                                  ctx->stream));
    RIGC_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    RIGC_HIP(hipStreamSynchronize(ctx->stream));
    float ms = 0;
    RIGC_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.assembly_ms = ms;
    ctx->stats.num_candidates = ctx->bp.n_ee + ctx->bp.n_fv;
    if (counts)
        for (int k = 0; k < 4; k++) counts[k] = ctx->as.counts[k];
    return RIGC_OK;
}

rigc_status rigc_get_constraint_set(rigc_ctx* ctx, int64_t capacity, int32_t* records, double* eps_x, int64_t* n) {
    if (!ctx || !n) return RIGC_ERR_INVALID_ARGUMENT;
    if (!ctx->as.have_cons) return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_get_constraint_set: no constraint set");
    *n = ctx->as.n_con;
    if (ctx->as.n_con > capacity) return ctx->fail(RIGC_ERR_CAPACITY, "rigc_get_constraint_set: capacity too small");
    if (ctx->as.n_con == 0) return RIGC_OK;
    if (!records) return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_get_constraint_set: records is NULL");
    std::vector<ConRec> h(ctx->as.n_con);
    RIGC_HIP(hipMemcpyAsync(h.data(), ctx->as.cons.p, sizeof(ConRec) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    RIGC_HIP(hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < h.size(); i++) {
        int32_t* r = records + 6 * i;
        r[0] = h[i].type; r[1] = h[i].a; r[2] = h[i].b; r[3] = h[i].mult; r[4] = h[i].body[0]; r[5] = h[i].body[1];
        if (eps_x) eps_x[i] = h[i].eps_x;
    }
    return RIGC_OK;
}

rigc_status rigc_compute_barrier_term(rigc_ctx* ctx, const double* poses, double dhat, double dmin, int32_t flags,
                                      double* energy, double* grad, rigc_block_hessian* hess) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_compute_barrier_term: no scene");
    if (!poses || !energy || !(dhat > 0) || !(dmin >= 0) || bad_hess(flags, grad, hess))
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_compute_barrier_term: bad arguments");
    RIGC_HIP(hipSetDevice(ctx->device));
    rigc_status s = upload_poses(ctx, poses, nullptr);
    if (s != RIGC_OK) return s;
    hipStream_t st = ctx->stream;
    RIGC_HIP(hipEventRecord(ctx->ev0, st));
    if (flags & RIGC_BUILD_CONSTRAINTS)
        RIGC_HIP(build_constraint_set(ctx->as, ctx->bp, ctx->scene, ctx->poses0.as<double>(), dhat, dmin,
                                      ctx->num_cus, st));
    if (!ctx->as.have_cons)
        return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_compute_barrier_term: no constraint set (build one first)");
    AssemblyOut out;
    RIGC_HIP(assemble_potential(ctx->as, ctx->scene, false, ctx->poses0.as<double>(), nullptr, dhat, dmin, 0.0,
                                (flags & RIGC_GRADIENT) != 0, (flags & RIGC_HESSIAN) != 0, ctx->num_cus, out, st));
    RIGC_HIP(hipEventRecord(ctx->ev1, st));
    s = copy_potential(ctx, out, flags, energy, grad, hess);
    float ms = 0;
    RIGC_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.assembly_ms = ms;
    return s;
}

rigc_status rigc_compute_minimum_distance(rigc_ctx* ctx, const double* poses, double dhat, double dmin,
                                          int32_t flags, double* min_distance) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_compute_minimum_distance: no scene");
    if (!poses || !min_distance || !(dhat > 0) || !(dmin >= 0))
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_compute_minimum_distance: bad arguments");
    RIGC_HIP(hipSetDevice(ctx->device));
    rigc_status s = upload_poses(ctx, poses, nullptr);
    if (s != RIGC_OK) return s;
    if (flags & RIGC_BUILD_CONSTRAINTS)
        RIGC_HIP(build_constraint_set(ctx->as, ctx->bp, ctx->scene, ctx->poses0.as<double>(), dhat, dmin,
                                      ctx->num_cus, ctx->stream));
    if (!ctx->as.have_cons)
        return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_compute_minimum_distance: no constraint set (build one first)");
    double m2 = HUGE_VAL;
    RIGC_HIP(rigc::minimum_distance(ctx->as, ctx->scene, ctx->poses0.as<double>(), m2, ctx->stream));
    *min_distance = std::sqrt(m2);
    return RIGC_OK;
}

rigc_status rigc_update_friction_constraints(rigc_ctx* ctx, const double* poses, double dhat, double dmin,
                                             double barrier_stiffness, double mu, int32_t flags, int64_t* count) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_update_friction_constraints: no scene");
    if (!poses || !(dhat > 0) || !(dmin >= 0) || !(barrier_stiffness >= 0) || !(mu >= 0))
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_update_friction_constraints: bad arguments");
    RIGC_HIP(hipSetDevice(ctx->device));
    rigc_status s = upload_poses(ctx, poses, nullptr);
    if (s != RIGC_OK) return s;
    if (flags & RIGC_BUILD_CONSTRAINTS)
        RIGC_HIP(build_constraint_set(ctx->as, ctx->bp, ctx->scene, ctx->poses0.as<double>(), dhat, dmin,
                                      ctx->num_cus, ctx->stream));
    if (!ctx->as.have_cons)
        return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_update_friction_constraints: no constraint set (build one first)");
    RIGC_HIP(build_friction_set(ctx->as, ctx->scene, ctx->poses0.as<double>(), dhat, dmin, barrier_stiffness, mu,
                                ctx->stream));
    RIGC_HIP(hipStreamSynchronize(ctx->stream));
    if (count) *count = ctx->as.n_fric;
    return RIGC_OK;
}

rigc_status rigc_get_friction_constraints(rigc_ctx* ctx, int64_t capacity, double* records, int64_t* n) {
    if (!ctx || !n) return RIGC_ERR_INVALID_ARGUMENT;
    if (!ctx->as.have_fric) return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_get_friction_constraints: no friction set");
    *n = ctx->as.n_fric;
    if (ctx->as.n_fric > capacity)
        return ctx->fail(RIGC_ERR_CAPACITY, "rigc_get_friction_constraints: capacity too small");
    if (ctx->as.n_fric == 0) return RIGC_OK;
    if (!records) return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_get_friction_constraints: records is NULL");
    std::vector<FricRec> h(ctx->as.n_fric);
    RIGC_HIP(hipMemcpyAsync(h.data(), ctx->as.fric.p, sizeof(FricRec) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    RIGC_HIP(hipStreamSynchronize(ctx->stream));
    for (size_t i = 0; i < h.size(); i++) {
        double* r = records + 9 * i;
        r[0] = h[i].normal_force;
        r[1] = h[i].coords[0];
        r[2] = h[i].coords[1];
        for (int k = 0; k < 6; k++) r[3 + k] = h[i].basis[k];
    }
    return RIGC_OK;
}

rigc_status rigc_compute_friction_term(rigc_ctx* ctx, const double* poses, const double* poses_t0,
                                       double epsv_times_h, int32_t flags, double* energy, double* grad,
                                       rigc_block_hessian* hess) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_compute_friction_term: no scene");
    if (!poses || !poses_t0 || !energy || !(epsv_times_h > 0) || bad_hess(flags, grad, hess))
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_compute_friction_term: bad arguments");
    if (!ctx->as.have_fric)
        return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_compute_friction_term: no friction set (update it first)");
    RIGC_HIP(hipSetDevice(ctx->device));
    rigc_status s = upload_poses(ctx, poses, poses_t0);
    if (s != RIGC_OK) return s;
    hipStream_t st = ctx->stream;
    RIGC_HIP(hipEventRecord(ctx->ev0, st));
    AssemblyOut out;
    RIGC_HIP(assemble_potential(ctx->as, ctx->scene, true, ctx->poses0.as<double>(), ctx->poses1.as<double>(), 0.0,
                                0.0, epsv_times_h, (flags & RIGC_GRADIENT) != 0, (flags & RIGC_HESSIAN) != 0,
                                ctx->num_cus, out, st));
    RIGC_HIP(hipEventRecord(ctx->ev1, st));
    s = copy_potential(ctx, out, flags, energy, grad, hess);
    float ms = 0;
    RIGC_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.assembly_ms = ms;
    return s;
}

}  // extern "C"
