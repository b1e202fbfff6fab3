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


// Selected scenes of a call: the device lists of SceneSel in one buffer,
// plus the host-side layout of the request order.
struct SelState {
    DevBuf buf;
    SceneSel sel;
    std::vector<int> req_start;  // [n_sel + 1] first body of each selected scene in request order
    HostBuf host;                // page-locked staging of the last upload
};

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
    std::vector<int> h_body_scene;  // [nb]

    // per-call buffers
    DevBuf poses0, poses1, stage0, stage1, out_rows, ee, fv, toi, hit, counters, qbest, qvisits, qflag;
    DevBuf pl_arena, pl_ids;  // piecewise-linear CCD lane slices and re-run query list
    bool pl_small_slices = false;
    DevBuf poses2;  // the friction term's start-of-step poses
    // Content tags of poses0 / poses1 / poses2 (0: unknown; equal non-zero
    // tags, equal contents) and the host copies of their last full uploads:
    // an upload of the same values is skipped, and world vertices computed
    // from them are reused.
    unsigned long long pose_tag[3] = { 0, 0, 0 }, tag_counter = 0;
    std::vector<double> h_last[3];
    DevBuf scene_best, scene_skip, scene_par;
    DevBuf enq_state;  // rigc_narrow_phase_enqueue's per-call state, used by no other call
    DevBuf scene_evals;                       // box evaluations per scene of the last earliest-TOI call (multi-scene contexts)
    std::vector<unsigned long long> h_scene_evals;
    // page-locked staging: poses (uploads), per-scene parameters, the
    // earliest-TOI bounds and results, and potential outputs; a staging
    // buffer is rewritten only after ev_stage, recorded behind the last
    // upload from it, has completed
    HostBuf h_pose0, h_pose1, h_pose2, h_par, h_toi, h_out;
    hipEvent_t ev_stage = nullptr;
    SelState sel_all, sel_req;
    BroadPhase bp;
    Assembly as;
    CcdWork ccdw;
    Intersections ix;
    // rigid CCD tunables (rigc_tuning; results do not depend on them)
    int walk_budget = 4096;
    int spec_levels = 5;
    double window_factor = 4.0;
    long long wide_queries = 4096;
    int detection = RIGC_DETECTION_HASH_GRID;  // broad phase of rigc_broad_phase and the earliest TOI
    std::vector<double> last_toi;  // per scene: answer of its previous earliest-TOI call
    hipEvent_t ev2 = nullptr;
    unsigned char* scratch = nullptr;  // this device's no-allocation array (rigc_narrow_phase_noalloc)

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
    ctx->h_body_scene.clear();
    ctx->has_scene = false;
}

// Waits until the uploads queued from the page-locked staging buffers have
// been made (normally long done: every entry point synchronises its stream).
hipError_t stage_wait(rigc_ctx* ctx) { return hipEventSynchronize(ctx->ev_stage); }

DevBuf& pose_buf(rigc_ctx* ctx, int k) { return k == 0 ? ctx->poses0 : (k == 1 ? ctx->poses1 : ctx->poses2); }

void share_tags(rigc_ctx* ctx) {
    for (int k = 0; k < 3; k++) {
        ctx->as.tag_ptr[k] = pose_buf(ctx, k).as<double>();
        ctx->as.tag[k] = ctx->pose_tag[k];
    }
}

// Pose array k was written other than by put_poses (or lost its contents).
void forget_poses(rigc_ctx* ctx, int k) {
    ctx->pose_tag[k] = 0;
    ctx->h_last[k].clear();
    share_tags(ctx);
}

// Builds and uploads the selection of scenes ids[0..n) (strictly ascending).
hipError_t make_sel(rigc_ctx* ctx, SelState& S, const int* ids, int n) {
    const DevScene& d = ctx->scene;
    const std::vector<int>& sb = d.h_scene_body;
    std::vector<int> scenes(ids, ids + n), bodies;
    std::vector<int4> tiles;
    S.req_start.assign(1, 0);
    for (int i = 0; i < n; i++) {
        const int b0 = sb[ids[i]], b1 = sb[ids[i] + 1];
        for (int b = b0; b < b1; b++) bodies.push_back(b);
        S.req_start.push_back((int)bodies.size());
        if (b1 - b0 < 2) continue;
        const int nt = (b1 - b0 + 63) / 64;
        for (int ti = 0; ti < nt; ti++)
            for (int tj = ti; tj < nt; tj++) tiles.push_back(make_int4(b0 + 64 * ti, b0 + 64 * tj, b1, 0));
    }
    // layout: tiles (16-byte aligned) | scenes | bodies
    const size_t tb = sizeof(int4) * tiles.size(), sbytes = sizeof(int) * scenes.size();
    const size_t bytes = tb + sbytes + sizeof(int) * bodies.size() + 16;
    hipError_t e = stage_wait(ctx);
    if (e == hipSuccess) e = S.host.ensure(bytes);
    if (e == hipSuccess) e = S.buf.ensure(bytes);
    if (e != hipSuccess) return e;
    char* host = S.host.as<char>();
    std::memset(host, 0, bytes);
    if (!tiles.empty()) std::memcpy(host, tiles.data(), tb);
    if (!scenes.empty()) std::memcpy(host + tb, scenes.data(), sbytes);
    if (!bodies.empty()) std::memcpy(host + tb + sbytes, bodies.data(), sizeof(int) * bodies.size());
    e = hipMemcpyAsync(S.buf.p, host, bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev_stage, ctx->stream);
    if (e != hipSuccess) return e;
    char* base = S.buf.as<char>();
    S.sel.n_scenes = n;
    S.sel.scenes = reinterpret_cast<const int*>(base + tb);
    S.sel.n_bodies = (int)bodies.size();
    S.sel.bodies = reinterpret_cast<const int*>(base + tb + sbytes);
    S.sel.n_tiles = (long long)tiles.size();
    S.sel.tiles = reinterpret_cast<const int4*>(base);
    S.sel.h_scenes = scenes;
    S.sel.all = n == d.n_scenes;
    return hipSuccess;
}

// Uploads scenes[0..n) as one concatenated scene set.
rigc_status load_scenes(rigc_ctx* ctx, int n, const rigc_scene* scenes, const char* who) {
    const std::string w(who);
    if (n < 1) return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, w + ": no scenes");
    int NB = 0, NV = 0, NE = 0, NF = 0;
    for (int k = 0; k < n; k++) {
        const rigc_scene* s = scenes + k;
        if (s->num_bodies < 1 || s->num_vertices < 0 || s->num_edges < 0 || s->num_faces < 0)
            return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, w + ": negative sizes");
        if (!s->vertices || !s->body_vertex_start || !s->body_edge_start || !s->body_face_start || !s->group_ids
            || !s->is_static || (s->num_edges && !s->edges) || (s->num_faces && (!s->faces || !s->face_edges)))
            return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, w + ": missing arrays");
        const int nb = s->num_bodies, nv = s->num_vertices, ne = s->num_edges, nf = s->num_faces;
        if (s->body_vertex_start[0] != 0 || s->body_vertex_start[nb] != nv || s->body_edge_start[0] != 0
            || s->body_edge_start[nb] != ne || s->body_face_start[0] != 0 || s->body_face_start[nb] != nf)
            return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, w + ": body ranges do not cover the arrays");
        for (int i = 0; i < 2 * ne; i++)
            if (s->edges[i] < 0 || s->edges[i] >= nv)
                return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, w + ": edge vertex out of range");
        for (int i = 0; i < 3 * nf; i++)
            if (s->faces[i] < 0 || s->faces[i] >= nv || s->face_edges[i] < 0 || s->face_edges[i] >= ne)
                return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, w + ": face index out of range");
        // every element refers to its own body's vertices (and a face to its
        // body's edges): the per-body kernels index body-local arrays
        for (int b = 0; b < nb; b++) {
            const int v0 = s->body_vertex_start[b], v1 = s->body_vertex_start[b + 1];
            const int e0 = s->body_edge_start[b], e1 = s->body_edge_start[b + 1];
            if (v1 < v0 || e1 < e0 || s->body_face_start[b + 1] < s->body_face_start[b])
                return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, w + ": body ranges not ascending");
            for (int i = 2 * e0; i < 2 * e1; i++)
                if (s->edges[i] < v0 || s->edges[i] >= v1)
                    return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, w + ": an edge refers to another body's vertex");
            for (int i = 3 * s->body_face_start[b]; i < 3 * s->body_face_start[b + 1]; i++)
                if (s->faces[i] < v0 || s->faces[i] >= v1 || s->face_edges[i] < e0 || s->face_edges[i] >= e1)
                    return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, w + ": a face refers to another body's element");
        }
        if ((long long)NV + nv > INT32_MAX / 2 || (long long)NE + ne > INT32_MAX / 2 || (long long)NF + nf > INT32_MAX / 2)
            return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, w + ": scenes too large");
        NB += nb; NV += nv; NE += ne; NF += nf;
    }

    // Concatenation with offset element ids, and the host-side derived data
    // of RigidBody's constructor (src/physics/rigid_body.cpp:124-142) and
    // RigidBody::edge_length (src/physics/rigid_body.hpp:151-156).
    std::vector<double> V, r_max, elen;
    std::vector<int> E, F, FE, v_body, e_body, f_body, bvs{ 0 }, bes{ 0 }, bfs{ 0 }, group, body_scene;
    std::vector<int> scene_body{ 0 }, scene_edge{ 0 };
    std::vector<uint8_t> is_static;
    V.reserve(3 * (size_t)NV);
    int gmax = 0;
    for (int k = 0; k < n; k++) {
        const rigc_scene* s = scenes + k;
        const int B0 = (int)group.size(), V0 = (int)(V.size() / 3), E0 = (int)(E.size() / 2);
        const int nb = s->num_bodies, nv = s->num_vertices, ne = s->num_edges, nf = s->num_faces;
        V.insert(V.end(), s->vertices, s->vertices + 3 * (size_t)nv);
        for (int i = 0; i < 2 * ne; i++) E.push_back(V0 + s->edges[i]);
        for (int i = 0; i < 3 * nf; i++) {
            F.push_back(V0 + s->faces[i]);
            FE.push_back(E0 + s->face_edges[i]);
        }
        // group ids of different scenes never compare equal
        int g0 = 0;
        for (int b = 0; b < nb; b++) g0 = std::max(g0, s->group_ids[b] + 1);
        for (int b = 0; b < nb; b++) {
            group.push_back(gmax + s->group_ids[b]);
            is_static.push_back(s->is_static[b]);
            body_scene.push_back(k);
            r_max.push_back(0.0);
            for (int i = s->body_vertex_start[b]; i < s->body_vertex_start[b + 1]; i++) {
                v_body.push_back(B0 + b);
                const double* v = s->vertices + 3 * i;
                r_max.back() = std::max(r_max.back(), std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]));
            }
            for (int i = s->body_edge_start[b]; i < s->body_edge_start[b + 1]; i++) {
                e_body.push_back(B0 + b);
                const double* a = s->vertices + 3 * s->edges[2 * i];
                const double* c = s->vertices + 3 * s->edges[2 * i + 1];
                const double d0 = c[0] - a[0], d1 = c[1] - a[1], d2 = c[2] - a[2];
                elen.push_back(std::sqrt((d0 * d0 + d1 * d1) + d2 * d2));
            }
            for (int i = s->body_face_start[b]; i < s->body_face_start[b + 1]; i++) f_body.push_back(B0 + b);
            bvs.push_back(V0 + s->body_vertex_start[b + 1]);
            bes.push_back(E0 + s->body_edge_start[b + 1]);
            bfs.push_back((int)(F.size() / 3) - nf + s->body_face_start[b + 1]);
        }
        gmax += std::max(g0, 1);
        scene_body.push_back((int)group.size());
        scene_edge.push_back((int)(E.size() / 2));
    }

    RIGC_HIP(hipSetDevice(ctx->device));
    RIGC_HIP(hipStreamSynchronize(ctx->stream));
    free_scene(ctx);
    // new geometry: no pose array or world-vertex array is reusable
    for (int k = 0; k < 3; k++) {
        ctx->pose_tag[k] = 0;
        ctx->h_last[k].clear();
        ctx->as.tag[k] = 0;
    }
    ctx->as.W_tag = ctx->as.W0_tag = ctx->as.F_tag = 0;
    DevScene& d = ctx->scene;
    d.nb = NB; d.nv = NV; d.ne = NE; d.nf = NF;
    d.n_scenes = n;
    RIGC_HIP(upload(ctx, &d.V, V.data(), V.size()));
    RIGC_HIP(upload(ctx, &d.E, E.data(), E.size()));
    RIGC_HIP(upload(ctx, &d.F, F.data(), F.size()));
    RIGC_HIP(upload(ctx, &d.FE, FE.data(), FE.size()));
    RIGC_HIP(upload(ctx, &d.v_body, v_body.data(), v_body.size()));
    RIGC_HIP(upload(ctx, &d.e_body, e_body.data(), e_body.size()));
    RIGC_HIP(upload(ctx, &d.f_body, f_body.data(), f_body.size()));
    RIGC_HIP(upload(ctx, &d.body_vstart, bvs.data(), bvs.size()));
    RIGC_HIP(upload(ctx, &d.body_estart, bes.data(), bes.size()));
    RIGC_HIP(upload(ctx, &d.body_fstart, bfs.data(), bfs.size()));
    RIGC_HIP(upload(ctx, &d.group, group.data(), group.size()));
    RIGC_HIP(upload(ctx, &d.is_static, is_static.data(), is_static.size()));
    RIGC_HIP(upload(ctx, &d.edge_len, elen.data(), elen.size()));
    RIGC_HIP(upload(ctx, &d.r_max, r_max.data(), r_max.size()));
    RIGC_HIP(upload(ctx, &d.body_scene, body_scene.data(), body_scene.size()));
    RIGC_HIP(upload(ctx, &d.scene_body, scene_body.data(), scene_body.size()));
    {
        // MeshSelector ownership for the BVH broad phase
        std::vector<int> v2f(NV, INT32_MAX), e2f(NE, INT32_MAX), on_edge(NV, 0);
        for (int f = 0; f < NF; f++)
            for (int j = 0; j < 3; j++) {
                v2f[F[3 * f + j]] = std::min(v2f[F[3 * f + j]], f);
                e2f[FE[3 * f + j]] = std::min(e2f[FE[3 * f + j]], f);
            }
        for (int e = 0; e < 2 * NE; e++) on_edge[E[e]] = 1;
        bool ok = true;
        for (int v = 0; v < NV && ok; v++) ok = on_edge[v] != 0;
        for (int e = 0; e < NE && ok; e++) ok = e2f[e] != INT32_MAX;
        d.bvh_ok = ok;
        d.max_body_vertices = 0;
        d.max_body_faces = 0;
        for (int b = 0; b < NB; b++) {
            d.max_body_vertices = std::max(d.max_body_vertices, bvs[b + 1] - bvs[b]);
            d.max_body_faces = std::max(d.max_body_faces, bfs[b + 1] - bfs[b]);
        }
        RIGC_HIP(upload(ctx, &d.v2f, v2f.data(), v2f.size()));
        RIGC_HIP(upload(ctx, &d.e2f, e2f.data(), e2f.size()));
        RIGC_HIP(hipStreamSynchronize(ctx->stream));  // the host arrays end with this block
    }
    d.h_scene_body = scene_body;
    d.h_scene_edge = scene_edge;
    ctx->h_body_scene = body_scene;
    std::vector<int> all(n);
    for (int k = 0; k < n; k++) all[k] = k;
    RIGC_HIP(make_sel(ctx, ctx->sel_all, all.data(), n));
    RIGC_HIP(ctx->scene_best.ensure(sizeof(unsigned long long) * n));
    RIGC_HIP(ctx->scene_skip.ensure((size_t)n));
    RIGC_HIP(ctx->scene_par.ensure(3 * sizeof(double) * n));
    RIGC_HIP(hipStreamSynchronize(ctx->stream));
    ctx->as.have_cons = ctx->as.have_fric = false;
    ctx->as.n_con = ctx->as.n_fric = 0;
    ctx->as.fric_off.assign(n + 1, 0);
    ctx->as.fric_valid.assign(n, 0);
    ctx->last_toi.assign(n, HUGE_VAL);
    ctx->has_scene = true;
    return RIGC_OK;
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
    if (e == hipSuccess) e = hipEventCreate(&ctx->ev_stage);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev_stage, ctx->stream);
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, device);
    void* scratch = nullptr;
    if (e == hipSuccess) e = noalloc_scratch(device, &scratch);  // device current
    if (e != hipSuccess) {
        delete ctx;
        return RIGC_ERR_DEVICE;
    }
    ctx->scratch = (unsigned char*)scratch;
    ctx->num_cus = prop.multiProcessorCount;
    // RIGC_CCD_WALK_BUDGET: box evaluations after which a lane hands its
    // pending subtrees to the next narrow-phase pass (0 = never)
    if (const char* wb = std::getenv("RIGC_CCD_WALK_BUDGET")) ctx->walk_budget = std::atoi(wb);
    // RIGC_BROAD_PHASE_MIN_CAPACITY: start the broad phase's candidate and
    // body-pair capacities from this floor instead of the defaults (a small
    // value drives the passes that regrow them; used by the tests)
    if (const char* mc = std::getenv("RIGC_BROAD_PHASE_MIN_CAPACITY")) ctx->bp.min_cap = std::atoll(mc);
    // RIGC_PLCCD_SMALL_SLICES: the piecewise-linear CCD skips its LDS pass
    // and starts with tiny HBM lane slices, so that most queries outgrow the
    // first two passes (used by the tests of the re-run passes)
    if (const char* ss = std::getenv("RIGC_PLCCD_SMALL_SLICES")) ctx->pl_small_slices = std::atoi(ss) != 0;
    *out = ctx;
    return RIGC_OK;
}

rigc_status rigc_destroy(rigc_ctx* ctx) {
    if (!ctx) return RIGC_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_scene(ctx);
    for (DevBuf* b : { &ctx->poses0, &ctx->poses1, &ctx->poses2, &ctx->stage0, &ctx->stage1, &ctx->out_rows, &ctx->ee, &ctx->fv,
                       &ctx->toi, &ctx->hit, &ctx->counters, &ctx->qbest, &ctx->qvisits, &ctx->qflag,
                       &ctx->scene_best, &ctx->scene_skip, &ctx->scene_par, &ctx->scene_evals, &ctx->sel_all.buf, &ctx->sel_req.buf,
                       &ctx->pl_arena, &ctx->pl_ids, &ctx->enq_state })
        b->release();
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->ev2) (void)hipEventDestroy(ctx->ev2);
    if (ctx->ev_stage) (void)hipEventDestroy(ctx->ev_stage);
    for (HostBuf* b : { &ctx->h_pose0, &ctx->h_pose1, &ctx->h_pose2, &ctx->h_par, &ctx->h_toi, &ctx->h_out, &ctx->sel_all.host,
                        &ctx->sel_req.host, &ctx->bp.h_counters })
        b->release();
    for (DevBuf* b : { &ctx->bp.box_raw, &ctx->bp.box_inf, &ctx->bp.inc, &ctx->bp.dom, &ctx->bp.vmin, &ctx->bp.vmax,
                       &ctx->bp.bhull, &ctx->bp.vc, &ctx->bp.vh, &ctx->bp.ec, &ctx->bp.eh, &ctx->bp.fc, &ctx->bp.fh,
                       &ctx->bp.pairs, &ctx->bp.keys_ee, &ctx->bp.keys_fv, &ctx->bp.tmp, &ctx->bp.hist,
                       &ctx->bp.counters, &ctx->bp.out_ee, &ctx->bp.out_fv })
        b->release();
    ctx->as.release();
    ctx->ccdw.release();
    ctx->ix.release();
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
    return load_scenes(ctx, 1, s, "rigc_set_scene");
}

rigc_status rigc_set_scenes(rigc_ctx* ctx, int32_t num_scenes, const rigc_scene* scenes) {
    if (!ctx || !scenes) return RIGC_ERR_INVALID_ARGUMENT;
    return load_scenes(ctx, num_scenes, scenes, "rigc_set_scenes");
}

rigc_status rigc_get_scene_count(const rigc_ctx* ctx, int32_t* num_scenes) {
    if (!ctx || !num_scenes) return RIGC_ERR_INVALID_ARGUMENT;
    *num_scenes = ctx->has_scene ? ctx->scene.n_scenes : 0;
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
    forget_poses(ctx, 0);
    forget_poses(ctx, 1);
    RIGC_HIP(hipMemcpyAsync(ctx->poses0.p, poses_t0, pose_bytes, hipMemcpyHostToDevice, st));
    RIGC_HIP(hipMemcpyAsync(ctx->poses1.p, poses_t1, pose_bytes, hipMemcpyHostToDevice, st));
    if (n_ee) RIGC_HIP(hipMemcpyAsync(ctx->ee.p, ee, 2 * sizeof(int32_t) * n_ee, hipMemcpyHostToDevice, st));
    if (n_fv) RIGC_HIP(hipMemcpyAsync(ctx->fv.p, fv, 2 * sizeof(int32_t) * n_fv, hipMemcpyHostToDevice, st));
    const unsigned long long init[4] = { 0, 0, 0, 0 };
    RIGC_HIP(hipMemcpyAsync(ctx->counters.p, init, sizeof(init), hipMemcpyHostToDevice, st));
    const std::vector<unsigned long long> inf_bits(d.n_scenes, 0x7ff0000000000000ULL);
    RIGC_HIP(hipMemcpyAsync(ctx->scene_best.p, inf_bits.data(), sizeof(unsigned long long) * d.n_scenes,
                            hipMemcpyHostToDevice, st));

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
    a.global_best = ctx->scene_best.as<unsigned long long>();
    a.body_scene = d.body_scene;
    a.scene_skip = nullptr;
    a.counter = ctx->counters.as<unsigned long long>() + 1;
    a.total_evals = ctx->counters.as<unsigned long long>() + 2;
    a.toi_out = ctx->toi.as<double>();
    a.hit_out = ctx->hit.as<signed char>();
    a.q_best = ctx->qbest.as<unsigned long long>();
    a.q_visits = ctx->qvisits.as<unsigned int>();
    a.q_flag = ctx->qflag.as<int>();

    const int block = 256;
    const int grid = ccd_grid(ctx->num_cus, (long long)(n_ee + n_fv), ctx->wide_queries);  // persistent
    RIGC_HIP(hipEventRecord(ctx->ev0, st));
    a.walk_budget = ctx->walk_budget;
    a.spec_levels = ctx->spec_levels;
    if (nq) RIGC_HIP(launch_ccd(a, grid, block, st, ctx->ccdw));
    RIGC_HIP(hipEventRecord(ctx->ev1, st));

    std::vector<double> tois(nq);
    std::vector<signed char> hits(nq);
    if (nq) {
        RIGC_HIP(hipMemcpyAsync(tois.data(), ctx->toi.p, sizeof(double) * nq, hipMemcpyDeviceToHost, st));
        RIGC_HIP(hipMemcpyAsync(hits.data(), ctx->hit.p, nq, hipMemcpyDeviceToHost, st));
    }
    unsigned long long res[4];
    std::vector<unsigned long long> sbest(d.n_scenes);
    RIGC_HIP(hipMemcpyAsync(res, ctx->counters.p, sizeof(res), hipMemcpyDeviceToHost, st));
    RIGC_HIP(hipMemcpyAsync(sbest.data(), ctx->scene_best.p, sizeof(unsigned long long) * d.n_scenes,
                            hipMemcpyDeviceToHost, st));
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
        double g = HUGE_VAL;
        for (int k = 0; k < d.n_scenes; k++) {
            double gk;
            std::memcpy(&gk, &sbest[k], sizeof(double));
            g = std::min(g, gk);
        }
        *earliest_toi = g;
    }
    if (toi_out) std::memcpy(toi_out, tois.data(), sizeof(double) * nq);
    if (hit_out) std::memcpy(hit_out, hits.data(), nq);
    if (unresolved)
        return ctx->fail(RIGC_ERR_UNRESOLVED, "rigc_narrow_phase: a query exceeded its evaluation budget");
    return RIGC_OK;
}

rigc_status rigc_narrow_phase_noalloc(rigc_ctx* ctx, const double* poses_t0, const double* poses_t1,
                                      const int32_t* ee, int64_t n_ee, const int32_t* fv, int64_t n_fv,
                                      double toi_tolerance, int32_t per_query, double* earliest_toi,
                                      double* toi_out, int8_t* hit_out) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_narrow_phase_noalloc: no scene");
    if (!poses_t0 || !poses_t1 || !earliest_toi || n_ee < 0 || n_fv < 0 || (n_ee && !ee) || (n_fv && !fv)
        || !(toi_tolerance > 0))
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_narrow_phase_noalloc: bad arguments");
    const DevScene& d = ctx->scene;
    for (int64_t i = 0; i < n_ee; i++)
        if (ee[2 * i] < 0 || ee[2 * i] >= d.ne || ee[2 * i + 1] < 0 || ee[2 * i + 1] >= d.ne)
            return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_narrow_phase_noalloc: edge id out of range");
    for (int64_t i = 0; i < n_fv; i++)
        if (fv[2 * i] < 0 || fv[2 * i] >= d.nf || fv[2 * i + 1] < 0 || fv[2 * i + 1] >= d.nv)
            return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_narrow_phase_noalloc: face/vertex id out of range");
    const NoallocLayout L = noalloc_layout(d.nb, n_ee, n_fv, d.n_scenes);
    if (L.total > NOALLOC_SCRATCH_BYTES)
        return ctx->fail(RIGC_ERR_CAPACITY, "rigc_narrow_phase_noalloc: the scene's bodies and the queries need more "
                                            "than RIGC_NOALLOC_SCRATCH_BYTES of scratch");
    RIGC_HIP(hipSetDevice(ctx->device));
    unsigned char* base = ctx->scratch;
    hipStream_t st = ctx->stream;
    const size_t nq = (size_t)(n_ee + n_fv);
    const size_t pose_bytes = 48 * (size_t)d.nb;
    RIGC_HIP(hipMemcpyAsync(base + L.poses0, poses_t0, pose_bytes, hipMemcpyHostToDevice, st));
    RIGC_HIP(hipMemcpyAsync(base + L.poses1, poses_t1, pose_bytes, hipMemcpyHostToDevice, st));
    if (n_ee) RIGC_HIP(hipMemcpyAsync(base + L.ee, ee, 8 * (size_t)n_ee, hipMemcpyHostToDevice, st));
    if (n_fv) RIGC_HIP(hipMemcpyAsync(base + L.fv, fv, 8 * (size_t)n_fv, hipMemcpyHostToDevice, st));

    CcdArgs a{};
    a.V = d.V; a.E = d.E; a.F = d.F; a.FE = d.FE;
    a.v_body = d.v_body; a.e_body = d.e_body; a.f_body = d.f_body; a.edge_len = d.edge_len;
    a.poses_t0 = (const double*)(base + L.poses0);
    a.poses_t1 = (const double*)(base + L.poses1);
    a.ee = (const int*)(base + L.ee);
    a.fv = (const int*)(base + L.fv);
    a.n_ee = n_ee;
    a.n_fv = n_fv;
    a.spec_levels = ctx->spec_levels;
    a.toi_tol = toi_tolerance;
    a.eval_budget = 0;
    a.use_global_best = per_query ? 0 : 1;
    a.global_best = (unsigned long long*)(base + L.scene_best);
    a.body_scene = d.body_scene;
    a.scene_skip = nullptr;
    unsigned long long* counters = (unsigned long long*)(base + L.counters);
    a.counter = counters + 1;
    a.total_evals = counters + 2;
    a.n_tasks_out = counters + 3;
    a.toi_out = (double*)(base + L.toi);
    a.hit_out = (signed char*)(base + L.hit);
    a.q_best = (unsigned long long*)(base + L.qbest);
    a.q_visits = (unsigned int*)(base + L.qvisits);
    a.q_flag = (int*)(base + L.qflag);

    RIGC_HIP(hipEventRecord(ctx->ev0, st));
    RIGC_HIP(noalloc_narrow_phase(a, ccd_grid(ctx->num_cus, (long long)nq, ctx->wide_queries), base, L, d.n_scenes,
                                  per_query ? 1 : 0, st));
    RIGC_HIP(hipEventRecord(ctx->ev1, st));
    unsigned long long res[2] = { 0, 0 }, evals = 0;
    RIGC_HIP(hipMemcpyAsync(res, base + L.result, sizeof(res), hipMemcpyDeviceToHost, st));
    if (nq) RIGC_HIP(hipMemcpyAsync(&evals, counters + 2, sizeof(evals), hipMemcpyDeviceToHost, st));
    if (nq && toi_out) RIGC_HIP(hipMemcpyAsync(toi_out, base + L.toi, 8 * nq, hipMemcpyDeviceToHost, st));
    if (nq && hit_out) RIGC_HIP(hipMemcpyAsync(hit_out, base + L.hit, nq, hipMemcpyDeviceToHost, st));
    RIGC_HIP(hipStreamSynchronize(st));
    float ms = 0;
    RIGC_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.narrow_phase_ms = ms;
    ctx->stats.box_evaluations = (int64_t)evals;
    ctx->stats.num_candidates = (int64_t)nq;
    ctx->stats.ccd_kernel_ms = ms;
    ctx->stats.ccd_launches = nq ? 1 : 0;
    std::memcpy(earliest_toi, &res[0], sizeof(double));
    if (res[1])
        return ctx->fail(RIGC_ERR_UNRESOLVED, "rigc_narrow_phase_noalloc: a query exceeded the search depth");
    return RIGC_OK;
}

namespace {

// Poses of every body of the context (scene order).
// Copies n_rows pose rows of p0 (and p1) through the page-locked staging
// into dst0 (dst1).
rigc_status stage_poses(rigc_ctx* ctx, const double* p0, const double* p1, size_t n_rows, void* dst0, void* dst1) {
    const size_t bytes = 6 * sizeof(double) * n_rows;
    RIGC_HIP(stage_wait(ctx));
    RIGC_HIP(ctx->h_pose0.ensure(bytes));
    std::memcpy(ctx->h_pose0.p, p0, bytes);
    RIGC_HIP(hipMemcpyAsync(dst0, ctx->h_pose0.p, bytes, hipMemcpyHostToDevice, ctx->stream));
    if (p1) {
        RIGC_HIP(ctx->h_pose1.ensure(bytes));
        std::memcpy(ctx->h_pose1.p, p1, bytes);
        RIGC_HIP(hipMemcpyAsync(dst1, ctx->h_pose1.p, bytes, hipMemcpyHostToDevice, ctx->stream));
    }
    RIGC_HIP(hipEventRecord(ctx->ev_stage, ctx->stream));
    return RIGC_OK;
}

// Full upload of pose array k, skipped when the array holds these values
// already (the same bytes as its last full upload).
rigc_status put_poses(rigc_ctx* ctx, int k, const double* p) {
    const size_t n = 6 * (size_t)ctx->scene.nb, bytes = sizeof(double) * n;
    DevBuf& d = pose_buf(ctx, k);
    const void* before = d.p;
    RIGC_HIP(d.ensure(bytes));
    std::vector<double>& last = ctx->h_last[k];
    if (d.p == before && ctx->pose_tag[k] && last.size() == n && std::memcmp(last.data(), p, bytes) == 0)
        return RIGC_OK;
    forget_poses(ctx, k);
    HostBuf& h = k == 0 ? ctx->h_pose0 : (k == 1 ? ctx->h_pose1 : ctx->h_pose2);
    RIGC_HIP(stage_wait(ctx));
    RIGC_HIP(h.ensure(bytes));
    std::memcpy(h.p, p, bytes);
    RIGC_HIP(hipMemcpyAsync(d.p, h.p, bytes, hipMemcpyHostToDevice, ctx->stream));
    RIGC_HIP(hipEventRecord(ctx->ev_stage, ctx->stream));
    last.assign(p, p + n);
    ctx->pose_tag[k] = ++ctx->tag_counter;
    share_tags(ctx);
    return RIGC_OK;
}

rigc_status upload_poses(rigc_ctx* ctx, const double* p0, const double* p1) {
    const size_t pose_bytes = 6 * sizeof(double) * (size_t)ctx->scene.nb;
    RIGC_HIP(ctx->poses0.ensure(pose_bytes));
    RIGC_HIP(ctx->poses1.ensure(pose_bytes));
    rigc_status s = put_poses(ctx, 0, p0);
    if (s == RIGC_OK && p1) s = put_poses(ctx, 1, p1);
    return s;
}

// Poses of the selected scenes' bodies, in request order, into the
// scene-order pose arrays (the other rows keep their values).
rigc_status upload_sel_poses(rigc_ctx* ctx, const SelState& S, const double* p0, const double* p1) {
    const size_t nb = (size_t)ctx->scene.nb, ns = (size_t)S.sel.n_bodies;
    RIGC_HIP(ctx->poses0.ensure(6 * sizeof(double) * nb));
    RIGC_HIP(ctx->poses1.ensure(6 * sizeof(double) * nb));
    if (S.sel.all) return upload_poses(ctx, p0, p1);
    RIGC_HIP(ctx->stage0.ensure(6 * sizeof(double) * std::max<size_t>(ns, 1)));
    RIGC_HIP(ctx->stage1.ensure(6 * sizeof(double) * std::max<size_t>(ns, 1)));
    hipStream_t st = ctx->stream;
    forget_poses(ctx, 0);
    if (p1) forget_poses(ctx, 1);
    const rigc_status s = stage_poses(ctx, p0, p1, ns, ctx->stage0.p, ctx->stage1.p);
    if (s != RIGC_OK) return s;
    scatter_rows(S.sel.bodies, (int)ns, 6, ctx->stage0.as<double>(), ctx->poses0.as<double>(), st);
    if (p1) scatter_rows(S.sel.bodies, (int)ns, 6, ctx->stage1.as<double>(), ctx->poses1.as<double>(), st);
    RIGC_HIP(hipGetLastError());
    return RIGC_OK;
}

// Checks and uploads a request's scene list.
rigc_status select_scenes(rigc_ctx* ctx, int32_t n, const int32_t* ids, const char* who) {
    const std::string w(who);
    if (!ctx->has_scene) return ctx->fail(RIGC_ERR_NO_SCENE, w + ": no scene");
    if (n < 1 || !ids) return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, w + ": empty scene list");
    for (int i = 0; i < n; i++)
        if (ids[i] < 0 || ids[i] >= ctx->scene.n_scenes || (i && ids[i] <= ids[i - 1]))
            return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, w + ": scene ids must be ascending and in range");
    RIGC_HIP(hipSetDevice(ctx->device));
    RIGC_HIP(make_sel(ctx, ctx->sel_req, ids, n));
    return RIGC_OK;
}

// The BVH broad phase is chosen for the context: the scene in force must
// support it (a scene set after the choice is checked here).
rigc_status check_detection(rigc_ctx* ctx, const char* who) {
    if (ctx->detection != RIGC_DETECTION_BVH) return RIGC_OK;
    if (!ctx->scene.bvh_ok || bvh_lds_bytes(ctx->scene) > 65536)
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, std::string(who) + ": the scene does not support the BVH broad "
                                                                       "phase (codimensional elements or a body too "
                                                                       "large)");
    return RIGC_OK;
}

CcdArgs ccd_args(rigc_ctx* ctx) {
    const DevScene& d = ctx->scene;
    CcdArgs a{};
    a.V = d.V; a.E = d.E; a.F = d.F; a.FE = d.FE;
    a.v_body = d.v_body; a.e_body = d.e_body; a.f_body = d.f_body; a.edge_len = d.edge_len;
    a.poses_t0 = ctx->poses0.as<double>();
    a.poses_t1 = ctx->poses1.as<double>();
    a.global_best = ctx->scene_best.as<unsigned long long>();
    a.body_scene = d.body_scene;
    a.counter = ctx->counters.as<unsigned long long>() + 1;
    a.total_evals = ctx->counters.as<unsigned long long>() + 2;
    a.spec_levels = ctx->spec_levels;
    return a;
}

// Earliest TOI of each selected scene (poses already uploaded): broad phase
// over the selection, then the narrow phase with a running minimum per
// scene. Windowed search: a first pass starts each scene's minimum at a
// bound T_s (window_factor, 4 unless tuned, times that scene's previous
// answer; successive line-search calls return similar times). It prunes every box with
// t.lower >= T_s, so if it accepts a box of scene s the minimum below T_s
// is s's answer; the scenes where it accepts none are searched again in
// full, with the queries of every other scene skipped. Either way each
// result is the exact minimum.
rigc_status earliest_toi_sel(rigc_ctx* ctx, const SelState& S, double msd, std::vector<double>& out) {
    const rigc_status cs = check_detection(ctx, "earliest TOI");
    if (cs != RIGC_OK) return cs;
    const SceneSel& sel = S.sel;
    const int ns = ctx->scene.n_scenes;
    hipStream_t s = ctx->stream;
    RIGC_HIP(hipEventRecord(ctx->ev0, s));
    // the earliest TOI is a minimum over the candidates: their order is free
    if (ctx->detection == RIGC_DETECTION_BVH)
        RIGC_HIP(run_broad_phase_bvh(ctx->bp, ctx->scene, sel, ctx->poses0.as<double>(), ctx->poses1.as<double>(),
                                     msd / 2.0, true, false, ctx->num_cus, s));
    else
        RIGC_HIP(run_broad_phase(ctx->bp, ctx->scene, sel, ctx->poses0.as<double>(), ctx->poses1.as<double>(),
                                 msd / 2.0, true, false, s));
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
    const int grid = ccd_grid(ctx->num_cus, nq, ctx->wide_queries);  // persistent
    a.walk_budget = ctx->walk_budget;
    // a multi-scene context counts each scene's box evaluations (what a
    // batch scheduler uses to keep deep searches out of shared calls)
    const bool count_scenes = ns > 1;
    if (count_scenes) {
        RIGC_HIP(ctx->scene_evals.ensure(sizeof(unsigned long long) * ns));
        RIGC_HIP(hipMemsetAsync(ctx->scene_evals.p, 0, sizeof(unsigned long long) * ns, s));
        a.scene_evals = ctx->scene_evals.as<unsigned long long>();
    }

    const double inf = HUGE_VAL;
    std::vector<double> window(sel.n_scenes, inf);
    bool any_window = false;
    for (int i = 0; i < sel.n_scenes; i++) {
        const double lt = ctx->last_toi[sel.h_scenes[i]];
        if (ctx->window_factor > 0 && lt * ctx->window_factor < 1.0) {
            window[i] = std::max(ctx->window_factor * lt, std::ldexp(1.0, -20));
            any_window = true;
        }
    }
    out.assign(sel.n_scenes, inf);
    long long evals = 0, launches = 0;
    double kernel_ms = 0.0;
    // page-locked staging: per-scene start bounds and results (bits),
    // the four counters read back, the skip flags
    RIGC_HIP(stage_wait(ctx));
    RIGC_HIP(ctx->h_toi.ensure(sizeof(unsigned long long) * (2 * (size_t)ns + 4) + (size_t)ns));
    unsigned long long* best_bits = ctx->h_toi.as<unsigned long long>();
    unsigned long long* res_bits = best_bits + ns;
    unsigned long long* res = res_bits + ns;
    unsigned char* skip = reinterpret_cast<unsigned char*>(res + 4);
    for (int k = 0; k < ns; k++) best_bits[k] = 0x7ff0000000000000ULL;
    std::memset(skip, 0, (size_t)ns);
    // bound[i]: the starting minimum of selected scene i (NaN: skip it)
    auto run = [&](const std::vector<double>& bound) -> hipError_t {
        bool any_skip = false;
        for (int i = 0; i < sel.n_scenes; i++) {
            const int k = sel.h_scenes[i];
            skip[k] = std::isnan(bound[i]) ? 1 : 0;
            any_skip = any_skip || skip[k];
            const double b = skip[k] ? 0.0 : bound[i];
            std::memcpy(&best_bits[k], &b, sizeof(double));
        }
        // the staging is rewritten only after the previous run's
        // synchronisation (launch_ccd synchronises after every pass); the
        // work and evaluation counters are cleared by launch_ccd
        hipError_t e = hipMemcpyAsync(ctx->scene_best.p, best_bits, sizeof(unsigned long long) * ns,
                                      hipMemcpyHostToDevice, s);
        if (e == hipSuccess && any_skip)
            e = hipMemcpyAsync(ctx->scene_skip.p, skip, (size_t)ns, hipMemcpyHostToDevice, s);
        a.scene_skip = any_skip ? ctx->scene_skip.as<unsigned char>() : nullptr;
        for (int k = 0; k < 4; k++) res[k] = 0;
        const CcdReadback rb[2] = { { res, ctx->counters.p, 4 * sizeof(unsigned long long) },
                                    { res_bits, ctx->scene_best.p, sizeof(unsigned long long) * ns } };
        if (e == hipSuccess) e = launch_ccd(a, grid, 256, s, ctx->ccdw, rb, 2);
        evals += (long long)res[2];
        launches += ctx->ccdw.passes;
        kernel_ms += ctx->ccdw.kernel_ms;
        return e;
    };
    auto result = [&](int i) {
        double g;
        std::memcpy(&g, &res_bits[sel.h_scenes[i]], sizeof(double));
        return g;
    };
    if (nq) {
        RIGC_HIP(run(window));
        std::vector<double> again(sel.n_scenes, std::nan(""));
        bool second = false;
        for (int i = 0; i < sel.n_scenes; i++) {
            const double g = result(i);
            if (window[i] < inf && !(g < window[i])) {
                again[i] = inf;
                second = true;
            } else {
                out[i] = g;
            }
        }
        if (second) {
            RIGC_HIP(run(again));
            for (int i = 0; i < sel.n_scenes; i++)
                if (!std::isnan(again[i])) out[i] = result(i);
        }
        (void)any_window;
    }
    RIGC_HIP(hipEventRecord(ctx->ev2, s));
    if (count_scenes) {
        ctx->h_scene_evals.assign(ns, 0ULL);
        RIGC_HIP(hipMemcpyAsync(ctx->h_scene_evals.data(), ctx->scene_evals.p, sizeof(unsigned long long) * ns,
                                hipMemcpyDeviceToHost, s));
    }
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
    for (int i = 0; i < sel.n_scenes; i++) ctx->last_toi[sel.h_scenes[i]] = out[i];
    return RIGC_OK;
}

// Queues the potential's outputs (rows of the selected bodies, in request
// order) into the page-locked staging, synchronises once, finishes `out`
// and copies to the caller's arrays. The off-diagonal blocks are staged up
// to min(records, capacity): their number is known only after the
// synchronisation and never exceeds the number of records.
rigc_status copy_potential(rigc_ctx* ctx, const SelState& S, AssemblyOut& out, int32_t flags, double* grad,
                           rigc_block_hessian* hess) {
    const DevScene& d = ctx->scene;
    hipStream_t st = ctx->stream;
    const bool want_hess = (flags & RIGC_HESSIAN) != 0 && hess;
    const bool want_grad = (flags & (RIGC_HESSIAN | RIGC_GRADIENT)) != 0 && grad;
    const size_t rows = S.sel.all ? (size_t)d.nb : (size_t)S.sel.n_bodies;
    const long long m = want_hess ? std::min<long long>(out.n_rec, std::max<long long>(hess->capacity, 0)) : 0;
    const size_t o_diag = 6 * sizeof(double) * rows, o_off = o_diag + 36 * sizeof(double) * rows;
    const size_t o_pairs = o_off + 36 * sizeof(double) * (size_t)m;
    RIGC_HIP(ctx->h_out.ensure(o_pairs + 2 * sizeof(int32_t) * (size_t)m));
    char* h = ctx->h_out.as<char>();
    if (!S.sel.all && (want_grad || want_hess))
        RIGC_HIP(ctx->out_rows.ensure(36 * sizeof(double) * std::max<size_t>(rows, 1)));
    // the two row gathers share out_rows; the stream orders each copy
    // before the next gather
    auto rows_out = [&](const DevBuf& src, int width, size_t off) -> hipError_t {
        const void* from = src.p;
        if (!S.sel.all) {
            gather_rows(S.sel.bodies, (int)rows, width, src.as<double>(), ctx->out_rows.as<double>(), st);
            from = ctx->out_rows.p;
        }
        return hipMemcpyAsync(h + off, from, width * sizeof(double) * rows, hipMemcpyDeviceToHost, st);
    };
    if (want_grad && rows) RIGC_HIP(rows_out(ctx->as.grad, 6, 0));
    if (want_hess && rows) RIGC_HIP(rows_out(ctx->as.diag, 36, o_diag));
    if (m > 0) {
        RIGC_HIP(hipMemcpyAsync(h + o_off, ctx->as.oblocks.p, 36 * sizeof(double) * (size_t)m, hipMemcpyDeviceToHost,
                                st));
        RIGC_HIP(hipMemcpyAsync(h + o_pairs, ctx->as.opairs.p, 2 * sizeof(int32_t) * (size_t)m,
                                hipMemcpyDeviceToHost, st));
    }
    RIGC_HIP(hipStreamSynchronize(st));
    finish_potential(ctx->as, out);
    if (want_grad) std::memcpy(grad, h, 6 * sizeof(double) * rows);
    bool fits = true;
    if (want_hess) {
        std::memcpy(hess->diag, h + o_diag, 36 * sizeof(double) * rows);
        hess->count = out.n_off;
        fits = out.n_off <= hess->capacity;
        if (fits && out.n_off) {
            std::memcpy(hess->off, h + o_off, 36 * sizeof(double) * (size_t)out.n_off);
            std::memcpy(hess->pairs, h + o_pairs, 2 * sizeof(int32_t) * (size_t)out.n_off);
            if (!S.sel.all) {
                // scene-order body ids -> request order
                std::vector<int> first(d.n_scenes, -1);
                for (int i = 0; i < S.sel.n_scenes; i++) first[S.sel.h_scenes[i]] = S.req_start[i];
                for (long long k = 0; k < 2 * out.n_off; k++) {
                    const int b = hess->pairs[k];
                    const int sidx = ctx->h_body_scene[b];
                    hess->pairs[k] = first[sidx] + (b - d.h_scene_body[sidx]);
                }
            }
        }
    }
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

double total(const std::vector<double>& v) {
    double t = 0.0;
    for (double x : v) t += x;
    return t;
}

rigc_status build_sel(rigc_ctx* ctx, const SelState& S, double dhat, double dmin) {
    RIGC_HIP(build_constraint_set(ctx->as, ctx->bp, ctx->scene, S.sel, ctx->poses0.as<double>(), dhat, dmin,
                                  ctx->num_cus, ctx->stream));
    ctx->stats.num_constraints = ctx->as.n_con;
    return RIGC_OK;
}

// Uploads per-scene parameters (rows 0..2 of scene_par: kappa, mu, epsv_h).
rigc_status upload_scene_par(rigc_ctx* ctx, int row, const SelState& S, const double* vals) {
    const size_t ns = (size_t)ctx->scene.n_scenes;
    RIGC_HIP(stage_wait(ctx));
    RIGC_HIP(ctx->h_par.ensure(3 * sizeof(double) * ns));
    double* h = ctx->h_par.as<double>() + (size_t)row * ns;
    for (size_t k = 0; k < ns; k++) h[k] = 0.0;
    for (int i = 0; i < S.sel.n_scenes; i++) h[S.sel.h_scenes[i]] = vals[i];
    RIGC_HIP(hipMemcpyAsync(ctx->scene_par.as<double>() + (size_t)row * ns, h, sizeof(double) * ns,
                            hipMemcpyHostToDevice, ctx->stream));
    RIGC_HIP(hipEventRecord(ctx->ev_stage, ctx->stream));
    return RIGC_OK;
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
    rigc_status st = check_detection(ctx, "rigc_broad_phase");
    if (st != RIGC_OK) return st;
    st = upload_poses(ctx, poses_t0, poses_t1);
    if (st != RIGC_OK) return st;
    RIGC_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    if (ctx->detection == RIGC_DETECTION_BVH)
        RIGC_HIP(run_broad_phase_bvh(ctx->bp, ctx->scene, ctx->sel_all.sel, ctx->poses0.as<double>(),
                                     ctx->poses1.as<double>(), inflation_radius, poses_t1 != nullptr, true,
                                     ctx->num_cus, ctx->stream));
    else
        RIGC_HIP(run_broad_phase(ctx->bp, ctx->scene, ctx->sel_all.sel, ctx->poses0.as<double>(),
                                 ctx->poses1.as<double>(), inflation_radius, poses_t1 != nullptr, true, ctx->stream));
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
    std::vector<double> out;
    st = earliest_toi_sel(ctx, ctx->sel_all, minimum_separation_distance, out);
    if (st != RIGC_OK) return st;
    double g = HUGE_VAL;
    for (double x : out) g = std::min(g, x);
    *earliest_toi = g;  // +inf when no candidate collides
    return RIGC_OK;
}

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
    s = build_sel(ctx, ctx->sel_all, dhat, dmin);
    if (s != RIGC_OK) return s;
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
    if (flags & RIGC_BUILD_CONSTRAINTS) {
        s = build_sel(ctx, ctx->sel_all, dhat, dmin);
        if (s != RIGC_OK) return s;
    }
    if (!ctx->as.have_cons)
        return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_compute_barrier_term: no constraint set (build one first)");
    AssemblyOut out;
    RIGC_HIP(assemble_potential(ctx->as, ctx->scene, ctx->sel_all.sel, false, ctx->poses0.as<double>(), nullptr,
                                dhat, dmin, 0.0, nullptr, (flags & RIGC_GRADIENT) != 0,
                                (flags & RIGC_HESSIAN) != 0, ctx->num_cus, out, st, (flags & RIGC_UNPROJECTED) == 0));
    RIGC_HIP(hipEventRecord(ctx->ev1, st));
    s = copy_potential(ctx, ctx->sel_all, out, flags, grad, hess);
    if (!out.pending) *energy = total(out.energy);
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
    if (flags & RIGC_BUILD_CONSTRAINTS) {
        s = build_sel(ctx, ctx->sel_all, dhat, dmin);
        if (s != RIGC_OK) return s;
    }
    if (!ctx->as.have_cons)
        return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_compute_minimum_distance: no constraint set (build one first)");
    std::vector<double> m2;
    RIGC_HIP(rigc::minimum_distance(ctx->as, ctx->scene, ctx->sel_all.sel, ctx->poses0.as<double>(), m2,
                                    ctx->stream));
    double m = HUGE_VAL;
    for (double x : m2) m = std::min(m, x);
    *min_distance = std::sqrt(m);
    return RIGC_OK;
}

rigc_status rigc_detect_intersections(rigc_ctx* ctx, const double* poses, int32_t* pairs, int64_t capacity,
                                      int64_t* count) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_detect_intersections: no scene");
    if (!poses || !count || capacity < 0)
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_detect_intersections: bad arguments");
    RIGC_HIP(hipSetDevice(ctx->device));
    rigc_status s = upload_poses(ctx, poses, nullptr);
    if (s != RIGC_OK) return s;
    RIGC_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    RIGC_HIP(rigc::detect_intersections(ctx->as, ctx->bp, ctx->ix, ctx->scene, ctx->sel_all.sel,
                                        ctx->poses0.as<double>(), ctx->stream));
    RIGC_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    *count = ctx->ix.n;
    const bool fits = !pairs || ctx->ix.n <= capacity;
    if (pairs && fits && ctx->ix.n)
        RIGC_HIP(hipMemcpyAsync(pairs, ctx->ix.out.p, 2 * sizeof(int32_t) * ctx->ix.n, hipMemcpyDeviceToHost,
                                ctx->stream));
    RIGC_HIP(hipStreamSynchronize(ctx->stream));
    float ms = 0;
    RIGC_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.broad_phase_ms = ms;
    if (!fits) return ctx->fail(RIGC_ERR_CAPACITY, "rigc_detect_intersections: output capacity too small");
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
    if (flags & RIGC_BUILD_CONSTRAINTS) {
        s = build_sel(ctx, ctx->sel_all, dhat, dmin);
        if (s != RIGC_OK) return s;
    }
    if (!ctx->as.have_cons)
        return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_update_friction_constraints: no constraint set (build one first)");
    const int ns = ctx->scene.n_scenes;
    const std::vector<double> kap(ns, barrier_stiffness), mus(ns, mu);
    s = upload_scene_par(ctx, 0, ctx->sel_all, kap.data());
    if (s == RIGC_OK) s = upload_scene_par(ctx, 1, ctx->sel_all, mus.data());
    if (s != RIGC_OK) return s;
    RIGC_HIP(build_friction_set(ctx->as, ctx->scene, ctx->sel_all.sel, ctx->poses0.as<double>(), dhat, dmin,
                                ctx->scene_par.as<double>(), ctx->scene_par.as<double>() + ns, ctx->stream));
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
    rigc_status s = upload_poses(ctx, poses, nullptr);
    if (s == RIGC_OK) s = put_poses(ctx, 2, poses_t0);
    if (s != RIGC_OK) return s;
    hipStream_t st = ctx->stream;
    RIGC_HIP(hipEventRecord(ctx->ev0, st));
    AssemblyOut out;
    RIGC_HIP(assemble_potential(ctx->as, ctx->scene, ctx->sel_all.sel, true, ctx->poses0.as<double>(),
                                ctx->poses2.as<double>(), 0.0, 0.0, epsv_times_h, nullptr,
                                (flags & RIGC_GRADIENT) != 0, (flags & RIGC_HESSIAN) != 0, ctx->num_cus, out, st,
                                (flags & RIGC_UNPROJECTED) == 0));
    RIGC_HIP(hipEventRecord(ctx->ev1, st));
    s = copy_potential(ctx, ctx->sel_all, out, flags, grad, hess);
    if (!out.pending) *energy = total(out.energy);
    float ms = 0;
    RIGC_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.assembly_ms = ms;
    return s;
}

rigc_status rigc_friction_adjoint(rigc_ctx* ctx, const double* poses, const double* poses_t0, double epsv_times_h,
                                  double dhat, double dmin, double barrier_stiffness, const double* lambda,
                                  double* out) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_friction_adjoint: no scene");
    if (!poses || !poses_t0 || !lambda || !out || !(epsv_times_h > 0) || !(dhat > 0) || !(dmin >= 0))
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_friction_adjoint: bad arguments");
    if (ctx->scene.n_scenes != 1)
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_friction_adjoint: single-scene contexts only");
    if (!ctx->as.have_fric)
        return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_friction_adjoint: no friction set (update it first)");
    RIGC_HIP(hipSetDevice(ctx->device));
    rigc_status s = upload_poses(ctx, poses, nullptr);
    if (s == RIGC_OK) s = put_poses(ctx, 2, poses_t0);
    if (s != RIGC_OK) return s;
    hipStream_t st = ctx->stream;
    const int nb = ctx->scene.nb;
    RIGC_HIP(ctx->out_rows.ensure(6 * sizeof(double) * (size_t)std::max(nb, 1)));
    RIGC_HIP(hipMemcpyAsync(ctx->out_rows.p, lambda, 6 * sizeof(double) * (size_t)nb, hipMemcpyHostToDevice, st));
    RIGC_HIP(friction_adjoint(ctx->as, ctx->scene, ctx->sel_all.sel, ctx->poses0.as<double>(),
                              ctx->poses2.as<double>(), epsv_times_h, dhat, dmin, barrier_stiffness,
                              ctx->out_rows.as<double>(), out, st));
    return RIGC_OK;
}

// ------------------------------------------------------------ batched scenes
rigc_status rigc_batch_earliest_toi(rigc_ctx* ctx, int32_t n, const int32_t* scene_ids, const double* poses_t0,
                                    const double* poses_t1, double minimum_separation_distance, double* toi) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!poses_t0 || !poses_t1 || !toi || !(minimum_separation_distance >= 0))
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_batch_earliest_toi: bad arguments");
    rigc_status s = select_scenes(ctx, n, scene_ids, "rigc_batch_earliest_toi");
    if (s != RIGC_OK) return s;
    s = upload_sel_poses(ctx, ctx->sel_req, poses_t0, poses_t1);
    if (s != RIGC_OK) return s;
    std::vector<double> out;
    s = earliest_toi_sel(ctx, ctx->sel_req, minimum_separation_distance, out);
    if (s != RIGC_OK) return s;
    std::memcpy(toi, out.data(), sizeof(double) * n);
    return RIGC_OK;
}

rigc_status rigc_get_tuning(const rigc_ctx* ctx, rigc_tuning* t) {
    if (!ctx || !t) return RIGC_ERR_INVALID_ARGUMENT;
    t->ccd_walk_budget = ctx->walk_budget;
    t->ccd_spec_levels = ctx->spec_levels;
    t->ccd_window_factor = ctx->window_factor;
    t->ccd_wide_queries = ctx->wide_queries;
    return RIGC_OK;
}

rigc_status rigc_set_tuning(rigc_ctx* ctx, const rigc_tuning* t) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!t || t->ccd_walk_budget < 0 || t->ccd_spec_levels < 0 || t->ccd_spec_levels > 5 ||
        !(t->ccd_window_factor >= 0) || !std::isfinite(t->ccd_window_factor) || t->ccd_wide_queries < 0)
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_set_tuning: bad tuning");
    ctx->walk_budget = t->ccd_walk_budget;
    ctx->spec_levels = t->ccd_spec_levels;
    ctx->window_factor = t->ccd_window_factor;
    ctx->wide_queries = t->ccd_wide_queries;
    return RIGC_OK;
}

rigc_status rigc_set_detection_method(rigc_ctx* ctx, int32_t method) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (method != RIGC_DETECTION_HASH_GRID && method != RIGC_DETECTION_BVH)
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_set_detection_method: unknown method");
    if (method == RIGC_DETECTION_BVH) {
        if (!ctx->has_scene) return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_set_detection_method: no scene");
        if (!ctx->scene.bvh_ok)
            return ctx->fail(RIGC_ERR_INVALID_ARGUMENT,
                             "rigc_set_detection_method: the BVH broad phase needs every vertex on an edge and every "
                             "edge on a face");
        if (bvh_lds_bytes(ctx->scene) > 65536)
            return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_set_detection_method: a body is too large for the BVH "
                                                        "broad phase's LDS staging");
    }
    ctx->detection = method;
    return RIGC_OK;
}

// Coordinate descent over the CCD tunables on the caller's sample calls:
// for each knob in turn (speculative levels, walk budget, window factor,
// the query count from which a launch takes two blocks per CU),
// every value of its grid runs the whole sample (from the same per-scene
// window state) three times with the other knobs at their current choice,
// and the median device time (broad + narrow phase) decides; a value
// replaces the current one only when it is at least 5% faster (smaller
// differences are the timing's noise). Every run must return the first
// run's TOIs bit for bit. The choice is set on the context.
rigc_status rigc_autotune_earliest_toi(rigc_ctx* ctx, int32_t num_samples, const double* poses_t0,
                                       const double* poses_t1, double minimum_separation_distance, int32_t knobs,
                                       rigc_tuning* best, double* best_ms) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_autotune_earliest_toi: no scene");
    if (num_samples < 1 || !poses_t0 || !poses_t1 || !(minimum_separation_distance >= 0) || knobs < 0 || knobs > 15)
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_autotune_earliest_toi: bad arguments");
    if (knobs == 0) knobs = 15;
    const size_t stride = 6 * (size_t)ctx->scene.nb;
    const std::vector<double> window0 = ctx->last_toi;
    const rigc_tuning start{ ctx->walk_budget, ctx->spec_levels, ctx->window_factor, ctx->wide_queries };
    std::vector<double> ref, got(num_samples);
    auto restore = [&](const rigc_tuning& t) {
        ctx->walk_budget = t.ccd_walk_budget;
        ctx->spec_levels = t.ccd_spec_levels;
        ctx->window_factor = t.ccd_window_factor;
        ctx->wide_queries = t.ccd_wide_queries;
        ctx->last_toi = window0;
    };
    // median device time of the sample under t (3 runs); RIGC_OK or an error
    auto time_of = [&](const rigc_tuning& t, double& med) -> rigc_status {
        double ms[3];
        for (int rep = 0; rep < 3; rep++) {
            restore(t);
            ms[rep] = 0;
            for (int i = 0; i < num_samples; i++) {
                const rigc_status st = rigc_compute_earliest_toi(ctx, poses_t0 + stride * i, poses_t1 + stride * i,
                                                                 minimum_separation_distance, &got[i]);
                if (st != RIGC_OK) return st;
                ms[rep] += ctx->stats.broad_phase_ms + ctx->stats.narrow_phase_ms;
            }
            if (ref.empty()) ref = got;
            if (std::memcmp(ref.data(), got.data(), sizeof(double) * num_samples) != 0)
                return ctx->fail(RIGC_ERR_DEVICE, "rigc_autotune_earliest_toi: a tuning changed a result");
        }
        std::sort(ms, ms + 3);
        med = ms[1];
        return RIGC_OK;
    };
    rigc_tuning cur = start;
    double cur_ms = 0;
    rigc_status st = time_of(cur, cur_ms);
    for (int knob = 0; knob < 4 && st == RIGC_OK; knob++) {
        if (!(knobs & (1 << knob))) continue;  // RIGC_TUNE_* bit of this knob
        const double grid[4][3] = { { 3, 4, 5 }, { 1024, 4096, 16384 }, { 0.0, 4.0, 16.0 }, { 0, 4096, 65536 } };
        rigc_tuning pick = cur;
        double pick_ms = cur_ms;
        for (int g = 0; g < 3 && st == RIGC_OK; g++) {
            rigc_tuning t = cur;
            if (knob == 0) t.ccd_spec_levels = (int32_t)grid[0][g];
            else if (knob == 1) t.ccd_walk_budget = (int32_t)grid[1][g];
            else if (knob == 2) t.ccd_window_factor = grid[2][g];
            else t.ccd_wide_queries = (int64_t)grid[3][g];
            if (t.ccd_spec_levels == cur.ccd_spec_levels && t.ccd_walk_budget == cur.ccd_walk_budget &&
                t.ccd_window_factor == cur.ccd_window_factor && t.ccd_wide_queries == cur.ccd_wide_queries)
                continue;
            double ms = 0;
            st = time_of(t, ms);
            if (st == RIGC_OK && ms < 0.95 * pick_ms) {
                pick = t;
                pick_ms = ms;
            }
        }
        cur = pick;
        cur_ms = pick_ms;
    }
    if (st != RIGC_OK) {
        restore(start);
        return st;
    }
    restore(cur);
    if (best) *best = cur;
    if (best_ms) *best_ms = cur_ms;
    return RIGC_OK;
}

rigc_status rigc_narrow_phase_enqueue(rigc_ctx* ctx, void* stream, const double* poses_t0, const double* poses_t1,
                                      const int32_t* ee, int64_t n_ee, const int32_t* fv, int64_t n_fv,
                                      double toi_tolerance, double* earliest_toi) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_narrow_phase_enqueue: no scene");
    if (!poses_t0 || !poses_t1 || !earliest_toi || n_ee < 0 || n_fv < 0 || (n_ee && !ee) || (n_fv && !fv)
        || !(toi_tolerance > 0))
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_narrow_phase_enqueue: bad arguments");
    const DevScene& d = ctx->scene;
    const size_t nq = (size_t)(n_ee + n_fv), nq1 = std::max<size_t>(nq, 1);
    auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
    // counters [4], scene minima [n_scenes], per-query best / visits / flags
    const size_t o_sb = up(4 * sizeof(unsigned long long)), o_qb = o_sb + up(8 * (size_t)d.n_scenes);
    const size_t o_qv = o_qb + up(8 * nq1), o_qf = o_qv + up(4 * nq1), total = o_qf + up(4 * nq1);
    hipStream_t st = static_cast<hipStream_t>(stream);
    RIGC_HIP(hipSetDevice(ctx->device));
    if (ctx->enq_state.bytes < total) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        RIGC_HIP(hipStreamIsCapturing(st, &cs));
        if (cs != hipStreamCaptureStatusNone)
            return ctx->fail(RIGC_ERR_CAPACITY, "rigc_narrow_phase_enqueue: more queries than a call outside "
                                                "capture has sized the state for");
        RIGC_HIP(hipStreamSynchronize(st));  // work already queued may still use the old state
        RIGC_HIP(ctx->enq_state.ensure(total));
    }
    unsigned char* base = ctx->enq_state.as<unsigned char>();
    CcdArgs a = ccd_args(ctx);
    a.poses_t0 = poses_t0;
    a.poses_t1 = poses_t1;
    a.ee = ee;
    a.fv = fv;
    a.n_ee = n_ee;
    a.n_fv = n_fv;
    a.toi_tol = toi_tolerance;
    a.eval_budget = 0;
    a.use_global_best = 1;
    unsigned long long* cnt = reinterpret_cast<unsigned long long*>(base);
    a.counter = cnt + 1;
    a.total_evals = cnt + 2;
    a.n_tasks_out = cnt + 3;
    a.global_best = reinterpret_cast<unsigned long long*>(base + o_sb);
    a.q_best = reinterpret_cast<unsigned long long*>(base + o_qb);
    a.q_visits = reinterpret_cast<unsigned int*>(base + o_qv);
    a.q_flag = reinterpret_cast<int*>(base + o_qf);
    a.toi_out = nullptr;
    a.hit_out = nullptr;
    a.scene_skip = nullptr;
    a.scene_evals = nullptr;
    a.spec_levels = ctx->spec_levels;
    const int grid = ccd_grid(ctx->num_cus, (long long)nq, ctx->wide_queries);
    RIGC_HIP(enqueue_narrow_phase(a, grid, a.global_best, d.n_scenes, earliest_toi, st));
    return RIGC_OK;
}

rigc_status rigc_batch_set_previous_toi(rigc_ctx* ctx, int32_t n, const int32_t* scene_ids, const double* toi) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_batch_set_previous_toi: no scene");
    if (n < 0 || (n > 0 && (!scene_ids || !toi)))
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_batch_set_previous_toi: bad arguments");
    for (int i = 0; i < n; i++) {
        const int k = scene_ids[i];
        if (k < 0 || k >= ctx->scene.n_scenes || !(toi[i] >= 0))
            return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_batch_set_previous_toi: bad scene id or TOI");
    }
    for (int i = 0; i < n; i++) ctx->last_toi[scene_ids[i]] = toi[i];
    return RIGC_OK;
}

rigc_status rigc_batch_ccd_evaluations(rigc_ctx* ctx, int32_t n, const int32_t* scene_ids, int64_t* evals) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (n < 0 || (n > 0 && (!scene_ids || !evals)))
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_batch_ccd_evaluations: bad arguments");
    for (int i = 0; i < n; i++) {
        const int k = scene_ids[i];
        if (k < 0 || k >= ctx->scene.n_scenes)
            return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_batch_ccd_evaluations: scene id out of range");
        evals[i] = k < (int)ctx->h_scene_evals.size() ? (int64_t)ctx->h_scene_evals[k] : 0;
    }
    return RIGC_OK;
}

rigc_status rigc_batch_barrier_term(rigc_ctx* ctx, int32_t n, const int32_t* scene_ids, const double* build_poses,
                                    const double* poses, double dhat, double dmin, int32_t flags, double* energy,
                                    double* grad, rigc_block_hessian* hess, int64_t* counts,
                                    double* min_distance) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!build_poses || !poses || !energy || !(dhat > 0) || !(dmin >= 0) || bad_hess(flags, grad, hess))
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_batch_barrier_term: bad arguments");
    rigc_status s = select_scenes(ctx, n, scene_ids, "rigc_batch_barrier_term");
    if (s != RIGC_OK) return s;
    hipStream_t st = ctx->stream;
    RIGC_HIP(hipEventRecord(ctx->ev0, st));
    s = upload_sel_poses(ctx, ctx->sel_req, build_poses, nullptr);
    if (s == RIGC_OK) s = build_sel(ctx, ctx->sel_req, dhat, dmin);
    if (s != RIGC_OK) return s;
    if (counts) {
        std::vector<long long> c;
        RIGC_HIP(constraint_counts(ctx->as, ctx->scene, ctx->sel_req.sel, c, st));
        for (size_t k = 0; k < c.size(); k++) counts[k] = c[k];
    }
    if (poses != build_poses) {
        s = upload_sel_poses(ctx, ctx->sel_req, poses, nullptr);
        if (s != RIGC_OK) return s;
    }
    AssemblyOut out;
    RIGC_HIP(assemble_potential(ctx->as, ctx->scene, ctx->sel_req.sel, false, ctx->poses0.as<double>(), nullptr, dhat,
                                dmin, 0.0, nullptr, (flags & RIGC_GRADIENT) != 0, (flags & RIGC_HESSIAN) != 0,
                                ctx->num_cus, out, st));
    RIGC_HIP(hipEventRecord(ctx->ev1, st));
    s = copy_potential(ctx, ctx->sel_req, out, flags, grad, hess);
    if (!out.pending) std::memcpy(energy, out.energy.data(), sizeof(double) * n);
    if (s == RIGC_OK && min_distance) {
        std::vector<double> m2;
        RIGC_HIP(rigc::minimum_distance(ctx->as, ctx->scene, ctx->sel_req.sel, ctx->poses0.as<double>(), m2, st));
        for (int i = 0; i < n; i++) min_distance[i] = std::sqrt(m2[i]);
    }
    float ms = 0;
    RIGC_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.assembly_ms = ms;
    return s;
}

rigc_status rigc_batch_minimum_distance(rigc_ctx* ctx, int32_t n, const int32_t* scene_ids,
                                        const double* build_poses, const double* poses, double dhat, double dmin,
                                        double* min_distance) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!build_poses || !poses || !min_distance || !(dhat > 0) || !(dmin >= 0))
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_batch_minimum_distance: bad arguments");
    rigc_status s = select_scenes(ctx, n, scene_ids, "rigc_batch_minimum_distance");
    if (s != RIGC_OK) return s;
    s = upload_sel_poses(ctx, ctx->sel_req, build_poses, nullptr);
    if (s == RIGC_OK) s = build_sel(ctx, ctx->sel_req, dhat, dmin);
    if (s == RIGC_OK && poses != build_poses) s = upload_sel_poses(ctx, ctx->sel_req, poses, nullptr);
    if (s != RIGC_OK) return s;
    std::vector<double> m2;
    RIGC_HIP(rigc::minimum_distance(ctx->as, ctx->scene, ctx->sel_req.sel, ctx->poses0.as<double>(), m2,
                                    ctx->stream));
    for (int i = 0; i < n; i++) min_distance[i] = std::sqrt(m2[i]);
    return RIGC_OK;
}

rigc_status rigc_batch_detect_intersections(rigc_ctx* ctx, int32_t n, const int32_t* scene_ids,
                                            const double* poses, int64_t* counts) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!poses || !counts) return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_batch_detect_intersections: bad arguments");
    rigc_status s = select_scenes(ctx, n, scene_ids, "rigc_batch_detect_intersections");
    if (s == RIGC_OK) s = upload_sel_poses(ctx, ctx->sel_req, poses, nullptr);
    if (s != RIGC_OK) return s;
    RIGC_HIP(rigc::detect_intersections(ctx->as, ctx->bp, ctx->ix, ctx->scene, ctx->sel_req.sel,
                                        ctx->poses0.as<double>(), ctx->stream));
    const long long m = ctx->ix.n;
    std::vector<int32_t> pairs(2 * (size_t)m);
    if (m)
        RIGC_HIP(hipMemcpyAsync(pairs.data(), ctx->ix.out.p, sizeof(int32_t) * 2 * m, hipMemcpyDeviceToHost,
                                ctx->stream));
    RIGC_HIP(hipStreamSynchronize(ctx->stream));
    // the pairs are sorted by edge, and each scene's edges are one range
    const std::vector<int>& se = ctx->scene.h_scene_edge;
    long long q = 0;
    for (int i = 0; i < n; i++) {
        const int lo = se[scene_ids[i]], hi = se[scene_ids[i] + 1];
        while (q < m && pairs[2 * q] < lo) q++;
        long long c = 0;
        while (q < m && pairs[2 * q] < hi) {
            c++;
            q++;
        }
        counts[i] = c;
    }
    return RIGC_OK;
}

rigc_status rigc_batch_update_friction_constraints(rigc_ctx* ctx, int32_t n, const int32_t* scene_ids,
                                                   const double* poses, double dhat, double dmin,
                                                   const double* barrier_stiffness, const double* mu,
                                                   int64_t* counts) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!poses || !barrier_stiffness || !mu || !(dhat > 0) || !(dmin >= 0))
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_batch_update_friction_constraints: bad arguments");
    for (int i = 0; i < n; i++)
        if (!(barrier_stiffness[i] >= 0) || !(mu[i] >= 0))
            return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_batch_update_friction_constraints: bad kappa or mu");
    rigc_status s = select_scenes(ctx, n, scene_ids, "rigc_batch_update_friction_constraints");
    if (s != RIGC_OK) return s;
    s = upload_sel_poses(ctx, ctx->sel_req, poses, nullptr);
    if (s == RIGC_OK) s = build_sel(ctx, ctx->sel_req, dhat, dmin);
    if (s == RIGC_OK) s = upload_scene_par(ctx, 0, ctx->sel_req, barrier_stiffness);
    if (s == RIGC_OK) s = upload_scene_par(ctx, 1, ctx->sel_req, mu);
    if (s != RIGC_OK) return s;
    const int ns = ctx->scene.n_scenes;
    RIGC_HIP(build_friction_set(ctx->as, ctx->scene, ctx->sel_req.sel, ctx->poses0.as<double>(), dhat, dmin,
                                ctx->scene_par.as<double>(), ctx->scene_par.as<double>() + ns, ctx->stream));
    if (counts)
        for (int i = 0; i < n; i++) {
            const int k = scene_ids[i];
            counts[i] = ctx->as.fric_off[k + 1] - ctx->as.fric_off[k];
        }
    return RIGC_OK;
}

rigc_status rigc_batch_friction_term(rigc_ctx* ctx, int32_t n, const int32_t* scene_ids, const double* poses,
                                     const double* poses_t0, const double* epsv_times_h, int32_t flags,
                                     double* energy, double* grad, rigc_block_hessian* hess) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!poses || !poses_t0 || !epsv_times_h || !energy || bad_hess(flags, grad, hess))
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_batch_friction_term: bad arguments");
    for (int i = 0; i < n; i++)
        if (!(epsv_times_h[i] > 0))
            return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_batch_friction_term: bad epsv_times_h");
    rigc_status s = select_scenes(ctx, n, scene_ids, "rigc_batch_friction_term");
    if (s != RIGC_OK) return s;
    for (int i = 0; i < n; i++)
        if (!ctx->as.fric_valid[scene_ids[i]])
            return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_batch_friction_term: a scene has no friction set");
    s = upload_sel_poses(ctx, ctx->sel_req, poses, poses_t0);
    if (s == RIGC_OK) s = upload_scene_par(ctx, 2, ctx->sel_req, epsv_times_h);
    if (s != RIGC_OK) return s;
    hipStream_t st = ctx->stream;
    RIGC_HIP(hipEventRecord(ctx->ev0, st));
    AssemblyOut out;
    RIGC_HIP(assemble_potential(ctx->as, ctx->scene, ctx->sel_req.sel, true, ctx->poses0.as<double>(),
                                ctx->poses1.as<double>(), 0.0, 0.0, 0.0,
                                ctx->scene_par.as<double>() + 2 * (size_t)ctx->scene.n_scenes,
                                (flags & RIGC_GRADIENT) != 0, (flags & RIGC_HESSIAN) != 0, ctx->num_cus, out, st));
    RIGC_HIP(hipEventRecord(ctx->ev1, st));
    s = copy_potential(ctx, ctx->sel_req, out, flags, grad, hess);
    if (!out.pending) std::memcpy(energy, out.energy.data(), sizeof(double) * n);
    float ms = 0;
    RIGC_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->stats.assembly_ms = ms;
    return s;
}

}  // extern "C"

// ------------------------------------------------------------ piecewise-linear CCD
namespace {

// Runs the piecewise-linear narrow phase over nq queries whose candidate
// arrays are already on the device. Pass 1 gives every lane a small slice;
// queries that outgrow it run again with larger slices on fewer lanes.
// Leaves per-query results in ctx->toi / ctx->hit and returns them on the host.
rigc_status plccd_run(rigc_ctx* ctx, const int* d_ee, long long n_ee, const int* d_fv, long long n_fv,
                      double earliest_toi, double msd, std::vector<double>& tois, std::vector<signed char>& hits) {
    const DevScene& d = ctx->scene;
    const long long nq = n_ee + n_fv;
    hipStream_t st = ctx->stream;
    tois.assign((size_t)nq, HUGE_VAL);
    hits.assign((size_t)nq, 0);
    ctx->stats.box_evaluations = 0;
    ctx->stats.num_candidates = nq;
    ctx->stats.ccd_launches = 0;
    ctx->stats.ccd_kernel_ms = 0;
    if (nq == 0) return RIGC_OK;
    RIGC_HIP(ctx->toi.ensure(sizeof(double) * (size_t)nq));
    RIGC_HIP(ctx->hit.ensure((size_t)nq));
    RIGC_HIP(ctx->counters.ensure(4 * sizeof(unsigned long long)));
    RIGC_HIP(hipMemsetAsync(ctx->counters.p, 0, 4 * sizeof(unsigned long long), st));
    // (tstack, heap) capacities and lane counts of the passes; the last
    // slice holds any traversal (<= 1e6 refinements, two pushes each)
    struct Pass { int tcap; long long hcap; long long lanes; };
    // pass 0: one query per wavefront (64-lane block) in LDS (stack 64,
    // heap 512), 16 blocks per CU requested
    const Pass normal[3] = { { 64, 512, (long long)ctx->num_cus * 16 }, { 2048, 1LL << 16, 256 },
                             { 8192, 1LL << 21, 8 } };
    // test knob: every pass in HBM slices, the first two tiny
    const Pass small[3] = { { 2, 8, (long long)ctx->num_cus * 16 }, { 4, 64, 256 }, { 8192, 1LL << 21, 8 } };
    const Pass* passes = ctx->pl_small_slices ? small : normal;
    PlccdArgs a{};
    a.V = d.V; a.E = d.E; a.F = d.F;
    a.v_body = d.v_body; a.e_body = d.e_body; a.f_body = d.f_body;
    a.poses_t0 = ctx->poses0.as<double>();
    a.poses_t1 = ctx->poses1.as<double>();
    a.ee = d_ee; a.fv = d_fv; a.n_ee = n_ee;
    a.earliest_toi = earliest_toi;
    a.msd = msd;
    a.toi_out = ctx->toi.as<double>();
    a.hit_out = ctx->hit.as<signed char>();
    a.counters = ctx->counters.as<unsigned long long>();
    std::vector<int> redo;
    float total_ms = 0;
    for (int p = 0; p < 3; p++) {
        const long long n_run = p == 0 ? nq : (long long)redo.size();
        if (n_run == 0) break;
        const bool in_lds = p == 0 && !ctx->pl_small_slices;
        const long long lanes = in_lds ? std::min<long long>(passes[p].lanes, n_run)
                                       : std::min<long long>(passes[p].lanes, (n_run + 63) / 64 * 64);
        if (!in_lds) {
            const size_t slice = (size_t)passes[p].tcap + 7 * (size_t)passes[p].hcap;
            RIGC_HIP(ctx->pl_arena.ensure(sizeof(double) * slice * (size_t)lanes));
        }
        a.arena = ctx->pl_arena.as<double>();
        a.n_lanes = lanes;
        a.tstack_cap = passes[p].tcap;
        a.heap_cap = passes[p].hcap;
        a.n_queries = n_run;
        a.query_ids = nullptr;
        if (p > 0) {
            RIGC_HIP(ctx->pl_ids.ensure(sizeof(int) * redo.size()));
            RIGC_HIP(hipMemcpyAsync(ctx->pl_ids.p, redo.data(), sizeof(int) * redo.size(), hipMemcpyHostToDevice, st));
            a.query_ids = ctx->pl_ids.as<int>();
        }
        RIGC_HIP(hipEventRecord(ctx->ev0, st));
        if (in_lds) RIGC_HIP(launch_plccd_wave(a, (int)lanes, st));
        else RIGC_HIP(launch_plccd(a, (int)((lanes + 255) / 256), st));
        RIGC_HIP(hipEventRecord(ctx->ev1, st));
        RIGC_HIP(hipMemcpyAsync(hits.data(), ctx->hit.p, (size_t)nq, hipMemcpyDeviceToHost, st));
        RIGC_HIP(hipStreamSynchronize(st));
        float ms = 0;
        RIGC_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        total_ms += ms;
        ctx->stats.ccd_launches++;
        redo.clear();
        for (long long q = 0; q < nq; q++)
            if (hits[(size_t)q] == 2) redo.push_back((int)q);
    }
    if (!redo.empty())
        return ctx->fail(RIGC_ERR_UNRESOLVED, "piecewise-linear CCD: a query outgrew the largest lane slice");
    unsigned long long cnt[2] = { 0, 0 };
    RIGC_HIP(hipMemcpyAsync(tois.data(), ctx->toi.p, sizeof(double) * (size_t)nq, hipMemcpyDeviceToHost, st));
    RIGC_HIP(hipMemcpyAsync(cnt, ctx->counters.p, sizeof(cnt), hipMemcpyDeviceToHost, st));
    RIGC_HIP(hipStreamSynchronize(st));
    ctx->stats.box_evaluations = (int64_t)cnt[0];
    ctx->stats.ccd_kernel_ms = total_ms;
    ctx->stats.narrow_phase_ms = total_ms;
    return RIGC_OK;
}

}  // namespace

rigc_status rigc_narrow_phase_piecewise_linear(rigc_ctx* ctx, const double* poses_t0, const double* poses_t1,
                                               const int32_t* ee, int64_t n_ee, const int32_t* fv, int64_t n_fv,
                                               double earliest_toi_in, double minimum_separation_distance,
                                               double* earliest_toi, double* toi_out, int8_t* hit_out) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_narrow_phase_piecewise_linear: no scene");
    if (!poses_t0 || !poses_t1 || !earliest_toi || n_ee < 0 || n_fv < 0 || (n_ee && !ee) || (n_fv && !fv)
        || !(earliest_toi_in > 0 && earliest_toi_in <= 1) || !(minimum_separation_distance >= 0)
        || n_ee + n_fv > (int64_t)INT32_MAX)
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_narrow_phase_piecewise_linear: bad arguments");
    const DevScene& d = ctx->scene;
    for (int64_t i = 0; i < n_ee; i++)
        if (ee[2 * i] < 0 || ee[2 * i] >= d.ne || ee[2 * i + 1] < 0 || ee[2 * i + 1] >= d.ne)
            return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_narrow_phase_piecewise_linear: edge id out of range");
    for (int64_t i = 0; i < n_fv; i++)
        if (fv[2 * i] < 0 || fv[2 * i] >= d.nf || fv[2 * i + 1] < 0 || fv[2 * i + 1] >= d.nv)
            return ctx->fail(RIGC_ERR_INVALID_ARGUMENT,
                             "rigc_narrow_phase_piecewise_linear: face/vertex id out of range");
    RIGC_HIP(hipSetDevice(ctx->device));
    rigc_status s = upload_poses(ctx, poses_t0, poses_t1);
    if (s != RIGC_OK) return s;
    hipStream_t st = ctx->stream;
    RIGC_HIP(ctx->ee.ensure(2 * sizeof(int32_t) * std::max<size_t>((size_t)n_ee, 1)));
    RIGC_HIP(ctx->fv.ensure(2 * sizeof(int32_t) * std::max<size_t>((size_t)n_fv, 1)));
    if (n_ee) RIGC_HIP(hipMemcpyAsync(ctx->ee.p, ee, 2 * sizeof(int32_t) * n_ee, hipMemcpyHostToDevice, st));
    if (n_fv) RIGC_HIP(hipMemcpyAsync(ctx->fv.p, fv, 2 * sizeof(int32_t) * n_fv, hipMemcpyHostToDevice, st));
    std::vector<double> tois;
    std::vector<signed char> hits;
    s = plccd_run(ctx, ctx->ee.as<int>(), n_ee, ctx->fv.as<int>(), n_fv, earliest_toi_in, minimum_separation_distance,
                  tois, hits);
    if (s != RIGC_OK) return s;
    double g = HUGE_VAL;
    for (size_t q = 0; q < tois.size(); q++) {
        if (hits[q] == 1) g = std::min(g, tois[q]);
        if (toi_out) toi_out[q] = tois[q];
        if (hit_out) hit_out[q] = (int8_t)hits[q];
    }
    *earliest_toi = g;
    return RIGC_OK;
}

rigc_status rigc_compute_earliest_toi_piecewise_linear(rigc_ctx* ctx, const double* poses_t0, const double* poses_t1,
                                                       double minimum_separation_distance, double* earliest_toi) {
    if (!ctx) return RIGC_ERR_INVALID_ARGUMENT;
    if (!ctx->has_scene) return ctx->fail(RIGC_ERR_NO_SCENE, "rigc_compute_earliest_toi_piecewise_linear: no scene");
    if (!poses_t0 || !poses_t1 || !earliest_toi || !(minimum_separation_distance >= 0))
        return ctx->fail(RIGC_ERR_INVALID_ARGUMENT, "rigc_compute_earliest_toi_piecewise_linear: bad arguments");
    RIGC_HIP(hipSetDevice(ctx->device));
    rigc_status s = upload_poses(ctx, poses_t0, poses_t1);
    if (s != RIGC_OK) return s;
    hipStream_t st = ctx->stream;
    if ((s = check_detection(ctx, "rigc_compute_earliest_toi_piecewise_linear")) != RIGC_OK) return s;
    RIGC_HIP(hipEventRecord(ctx->ev0, st));
    if (ctx->detection == RIGC_DETECTION_BVH)
        RIGC_HIP(run_broad_phase_bvh(ctx->bp, ctx->scene, ctx->sel_all.sel, ctx->poses0.as<double>(),
                                     ctx->poses1.as<double>(), minimum_separation_distance / 2.0, true, false,
                                     ctx->num_cus, st));
    else
        RIGC_HIP(run_broad_phase(ctx->bp, ctx->scene, ctx->sel_all.sel, ctx->poses0.as<double>(),
                                 ctx->poses1.as<double>(), minimum_separation_distance / 2.0, true, false, st));
    RIGC_HIP(hipEventRecord(ctx->ev1, st));
    RIGC_HIP(hipStreamSynchronize(st));
    float bp_ms = 0;
    RIGC_HIP(hipEventElapsedTime(&bp_ms, ctx->ev0, ctx->ev1));
    const long long n_ee = ctx->bp.n_ee, n_fv = ctx->bp.n_fv;
    if (n_ee + n_fv > (long long)INT32_MAX)
        return ctx->fail(RIGC_ERR_CAPACITY, "rigc_compute_earliest_toi_piecewise_linear: too many candidates");
    std::vector<double> tois;
    std::vector<signed char> hits;
    s = plccd_run(ctx, ctx->bp.out_ee.as<int>(), n_ee, ctx->bp.out_fv.as<int>(), n_fv, 1.0,
                  minimum_separation_distance, tois, hits);
    if (s != RIGC_OK) return s;
    ctx->stats.broad_phase_ms = bp_ms;
    double g = HUGE_VAL;
    for (size_t q = 0; q < tois.size(); q++)
        if (hits[q] == 1) g = std::min(g, tois[q]);
    *earliest_toi = g;
    return RIGC_OK;
}
