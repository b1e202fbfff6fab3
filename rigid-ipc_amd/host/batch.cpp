// Batched time stepping of many independent scenes (rigs.h, rigs_batch_*).
//
// Fused mode: every scene runs the host stepper (simulation.cpp, unchanged)
// on a thread of its own, over a proxy contact backend. A proxy call parks
// its thread with a request. Executor threads, each with its own
// multi-scene device context (and so its own HIP stream), take the waiting
// requests and execute them as fused device calls (one rigc_batch_* call
// per kind of request and argument set: earliest TOI, barrier, minimum
// distance, friction update, friction potential), hand each scene its part
// of the results and wake the scenes. An executor takes the requests once
// every scene still stepping is parked (waiting, or served by another
// executor), or once the oldest waiting request has waited a gather window.
// So a fused call that runs long (a scene's deep bisection) holds only the
// executor serving it: the requests that come meanwhile go to another
// executor, on another stream, and their kernels run beside it on the idle
// part of the device. The friction kinds go to one executor of their own, whose context
// holds the lagged friction sets; every other kind of request carries all
// its inputs and may go to any general executor. Each scene's Newton loop
// proceeds independently (different iteration counts, line-search lengths
// and step ends). Per-scene results are bit-identical to a single-scene
// context, whatever the grouping and whichever context serves a call, so
// every scene follows exactly the trajectory it follows alone.
//
// Sequential mode: the plain path, one scene after another, each with its
// own single-scene context (the comparison point for the fused mode).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "rigs.h"

namespace rigs {
// host/simulation.cpp: whether a stepper computes the body energy on a
// helper thread beside its device calls, and the hooks its friction helper
// calls around a call it makes beside the stepper's own
void set_host_overlap(rigs_sim* sim, bool on);
void set_fork_hooks(rigs_sim* sim, void (*join)(void*), void (*part)(void*), void* arg);
}  // namespace rigs

namespace {

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

enum ReqKind { REQ_TOI = 0, REQ_BARRIER, REQ_MIND, REQ_FUPDATE, REQ_FRICTION, REQ_ISECT };

struct Request {
    int kind = 0;
    int scene = 0;
    // inputs (owned by the parked caller)
    const double* p0 = nullptr;
    const double* p1 = nullptr;
    double a = 0, b = 0, c = 0;  // TOI: msd; BARRIER / MIND: dhat, dmin; FRICTION: epsv_h; FUPDATE: dhat, dmin, kappa
    double d = 0;                // FUPDATE: mu
    int32_t flags = 0;
    // outputs
    double* value = nullptr;     // toi / energy / min distance
    double* mind = nullptr;      // BARRIER: also the minimum distance (optional)
    double* grad = nullptr;
    rigc_block_hessian* hess = nullptr;
    int64_t* count = nullptr;
    rigc_status status = RIGC_OK;
    std::string error;
    bool done = false;
    double t_submit = 0;         // when it was queued (ms)
};

bool is_friction_kind(int kind) { return kind == REQ_FUPDATE || kind == REQ_FRICTION; }

struct Batch;

struct Proxy {
    Batch* B = nullptr;
    int scene = 0;
    int nb = 0;
    std::vector<double> build_poses;  // poses of the scene's last construct_constraint_set
    int64_t counts[4] = { 0, 0, 0, 0 };
    // minimum distance returned with the last barrier evaluated at its
    // build poses (the stepper asks for it there after accepting a step)
    std::vector<double> mind_poses;
    double mind = 0, mind_new = 0, mind_dhat = 0, mind_dmin = 0;
    bool mind_valid = false;
    std::string err;
};

rigc_status submit(Proxy* P, Request& r);

// ------------------------------------------------------------ proxy backend
rigc_status px_set_scene(void*, const rigc_scene*) { return RIGC_OK; }

rigc_status px_toi(void* s, const double* p0, const double* p1, double msd, double* toi) {
    Proxy* P = static_cast<Proxy*>(s);
    Request r;
    r.kind = REQ_TOI;
    r.p0 = p0;
    r.p1 = p1;
    r.a = msd;
    r.value = toi;
    return submit(P, r);
}

// The set is built lazily, by the call that uses it: remember the poses.
rigc_status px_build(void* s, const double* poses, double, double, int64_t* counts) {
    Proxy* P = static_cast<Proxy*>(s);
    P->build_poses.assign(poses, poses + 6 * (size_t)P->nb);
    if (counts)
        for (int k = 0; k < 4; k++) counts[k] = P->counts[k];  // of the previous barrier call
    return RIGC_OK;
}

rigc_status px_barrier(void* s, const double* poses, double dhat, double dmin, int32_t flags, double* e, double* g,
                       rigc_block_hessian* h) {
    Proxy* P = static_cast<Proxy*>(s);
    if (flags & RIGC_BUILD_CONSTRAINTS) P->build_poses.assign(poses, poses + 6 * (size_t)P->nb);
    if (P->build_poses.empty()) {
        P->err = "no constraint set (build one first)";
        return RIGC_ERR_NO_SCENE;
    }
    Request r;
    r.kind = REQ_BARRIER;
    r.p0 = P->build_poses.data();
    r.p1 = poses;
    r.a = dhat;
    r.b = dmin;
    r.flags = flags & (RIGC_GRADIENT | RIGC_HESSIAN);
    r.value = e;
    r.grad = g;
    r.hess = h;
    const size_t bytes = sizeof(double) * 6 * (size_t)P->nb;
    const bool at_build = std::memcmp(P->build_poses.data(), poses, bytes) == 0;
    P->mind_valid = false;
    if (at_build) r.mind = &P->mind_new;
    rigc_status st = submit(P, r);
    if (st == RIGC_OK && at_build) {
        P->mind = P->mind_new;
        P->mind_poses.assign(poses, poses + 6 * (size_t)P->nb);
        P->mind_dhat = dhat;
        P->mind_dmin = dmin;
        P->mind_valid = true;
    }
    return st;
}

rigc_status px_mind(void* s, const double* poses, double dhat, double dmin, int32_t flags, double* out) {
    Proxy* P = static_cast<Proxy*>(s);
    if (flags & RIGC_BUILD_CONSTRAINTS) P->build_poses.assign(poses, poses + 6 * (size_t)P->nb);
    const size_t bytes = sizeof(double) * 6 * (size_t)P->nb;
    if (P->mind_valid && dhat == P->mind_dhat && dmin == P->mind_dmin && !P->build_poses.empty() &&
        std::memcmp(P->mind_poses.data(), poses, bytes) == 0 &&
        std::memcmp(P->mind_poses.data(), P->build_poses.data(), bytes) == 0) {
        *out = P->mind;  // same set, same poses: the barrier call's answer
        return RIGC_OK;
    }
    if (P->build_poses.empty()) {
        P->err = "no constraint set (build one first)";
        return RIGC_ERR_NO_SCENE;
    }
    Request r;
    r.kind = REQ_MIND;
    r.p0 = P->build_poses.data();
    r.p1 = poses;
    r.a = dhat;
    r.b = dmin;
    r.value = out;
    return submit(P, r);
}

rigc_status px_fupdate(void* s, const double* poses, double dhat, double dmin, double kappa, double mu, int32_t flags,
                       int64_t* n) {
    Proxy* P = static_cast<Proxy*>(s);
    if ((flags & RIGC_BUILD_CONSTRAINTS) || P->build_poses.empty())
        P->build_poses.assign(poses, poses + 6 * (size_t)P->nb);
    Request r;
    r.kind = REQ_FUPDATE;
    r.p0 = P->build_poses.data();
    r.a = dhat;
    r.b = dmin;
    r.c = kappa;
    r.d = mu;
    r.count = n;
    return submit(P, r);
}

rigc_status px_friction(void* s, const double* poses, const double* poses_t0, double eh, int32_t flags, double* e,
                        double* g, rigc_block_hessian* h) {
    Proxy* P = static_cast<Proxy*>(s);
    Request r;
    r.kind = REQ_FRICTION;
    r.p0 = poses;
    r.p1 = poses_t0;
    r.c = eh;
    r.flags = flags & (RIGC_GRADIENT | RIGC_HESSIAN);
    r.value = e;
    r.grad = g;
    r.hess = h;
    return submit(P, r);
}

const char* px_error(void* s) { return static_cast<Proxy*>(s)->err.c_str(); }

// The executors' contexts were given the batch's detection method at
// creation; a scene asks for the same one.
rigc_status px_detection(void* s, int32_t method);

// ------------------------------------------------------------ combiner
// One executor: a device context holding every scene, the thread that makes
// its fused calls, and that thread's staging.
struct Exec {
    Batch* B = nullptr;
    rigc_ctx* ctx = nullptr;
    bool friction = false;  // serves the friction kinds
    std::thread th;
    std::vector<int32_t> ids;
    std::vector<double> c0, c1, vals, kap, mus, grad, diag, off;
    std::vector<int32_t> pairs;
    std::vector<int64_t> cnt;
    int64_t cap = 0;
    rigs_batch_stats st{};  // this executor's share of the batch statistics

    void loop();
    void execute(std::vector<Request*>& reqs);
    void run_group(int kind, std::vector<Request*>& g);
};

struct Batch {
    int n = 0;
    std::vector<int> nb;
    std::vector<Proxy> proxies;
    std::vector<std::unique_ptr<Exec>> execs;  // general executors, then the friction executor (if any)
    bool has_friction_exec = false;
    std::mutex m;
    std::condition_variable cv;       // callers: a request is done
    std::condition_variable cv_exec;  // executors: a request came, a caller left, or stop
    std::vector<Request*> pending;
    int in_flight = 0;  // requests taken by an executor and not yet done
    int active = 0;     // callers taking part (stepping scenes and their helpers)
    bool stop = false;
    double gather_ms = 0.2;
    std::vector<double> last_toi;  // each scene's previous earliest TOI (any executor)
    int detection = RIGC_DETECTION_HASH_GRID;  // broad phase of every scene's CCD calls
    rigc_tuning tuning{};  // what the autotuner chose (zeros when it did not run)
    double tune_ms = 0;    // its device time over the sample

    bool serves(const Exec& e, const Request* r) const {
        return has_friction_exec ? (is_friction_kind(r->kind) == e.friction) : true;
    }
    void start();
    void shutdown();
};

rigc_status px_detection(void* s, int32_t method) {
    Proxy* P = static_cast<Proxy*>(s);
    if (method == P->B->detection) return RIGC_OK;
    P->err = "the batch's scenes must share one detection method";
    return RIGC_ERR_INVALID_ARGUMENT;
}

rigc_status submit(Proxy* P, Request& r) {
    Batch& B = *P->B;
    r.scene = P->scene;
    std::unique_lock<std::mutex> lk(B.m);
    r.t_submit = now_ms();
    B.pending.push_back(&r);
    B.cv_exec.notify_all();
    B.cv.wait(lk, [&] { return r.done; });
    if (r.status != RIGC_OK) P->err = r.error;
    return r.status;
}

// Intersection check: in a round like the other calls while the scenes
// step; outside stepping (the check of the initial state when a scene's
// stepper is created, before any round can run) a direct one-scene call on
// the first executor's context, which no executor uses meanwhile.
rigc_status px_isect(void* s, const double* poses, int64_t* n) {
    Proxy* P = static_cast<Proxy*>(s);
    Batch& B = *P->B;
    {
        std::lock_guard<std::mutex> lk(B.m);
        if (B.active == 0) {
            const int32_t id = P->scene;
            rigc_ctx* ctx = B.execs[0]->ctx;
            const rigc_status st = rigc_batch_detect_intersections(ctx, 1, &id, poses, n);
            if (st != RIGC_OK) P->err = rigc_last_error(ctx);
            return st;
        }
    }
    Request r;
    r.kind = REQ_ISECT;
    r.p0 = poses;
    r.count = n;
    return submit(P, r);
}

// A scene finished stepping (or failed), or one of its two callers is done
// or waiting for the other: one caller fewer takes part in rounds.
void leave(Batch& B) {
    std::lock_guard<std::mutex> lk(B.m);
    B.active--;
    B.cv_exec.notify_all();
}

// A stepper's friction helper is about to call beside the stepper's own
// thread (or that thread resumes): one caller more, so that a round does
// not start without its request unless the gather window has passed.
void fork_join(void* arg) {
    Batch& B = *static_cast<Batch*>(arg);
    std::lock_guard<std::mutex> lk(B.m);
    B.active++;
}
void fork_part(void* arg) { leave(*static_cast<Batch*>(arg)); }

// Executor thread: take this executor's kind of waiting requests once every
// caller is parked or the oldest has waited the gather window, execute them
// outside the lock, and wake each group's callers as soon as its call
// returns.
void Exec::loop() {
    Batch& b = *B;
    std::unique_lock<std::mutex> lk(b.m);
    for (;;) {
        double oldest = -1;
        for (const Request* r : b.pending)
            if (b.serves(*this, r) && (oldest < 0 || r->t_submit < oldest)) oldest = r->t_submit;
        if (oldest < 0) {
            if (b.stop) return;
            b.cv_exec.wait(lk);
            continue;
        }
        const bool all_parked = (int)b.pending.size() + b.in_flight >= b.active;
        if (!all_parked && !b.stop && b.gather_ms != 0) {
            if (b.gather_ms < 0) {
                b.cv_exec.wait(lk);
                continue;
            }
            const double wait_ms = oldest + b.gather_ms - now_ms();
            if (wait_ms > 0) {
                b.cv_exec.wait_for(lk, std::chrono::duration<double, std::milli>(wait_ms));
                continue;  // look again: another executor may have taken them
            }
        }
        std::vector<Request*> reqs, keep;
        for (Request* r : b.pending) (b.serves(*this, r) ? reqs : keep).push_back(r);
        b.pending.swap(keep);
        b.in_flight += (int)reqs.size();
        lk.unlock();
        const double t0 = now_ms();
        execute(reqs);
        st.device_ms += now_ms() - t0;
        st.rounds++;
        lk.lock();
    }
}

// Fused calls for reqs, one per kind and shared scalar arguments; each
// group's callers are woken when its call returns.
void Exec::execute(std::vector<Request*>& reqs) {
    std::map<std::tuple<int, int32_t, double, double>, std::vector<Request*>> groups;
    for (Request* r : reqs) {
        const double k1 = r->kind == REQ_TOI ? r->a : (r->kind == REQ_FRICTION ? 0.0 : r->a);
        const double k2 = (r->kind == REQ_BARRIER || r->kind == REQ_MIND || r->kind == REQ_FUPDATE) ? r->b : 0.0;
        groups[std::make_tuple(r->kind, r->flags, k1, k2)].push_back(r);
    }
    for (auto& kv : groups) {
        std::vector<Request*>& g = kv.second;
        std::sort(g.begin(), g.end(), [](const Request* x, const Request* y) { return x->scene < y->scene; });
        run_group(std::get<0>(kv.first), g);
        st.fused_calls++;
        st.requests += (int64_t)g.size();
        std::lock_guard<std::mutex> lk(B->m);
        B->in_flight -= (int)g.size();
        for (Request* r : g) r->done = true;
        B->cv.notify_all();
        B->cv_exec.notify_all();
    }
}

void Batch::start() {
    for (auto& e : execs) e->th = std::thread([p = e.get()] { p->loop(); });
}

void Batch::shutdown() {
    {
        std::lock_guard<std::mutex> lk(m);
        stop = true;
    }
    cv_exec.notify_all();
    for (auto& e : execs)
        if (e->th.joinable()) e->th.join();
    for (auto& e : execs)
        if (e->ctx) {
            rigc_destroy(e->ctx);
            e->ctx = nullptr;
        }
}

void fail_all(std::vector<Request*>& g, rigc_status s, const char* msg) {
    for (Request* r : g) {
        r->status = s;
        r->error = msg;
    }
}

void Exec::run_group(int kind, std::vector<Request*>& g) {
    const int k = (int)g.size();
    ids.resize(k);
    std::vector<int> start(k + 1, 0);
    for (int i = 0; i < k; i++) {
        ids[i] = g[i]->scene;
        start[i + 1] = start[i] + B->nb[g[i]->scene];
    }
    const int tot = start[k];
    auto concat = [&](std::vector<double>& dst, bool second) {
        dst.resize(6 * (size_t)tot);
        for (int i = 0; i < k; i++) {
            const double* src = second ? g[i]->p1 : g[i]->p0;
            std::memcpy(&dst[6 * (size_t)start[i]], src, sizeof(double) * 6 * B->nb[g[i]->scene]);
        }
    };
    vals.assign(k, 0.0);
    rigc_status s = RIGC_OK;
    if (kind == REQ_TOI) {
        concat(c0, false);
        concat(c1, true);
        // each scene's previous TOI, from whichever executor served it
        kap.resize(k);
        {
            std::lock_guard<std::mutex> lk(B->m);
            for (int i = 0; i < k; i++) kap[i] = B->last_toi[ids[i]];
        }
        s = rigc_batch_set_previous_toi(ctx, k, ids.data(), kap.data());
        if (s == RIGC_OK)
            s = rigc_batch_earliest_toi(ctx, k, ids.data(), c0.data(), c1.data(), g[0]->a, vals.data());
        if (s == RIGC_OK) {
            rigc_stats cs{};
            rigc_get_stats(ctx, &cs);
            st.ccd_candidates += cs.num_candidates;
            st.ccd_kernel_ms += cs.ccd_kernel_ms;
            st.ccd_launches += cs.ccd_launches;
            st.ccd_box_evaluations += cs.box_evaluations;
            st.ccd_calls++;
            for (int i = 0; i < k; i++) *g[i]->value = vals[i];
            std::lock_guard<std::mutex> lk(B->m);
            for (int i = 0; i < k; i++) B->last_toi[ids[i]] = vals[i];
        }
    } else if (kind == REQ_MIND) {
        concat(c0, false);
        concat(c1, true);
        s = rigc_batch_minimum_distance(ctx, k, ids.data(), c0.data(), c1.data(), g[0]->a, g[0]->b, vals.data());
        if (s == RIGC_OK)
            for (int i = 0; i < k; i++) *g[i]->value = vals[i];
    } else if (kind == REQ_ISECT) {
        concat(c0, false);
        cnt.assign(k, 0);
        s = rigc_batch_detect_intersections(ctx, k, ids.data(), c0.data(), cnt.data());
        if (s == RIGC_OK)
            for (int i = 0; i < k; i++) *g[i]->count = cnt[i];
    } else if (kind == REQ_FUPDATE) {
        concat(c0, false);
        kap.resize(k);
        mus.resize(k);
        for (int i = 0; i < k; i++) {
            kap[i] = g[i]->c;
            mus[i] = g[i]->d;
        }
        cnt.assign(k, 0);
        s = rigc_batch_update_friction_constraints(ctx, k, ids.data(), c0.data(), g[0]->a, g[0]->b, kap.data(),
                                                   mus.data(), cnt.data());
        if (s == RIGC_OK)
            for (int i = 0; i < k; i++)
                if (g[i]->count) *g[i]->count = cnt[i];
    } else {
        // barrier or friction potential
        const int32_t flags = g[0]->flags;
        const bool want_hess = (flags & RIGC_HESSIAN) != 0;
        const bool want_grad = want_hess || (flags & RIGC_GRADIENT) != 0;
        concat(c0, false);
        concat(c1, true);
        if (want_grad) grad.assign(6 * (size_t)tot, 0.0);
        std::vector<double> eh;
        if (kind == REQ_FRICTION) {
            eh.resize(k);
            for (int i = 0; i < k; i++) eh[i] = g[i]->c;
        }
        cnt.assign(4 * (size_t)k, 0);
        bool any_mind = false;
        for (Request* r : g) any_mind = any_mind || r->mind;
        std::vector<double> md(any_mind ? k : 0);
        rigc_block_hessian h{};
        for (int attempt = 0; attempt < 2; attempt++) {
            if (want_hess) {
                diag.resize(36 * (size_t)tot);
                pairs.resize(2 * (size_t)std::max<int64_t>(cap, 1));
                off.resize(36 * (size_t)std::max<int64_t>(cap, 1));
                h = rigc_block_hessian{ diag.data(), pairs.data(), off.data(), cap, 0 };
            }
            if (kind == REQ_BARRIER)
                s = rigc_batch_barrier_term(ctx, k, ids.data(), c0.data(), c1.data(), g[0]->a, g[0]->b, flags,
                                            vals.data(), want_grad ? grad.data() : nullptr,
                                            want_hess ? &h : nullptr, cnt.data(), any_mind ? md.data() : nullptr);
            else
                s = rigc_batch_friction_term(ctx, k, ids.data(), c0.data(), c1.data(), eh.data(), flags, vals.data(),
                                             want_grad ? grad.data() : nullptr, want_hess ? &h : nullptr);
            if (s == RIGC_ERR_CAPACITY && want_hess && h.count > cap) {
                cap = h.count;
                continue;
            }
            break;
        }
        if (s == RIGC_OK) {
            int64_t q = 0;  // pairs are sorted by first body: contiguous per scene
            for (int i = 0; i < k; i++) {
                Request* r = g[i];
                *r->value = vals[i];
                if (kind == REQ_BARRIER) {
                    Proxy& P = B->proxies[r->scene];
                    for (int c = 0; c < 4; c++) P.counts[c] = cnt[4 * i + c];
                    if (r->mind) *r->mind = md[i];
                }
                if (want_grad && r->grad)
                    std::memcpy(r->grad, &grad[6 * (size_t)start[i]], sizeof(double) * 6 * B->nb[r->scene]);
                if (want_hess && r->hess) {
                    std::memcpy(r->hess->diag, &diag[36 * (size_t)start[i]], sizeof(double) * 36 * B->nb[r->scene]);
                    int64_t q1 = q;
                    while (q1 < h.count && pairs[2 * q1] < start[i + 1]) q1++;
                    const int64_t m = q1 - q;
                    r->hess->count = m;
                    if (m > r->hess->capacity) {
                        r->status = RIGC_ERR_CAPACITY;
                        r->error = "off-diagonal Hessian capacity too small";
                    } else {
                        for (int64_t e = 0; e < m; e++) {
                            r->hess->pairs[2 * e] = pairs[2 * (q + e)] - start[i];
                            r->hess->pairs[2 * e + 1] = pairs[2 * (q + e) + 1] - start[i];
                        }
                        if (m) std::memcpy(r->hess->off, &off[36 * (size_t)q], sizeof(double) * 36 * m);
                    }
                    q = q1;
                }
            }
            return;
        }
    }
    if (s != RIGC_OK) fail_all(g, s, rigc_last_error(ctx));
}

}  // namespace

struct rigs_batch {
    int mode = RIGS_BATCH_FUSED;
    int n = 0;
    Batch B;
    std::vector<rigs_sim*> sims;
    std::vector<rigs_backend> tables;
    std::vector<std::string> errors;
    std::string err;
};

extern "C" {

rigc_status rigs_batch_create(rigs_batch** out, int32_t device, int32_t num_scenes, const rigc_scene* scenes,
                              const rigs_bodies* bodies, const rigs_params* params, int32_t mode) {
    if (!out || num_scenes < 1 || !scenes || !bodies || !params
        || (mode != RIGS_BATCH_FUSED && mode != RIGS_BATCH_SEQUENTIAL))
        return RIGC_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (mode == RIGS_BATCH_FUSED)
        for (int i = 0; i < num_scenes; i++)
            if (params[i].trajectory_type != RIGS_TRAJECTORY_RIGID)
                return RIGC_ERR_INVALID_ARGUMENT;  // the fused rounds serve the rigid CCD only
    rigs_batch* b = new rigs_batch();
    b->mode = mode;
    b->n = num_scenes;
    b->sims.assign(num_scenes, nullptr);
    b->errors.assign(num_scenes, "");
    rigc_status s = RIGC_OK;
    if (mode == RIGS_BATCH_FUSED) {
        Batch& B = b->B;
        B.n = num_scenes;
        // general executors (RIGS_BATCH_EXECUTORS, default 3), plus one for
        // the friction kinds when a scene has friction
        int n_general = 3;
        if (const char* e = std::getenv("RIGS_BATCH_EXECUTORS")) n_general = std::max(1, std::min(8, std::atoi(e)));
        if (const char* g = std::getenv("RIGS_BATCH_GATHER_US")) B.gather_ms = std::atof(g) * 1e-3;
        bool any_friction = false;
        for (int i = 0; i < num_scenes; i++) any_friction = any_friction || params[i].coefficient_friction > 0;
        B.has_friction_exec = any_friction;
        B.detection = params[0].detection_method;
        for (int i = 1; i < num_scenes; i++)
            if (params[i].detection_method != B.detection) {
                delete b;
                return RIGC_ERR_INVALID_ARGUMENT;  // one detection method per batch (the contexts hold it)
            }
        B.last_toi.assign(num_scenes, HUGE_VAL);
        const int n_exec = n_general + (any_friction ? 1 : 0);
        for (int k = 0; k < n_exec && s == RIGC_OK; k++) {
            B.execs.emplace_back(new Exec());
            Exec& e = *B.execs.back();
            e.B = &B;
            e.friction = any_friction && k == n_general;
            s = rigc_create(&e.ctx, device);
            if (s == RIGC_OK) s = rigc_set_scenes(e.ctx, num_scenes, scenes);
            if (s == RIGC_OK && B.detection != RIGC_DETECTION_HASH_GRID) s = rigc_set_detection_method(e.ctx, B.detection);
            if (s != RIGC_OK) b->err = e.ctx ? rigc_last_error(e.ctx) : "rigc_create failed";
        }
        if (s == RIGC_OK) {
            B.proxies.resize(num_scenes);
            B.nb.resize(num_scenes);
            b->tables.resize(num_scenes);
            for (int i = 0; i < num_scenes; i++) {
                B.nb[i] = scenes[i].num_bodies;
                Proxy& P = B.proxies[i];
                P.B = &B;
                P.scene = i;
                P.nb = scenes[i].num_bodies;
                b->tables[i] = rigs_backend{ &P, px_set_scene, px_toi, px_build, px_barrier, px_mind, px_fupdate,
                                             px_friction, px_error, nullptr, px_isect, nullptr, nullptr,
                                             px_detection };
            }
            // autotuner (RIGS_BATCH_AUTOTUNE, default on): the CCD's launch
            // width, walk budget and speculative levels timed on a
            // line-search-like sequence of earliest-TOI calls over all scenes
            // (each body's start pose, moved at its initial velocity over
            // four time steps, then over half and a quarter of that), then
            // set on every executor's context. The window factor keeps its
            // default: it acts through each scene's previous TOI, which this
            // sample at the start poses does not reproduce.
            const char* at = std::getenv("RIGS_BATCH_AUTOTUNE");
            if (!at || std::atoi(at) != 0) {
                size_t tot = 0;
                for (int i = 0; i < num_scenes; i++) tot += 6 * (size_t)scenes[i].num_bodies;
                const int n_samples = 3;
                std::vector<double> p0(n_samples * tot), p1(n_samples * tot);
                for (int rep = 0; rep < n_samples; rep++) {
                    size_t o = rep * tot;
                    for (int i = 0; i < num_scenes; i++) {
                        const double h = std::ldexp(4.0, -rep) * params[i].timestep;
                        for (int j = 0; j < 6 * scenes[i].num_bodies; j++) {
                            p0[o + j] = bodies[i].poses[j];
                            p1[o + j] = bodies[i].poses[j] + h * bodies[i].velocities[j];
                        }
                        o += 6 * (size_t)scenes[i].num_bodies;
                    }
                }
                rigc_tuning best{};
                s = rigc_autotune_earliest_toi(B.execs[0]->ctx, n_samples, p0.data(), p1.data(), 0.0,
                                               RIGC_TUNE_SPEC_LEVELS | RIGC_TUNE_WALK_BUDGET | RIGC_TUNE_WIDE_QUERIES,
                                               &best, &B.tune_ms);
                for (size_t k = 1; k < B.execs.size() && s == RIGC_OK; k++) s = rigc_set_tuning(B.execs[k]->ctx, &best);
                if (s != RIGC_OK) b->err = std::string("autotuner: ") + rigc_last_error(B.execs[0]->ctx);
                B.tuning = best;
                const char* lg = std::getenv("RIGS_BATCH_LOG");
                if (s == RIGC_OK && lg && std::atoi(lg) != 0) {
                    std::fprintf(stderr,
                                 "batch: autotuned CCD walk budget %d, speculative levels %d, wide launches from %lld "
                                 "queries (%.2f ms)\n",
                                 best.ccd_walk_budget, best.ccd_spec_levels, (long long)best.ccd_wide_queries,
                                 B.tune_ms);
                    std::fflush(stderr);
                }
            }
            if (s == RIGC_OK) B.start();
        }
    }
    for (int i = 0; i < num_scenes && s == RIGC_OK; i++) {
        s = rigs_create(&b->sims[i], mode == RIGS_BATCH_FUSED ? &b->tables[i] : nullptr, device, &scenes[i],
                        &bodies[i], &params[i]);
        if (s != RIGC_OK) {
            b->err = "rigs_create failed for scene " + std::to_string(i);
        } else {
            // fused: every scene's stepper already runs on its own thread,
            // and a scene's friction term joins the round of its barrier
            // term; sequential: each scene runs as the single-scene stepper
            if (mode == RIGS_BATCH_FUSED) {
                rigs::set_host_overlap(b->sims[i], false);
                rigs::set_fork_hooks(b->sims[i], fork_join, fork_part, &b->B);
            }
        }
    }
    if (s != RIGC_OK) {
        rigs_batch_destroy(b);
        return s;
    }
    *out = b;
    return RIGC_OK;
}

rigc_status rigs_batch_destroy(rigs_batch* b) {
    if (!b) return RIGC_OK;
    for (rigs_sim* s : b->sims)
        if (s) rigs_destroy(s);
    b->B.shutdown();
    delete b;
    return RIGC_OK;
}

const char* rigs_batch_last_error(const rigs_batch* b) { return b ? b->err.c_str() : "null batch"; }

rigc_status rigs_batch_step(rigs_batch* b, int32_t num_steps) {
    if (!b || num_steps < 0) return RIGC_ERR_INVALID_ARGUMENT;
    std::vector<rigc_status> res(b->n, RIGC_OK);
    if (b->mode == RIGS_BATCH_SEQUENTIAL) {
        for (int i = 0; i < b->n; i++) {
            res[i] = rigs_step(b->sims[i], num_steps);
            if (res[i] != RIGC_OK) b->errors[i] = rigs_last_error(b->sims[i]);
        }
    } else {
        Batch& B = b->B;
        {
            std::lock_guard<std::mutex> lk(B.m);
            B.active = b->n;
        }
        std::vector<std::thread> threads;
        threads.reserve(b->n);
        const char* lg = std::getenv("RIGS_BATCH_LOG");
        const bool log = lg && std::atoi(lg) != 0;
        const double t0 = now_ms();
        int finished = 0;
        std::mutex log_m;
        for (int i = 0; i < b->n; i++)
            threads.emplace_back([b, i, num_steps, &res, log, t0, &finished, &log_m] {
                res[i] = rigs_step(b->sims[i], num_steps);
                if (res[i] != RIGC_OK) b->errors[i] = rigs_last_error(b->sims[i]);
                leave(b->B);
                if (log) {
                    // progress of a long fused run: a line each time another
                    // sixteenth of the scenes has finished, and for the last
                    std::lock_guard<std::mutex> lk(log_m);
                    ++finished;
                    if (finished % std::max(1, b->n / 16) == 0 || finished == b->n) {
                        rigs_stats s{};
                        rigs_get_stats(b->sims[i], &s);
                        std::fprintf(stderr, "batch: %d/%d scenes done after %.2f s (last: scene %d, %lld Newton iterations)\n",
                                     finished, b->n, 1e-3 * (now_ms() - t0), i, (long long)s.newton_iterations);
                        std::fflush(stderr);
                    }
                }
            });
        for (auto& t : threads) t.join();
    }
    for (int i = 0; i < b->n; i++)
        if (res[i] != RIGC_OK) {
            b->err = "scene " + std::to_string(i) + ": " + b->errors[i];
            return res[i];
        }
    return RIGC_OK;
}

rigc_status rigs_batch_get_state(const rigs_batch* b, int32_t scene, double* poses, double* velocities) {
    if (!b || scene < 0 || scene >= b->n) return RIGC_ERR_INVALID_ARGUMENT;
    return rigs_get_state(b->sims[scene], poses, velocities);
}

rigc_status rigs_batch_get_stats(const rigs_batch* b, int32_t scene, rigs_stats* stats) {
    if (!b || scene < 0 || scene >= b->n) return RIGC_ERR_INVALID_ARGUMENT;
    return rigs_get_stats(b->sims[scene], stats);
}

rigc_status rigs_batch_get_batch_stats(const rigs_batch* b, rigs_batch_stats* stats) {
    if (!b || !stats) return RIGC_ERR_INVALID_ARGUMENT;
    rigs_batch_stats t{};
    for (const auto& e : b->B.execs) {
        t.rounds += e->st.rounds;
        t.fused_calls += e->st.fused_calls;
        t.requests += e->st.requests;
        t.device_ms += e->st.device_ms;
        t.ccd_calls += e->st.ccd_calls;
        t.ccd_candidates += e->st.ccd_candidates;
        t.ccd_kernel_ms += e->st.ccd_kernel_ms;
        t.ccd_launches += e->st.ccd_launches;
        t.ccd_box_evaluations += e->st.ccd_box_evaluations;
    }
    t.executors = (int32_t)b->B.execs.size();
    t.tuning = b->B.tuning;
    t.autotune_ms = b->B.tune_ms;
    *stats = t;
    return RIGC_OK;
}

rigc_status rigs_batch_set_gather(rigs_batch* b, double gather_us) {
    if (!b || std::isnan(gather_us)) return RIGC_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lk(b->B.m);
    b->B.gather_ms = gather_us < 0 ? -1.0 : gather_us * 1e-3;
    b->B.cv_exec.notify_all();
    return RIGC_OK;
}

}  // extern "C"
