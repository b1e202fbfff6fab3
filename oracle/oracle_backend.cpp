// ORACLE — test infrastructure only (see oracle_interval.hpp).
//
// The CPU restatement behind the host stepper's contact-backend table
// (rigid-ipc_amd/host/rigs.h), so that the stepper tests and the CPU
// baseline run the same Newton loop over the oracle instead of the device
// library. Per-constraint 12x12 blocks are summed into the same
// block-sparse Hessian layout (diag / body pairs) as the device library.
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../rigid-ipc_amd/host/rigs.h"
#include "oracle_barrier.hpp"
#include "oracle_scene.hpp"

using namespace oracle;

extern "C" double oracle_earliest_toi(const double* p0, const double* p1, double msd, long* n_candidates);
extern "C" double oracle_earliest_toi_trajectory(const double* p0, const double* p1, double msd, int trajectory,
                                                 long* n_candidates);

extern "C" void oracle_set_detection_method(int m);

namespace {

struct OracleBackend {
    long last_candidates = 0;
    std::vector<Constraint> cons;
    std::vector<FrictionConstraint> fric;
    bool have_cons = false, have_fric = false;
    std::string err;
};

OracleBackend& state() {
    static OracleBackend b;
    return b;
}

rigc_status ob_set_scene(void*, const rigc_scene* sc) {
    oracle_set_detection_method(0);  // a new scene starts with the hash grid
    Scene& s = scene();
    s.nb = sc->num_bodies; s.nv = sc->num_vertices; s.ne = sc->num_edges; s.nf = sc->num_faces;
    s.V.assign(sc->vertices, sc->vertices + 3 * (size_t)s.nv);
    s.E.assign(sc->edges, sc->edges + 2 * (size_t)s.ne);
    s.F.assign(sc->faces, sc->faces + 3 * (size_t)s.nf);
    s.FE.assign(sc->face_edges, sc->face_edges + 3 * (size_t)s.nf);
    s.bv.assign(sc->body_vertex_start, sc->body_vertex_start + s.nb + 1);
    s.be.assign(sc->body_edge_start, sc->body_edge_start + s.nb + 1);
    s.bf.assign(sc->body_face_start, sc->body_face_start + s.nb + 1);
    s.group.assign(sc->group_ids, sc->group_ids + s.nb);
    s.is_static.assign(sc->is_static, sc->is_static + s.nb);
    scene_derive(s);
    state() = OracleBackend{};
    return RIGC_OK;
}

rigc_status ob_toi(void*, const double* p0, const double* p1, double msd, double* toi) {
    long n = 0;
    *toi = oracle_earliest_toi(p0, p1, msd, &n);
    state().last_candidates = n;
    return RIGC_OK;
}
rigc_status ob_toi_pl(void*, const double* p0, const double* p1, double msd, double* toi) {
    long n = 0;
    *toi = oracle_earliest_toi_trajectory(p0, p1, msd, 1, &n);
    state().last_candidates = n;
    return RIGC_OK;
}

rigc_status ob_stats(void*, rigc_stats* st) {
    *st = rigc_stats{};
    st->num_candidates = state().last_candidates;
    return RIGC_OK;
}

rigc_status ob_build(void*, const double* poses, double dhat, double dmin, int64_t* counts) {
    OracleBackend& B = state();
    construct_constraint_set(scene(), poses, dhat, dmin, B.cons);
    B.have_cons = true;
    if (counts) {
        for (int k = 0; k < 4; k++) counts[k] = 0;
        for (const auto& c : B.cons) counts[c.type]++;
    }
    return RIGC_OK;
}

// Sum per-constraint 12x12 blocks into diag / pairs / off.
rigc_status blocks_out(const std::vector<std::array<double, 144>>& lh, const std::vector<const Constraint*>& C,
                       rigc_block_hessian* h) {
    const int nb = scene().nb;
    std::memset(h->diag, 0, sizeof(double) * 36 * (size_t)nb);
    std::map<std::pair<int, int>, std::array<double, 36>> off;
    for (size_t c = 0; c < lh.size(); c++) {
        const int b[2] = { C[c]->body[0], C[c]->body[1] };
        for (int si = 0; si < 2; si++)
            for (int sj = 0; sj < 2; sj++) {
                if (b[si] == b[sj]) {
                    for (int p = 0; p < 6; p++)
                        for (int q = 0; q < 6; q++) h->diag[36 * b[si] + 6 * p + q] += lh[c][12 * (6 * si + p) + 6 * sj + q];
                } else if (b[si] < b[sj]) {
                    auto& blk = off[{ b[si], b[sj] }];
                    for (int p = 0; p < 6; p++)
                        for (int q = 0; q < 6; q++) blk[6 * p + q] += lh[c][12 * (6 * si + p) + 6 * sj + q];
                }
            }
    }
    h->count = (int64_t)off.size();
    if (h->count > h->capacity) return RIGC_ERR_CAPACITY;
    int64_t k = 0;
    for (const auto& kv : off) {
        h->pairs[2 * k] = kv.first.first;
        h->pairs[2 * k + 1] = kv.first.second;
        std::memcpy(h->off + 36 * k, kv.second.data(), sizeof(double) * 36);
        k++;
    }
    return RIGC_OK;
}

rigc_status ob_barrier(void*, const double* poses, double dhat, double dmin, int32_t flags, double* e, double* g,
                       rigc_block_hessian* h) {
    OracleBackend& B = state();
    if (flags & RIGC_BUILD_CONSTRAINTS) ob_build(nullptr, poses, dhat, dmin, nullptr);
    if (!B.have_cons) {
        B.err = "no constraint set";
        return RIGC_ERR_NO_SCENE;
    }
    const bool want_hess = (flags & RIGC_HESSIAN) != 0;
    const bool want_grad = want_hess || (flags & RIGC_GRADIENT) != 0;
    const int nb = scene().nb;
    std::vector<double> grad(6 * (size_t)nb, 0.0);
    std::vector<std::array<double, 144>> lh;
    const bool proj = psd_projection_enabled();
    psd_projection_enabled() = (flags & RIGC_UNPROJECTED) == 0;
    *e = barrier_term(scene(), poses, B.cons, dhat, dmin, want_grad ? grad.data() : nullptr, want_hess ? &lh : nullptr);
    psd_projection_enabled() = proj;
    if (want_grad) std::memcpy(g, grad.data(), sizeof(double) * grad.size());
    if (want_hess) {
        std::vector<const Constraint*> C;
        for (const auto& c : B.cons) C.push_back(&c);
        rigc_status s = blocks_out(lh, C, h);
        if (s != RIGC_OK) B.err = "off-diagonal Hessian capacity too small";
        return s;
    }
    return RIGC_OK;
}

rigc_status ob_mind(void*, const double* poses, double dhat, double dmin, int32_t flags, double* out) {
    OracleBackend& B = state();
    if (flags & RIGC_BUILD_CONSTRAINTS) ob_build(nullptr, poses, dhat, dmin, nullptr);
    *out = std::sqrt(minimum_distance(scene(), poses, B.cons));
    return RIGC_OK;
}

rigc_status ob_fupdate(void*, const double* poses, double dhat, double dmin, double kappa, double mu, int32_t flags,
                       int64_t* n) {
    OracleBackend& B = state();
    if (flags & RIGC_BUILD_CONSTRAINTS) ob_build(nullptr, poses, dhat, dmin, nullptr);
    construct_friction_constraint_set(scene(), poses, B.cons, dhat, dmin, kappa, mu, B.fric);
    B.have_fric = true;
    if (n) *n = (int64_t)B.fric.size();
    return RIGC_OK;
}

rigc_status ob_friction(void*, const double* poses, const double* poses_t0, double eh, int32_t flags, double* e,
                        double* g, rigc_block_hessian* h) {
    OracleBackend& B = state();
    if (!B.have_fric) {
        B.err = "no friction set";
        return RIGC_ERR_NO_SCENE;
    }
    const bool want_hess = (flags & RIGC_HESSIAN) != 0;
    const bool want_grad = want_hess || (flags & RIGC_GRADIENT) != 0;
    const int nb = scene().nb;
    std::vector<double> grad(6 * (size_t)nb, 0.0);
    std::vector<std::array<double, 144>> lh;
    const bool proj = psd_projection_enabled();
    psd_projection_enabled() = (flags & RIGC_UNPROJECTED) == 0;
    *e = friction_term(scene(), poses, poses_t0, B.fric, eh, want_grad ? grad.data() : nullptr,
                       want_hess ? &lh : nullptr);
    psd_projection_enabled() = proj;
    if (want_grad) std::memcpy(g, grad.data(), sizeof(double) * grad.size());
    if (want_hess) {
        std::vector<const Constraint*> C;
        for (const auto& f : B.fric) C.push_back(&f.c);
        rigc_status s = blocks_out(lh, C, h);
        if (s != RIGC_OK) B.err = "off-diagonal Hessian capacity too small";
        return s;
    }
    return RIGC_OK;
}

rigc_status ob_fadjoint(void*, const double* poses, const double* poses_t0, double eh, double dhat, double dmin,
                        double kappa, const double* lambda, double* out) {
    OracleBackend& B = state();
    if (!B.have_fric) {
        B.err = "no friction set";
        return RIGC_ERR_NO_SCENE;
    }
    friction_adjoint(scene(), poses, poses_t0, B.fric, eh, dhat, dmin, kappa, lambda, out);
    return RIGC_OK;
}

const char* ob_error(void*) { return state().err.c_str(); }

rigc_status ob_isect(void*, const double* poses, int64_t* n) {
    *n = (int64_t)detect_intersections(scene(), poses).size();
    return RIGC_OK;
}

}  // namespace

extern "C" {

// Fills a rigs_backend table with the oracle's entry points. The detection
// method is the oracle's (process-wide) setting.
void oracle_backend(rigs_backend* out) {
    *out = rigs_backend{ nullptr, ob_set_scene, ob_toi, ob_build, ob_barrier, ob_mind, ob_fupdate, ob_friction,
                         ob_error, ob_stats, ob_isect, ob_toi_pl, ob_fadjoint,
                         [](void*, int32_t m) -> rigc_status {
                             oracle_set_detection_method(m);
                             return RIGC_OK;
                         } };
}

}  // extern "C"
