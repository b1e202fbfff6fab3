// ORACLE — test infrastructure only (see oracle_interval.hpp).
//
// Restates the rigid hash-grid broad phase:
//   close_bodies / close_bodies_bvh     src/physics/rigid_body_assembler.cpp:268-356
//   RigidBody::compute_bounding_box     src/physics/rigid_body.cpp:246-272
//   body_pairs_to_body_ids              src/ccd/rigid/rigid_body_hash_grid.cpp:17-34
//   compute_scene_conservative_bbox     :64-88, RigidBodyHashGrid::resize :36-62, :90-119
//   compute_vertices_intervals          :164-246 (forced split + "fits" recursion)
//   intervals_to_AABB / addBodies       :248-344 (continuous), :122-162 (discrete)
//   HashGrid::get{EdgeEdge,FaceVertex}Pairs (ipc-toolkit @d41d6ca, not in the
//     reference tree): items sharing a cell, AABB overlap test in the
//     center / half-extent form, group filter, deduplicated. The resulting
//     set does not depend on the cell size: it is every pair of included
//     elements whose AABBs overlap and whose bodies' groups differ.
// The SimpleBVH box query of close_bodies_bvh is restated as the inclusive
// min/max overlap of all body-box pairs.
#include <algorithm>
#include <cmath>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle_scene.hpp"

namespace oracle {

Scene& scene() {
    static Scene s;
    return s;
}

// Vertex AABBs of the most recent broad-phase call, exported for tests.
std::vector<double>& last_vertex_aabbs() {
    static std::vector<double> v;
    return v;
}

void scene_derive(Scene& s) {
    s.v_body.assign(s.nv, 0);
    s.e_body.assign(s.ne, 0);
    s.f_body.assign(s.nf, 0);
    s.r_max.assign(s.nb, 0.0);
    s.edge_len.assign(s.ne, 0.0);
    double avg_sum = 0;
    for (int b = 0; b < s.nb; b++) {
        // r_max = vertices.rowwise().norm().maxCoeff() (rigid_body.cpp:129)
        for (int i = s.bv[b]; i < s.bv[b + 1]; i++) {
            s.v_body[i] = b;
            const double* v = &s.V[3 * i];
            s.r_max[b] = std::max(s.r_max[b], std::sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]));
        }
        // average_edge_length (rigid_body.cpp:131-140), then the assembler's
        // edge-count weighted mean (rigid_body_assembler.cpp:103-107)
        double avg = 0;
        for (int i = s.be[b]; i < s.be[b + 1]; i++) {
            s.e_body[i] = b;
            const double* a = &s.V[3 * s.E[2 * i]];
            const double* c = &s.V[3 * s.E[2 * i + 1]];
            const double d0 = c[0] - a[0], d1 = c[1] - a[1], d2 = c[2] - a[2];
            s.edge_len[i] = std::sqrt((d0 * d0 + d1 * d1) + d2 * d2);
            avg += s.edge_len[i];
        }
        const int ne_b = s.be[b + 1] - s.be[b];
        if (ne_b > 0) avg /= ne_b;
        avg_sum += ne_b * avg;
        for (int i = s.bf[b]; i < s.bf[b + 1]; i++) s.f_body[i] = b;
    }
    s.average_edge_length = s.ne ? avg_sum / s.ne : 0.0;
}

void world_vertices(const Scene& s, int b, const double pose[6], std::vector<double>& out) {
    const Mat3d R = rotation_matrix(pose + 3);
    const int n = s.bv[b + 1] - s.bv[b];
    out.resize(3 * (size_t)n);
    for (int i = 0; i < n; i++) world_vertex(R, pose, &s.V[3 * (s.bv[b] + i)], &out[3 * i]);
}

namespace {

struct Box { double mn[3], mx[3]; };

// RigidBody::compute_bounding_box (rigid_body.cpp:246-272)
Box body_box(const Scene& s, int b, const double* p0, const double* p1) {
    Box bx;
    const bool same_rot = p0[3] == p1[3] && p0[4] == p1[4] && p0[5] == p1[5];
    if (s.is_static[b] || same_rot) {
        std::vector<double> w;
        for (int k = 0; k < 3; k++) { bx.mn[k] = INFINITY; bx.mx[k] = -INFINITY; }
        for (const double* p : { p0, p1 }) {
            world_vertices(s, b, p, w);
            for (size_t i = 0; i < w.size() / 3; i++)
                for (int k = 0; k < 3; k++) {
                    bx.mn[k] = std::min(bx.mn[k], w[3 * i + k]);
                    bx.mx[k] = std::max(bx.mx[k], w[3 * i + k]);
                }
        }
    } else {
        for (int k = 0; k < 3; k++) {
            bx.mn[k] = std::min(p0[k], p1[k]) - s.r_max[b];
            bx.mx[k] = std::max(p0[k], p1[k]) + s.r_max[b];
        }
    }
    return bx;
}

// close_bodies -> body_pairs_to_body_ids: the bodies in any overlapping pair.
std::vector<char> included_bodies(const Scene& s, const double* p0, const double* p1, double r,
                                  std::vector<Box>* raw_boxes) {
    std::vector<Box> raw(s.nb), inf(s.nb);
#pragma omp parallel for schedule(dynamic, 16)
    for (int b = 0; b < s.nb; b++) {
        raw[b] = body_box(s, b, p0 + 6 * b, p1 + 6 * b);
        for (int k = 0; k < 3; k++) {
            inf[b].mn[k] = raw[b].mn[k] - r;
            inf[b].mx[k] = raw[b].mx[k] + r;
        }
    }
    std::vector<char> inc(s.nb, 0);
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < s.nb; i++) {
        for (int j = 0; j < s.nb && !inc[i]; j++) {
            if (j == i || s.group[i] == s.group[j]) continue;
            bool ov = true;
            for (int k = 0; k < 3; k++) ov = ov && inf[i].mn[k] <= inf[j].mx[k] && inf[j].mn[k] <= inf[i].mx[k];
            if (ov) inc[i] = 1;
        }
    }
    if (raw_boxes) *raw_boxes = raw;
    return inc;
}

// ipc-toolkit AABB: center / half-extent overlap.
struct AABB {
    double mn[3], mx[3], c[3], h[3];
    void finish() {
        for (int k = 0; k < 3; k++) {
            h[k] = (mx[k] - mn[k]) / 2;
            c[k] = mn[k] + h[k];
        }
    }
};
inline bool overlapping(const AABB& a, const AABB& b) {
    for (int k = 0; k < 3; k++)
        if (!(std::fabs(a.c[k] - b.c[k]) <= a.h[k] + b.h[k])) return false;
    return true;
}

// Pairs of included elements with overlapping AABBs (ipc HashGrid semantics),
// found with a uniform grid; cell_size only affects speed.
void emit_pairs(const Scene& s, const std::vector<AABB>& va, const std::vector<char>& vin,
                const std::vector<AABB>& ea, const std::vector<char>& ein, const std::vector<AABB>& fa,
                const std::vector<char>& fin, double cell_size, Candidates& out) {
    double dmin[3] = { INFINITY, INFINITY, INFINITY };
    for (int i = 0; i < s.nv; i++)
        if (vin[i])
            for (int k = 0; k < 3; k++) dmin[k] = std::min(dmin[k], va[i].mn[k]);
    if (!(cell_size > 0) || !std::isfinite(dmin[0])) return;
    struct Item { unsigned long long key; int type, id; };
    auto cell = [&](double x, int k) { return (long long)std::floor((x - dmin[k]) / cell_size); };
    std::vector<Item> items;
    auto add = [&](const AABB& a, int type, int id) {
        long long lo[3], hi[3];
        for (int k = 0; k < 3; k++) { lo[k] = cell(a.mn[k], k); hi[k] = cell(a.mx[k], k); }
        for (long long x = lo[0]; x <= hi[0]; x++)
            for (long long y = lo[1]; y <= hi[1]; y++)
                for (long long z = lo[2]; z <= hi[2]; z++) {
                    const unsigned long long key = ((unsigned long long)(x & 0x1FFFFF) << 42)
                        | ((unsigned long long)(y & 0x1FFFFF) << 21) | (unsigned long long)(z & 0x1FFFFF);
                    items.push_back({ key, type, id });
                }
    };
    for (int i = 0; i < s.nv; i++) if (vin[i]) add(va[i], 0, i);
    for (int i = 0; i < s.ne; i++) if (ein[i]) add(ea[i], 1, i);
    for (int i = 0; i < s.nf; i++) if (fin[i]) add(fa[i], 2, i);
    std::sort(items.begin(), items.end(), [](const Item& a, const Item& b) {
        return a.key != b.key ? a.key < b.key : (a.type != b.type ? a.type < b.type : a.id < b.id);
    });
    std::vector<size_t> starts;
    for (size_t i = 0; i < items.size(); i++)
        if (i == 0 || items[i].key != items[i - 1].key) starts.push_back(i);
    starts.push_back(items.size());
    Candidates& C = out;
    std::vector<std::vector<std::array<int, 2>>> ee_t, fv_t;
    int nt = 1;
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    ee_t.resize(nt);
    fv_t.resize(nt);
#pragma omp parallel for schedule(dynamic, 64)
    for (long long c = 0; c < (long long)starts.size() - 1; c++) {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        const size_t a0 = starts[c], a1 = starts[c + 1];
        for (size_t i = a0; i < a1; i++) {
            const Item& I = items[i];
            for (size_t j = i + 1; j < a1; j++) {
                const Item& J = items[j];
                if (I.type == 1 && J.type == 1) {
                    const int e0 = std::min(I.id, J.id), e1 = std::max(I.id, J.id);
                    if (s.group[s.e_body[e0]] == s.group[s.e_body[e1]]) continue;
                    if (!overlapping(ea[e0], ea[e1])) continue;
                    ee_t[tid].push_back({ e0, e1 });
                } else if ((I.type == 0 && J.type == 2) || (I.type == 2 && J.type == 0)) {
                    const int f = I.type == 2 ? I.id : J.id, v = I.type == 0 ? I.id : J.id;
                    if (s.group[s.f_body[f]] == s.group[s.v_body[v]]) continue;
                    if (!overlapping(fa[f], va[v])) continue;
                    fv_t[tid].push_back({ f, v });
                }
            }
        }
    }
    for (int t = 0; t < nt; t++) {
        C.ee.insert(C.ee.end(), ee_t[t].begin(), ee_t[t].end());
        C.fv.insert(C.fv.end(), fv_t[t].begin(), fv_t[t].end());
    }
    std::sort(C.ee.begin(), C.ee.end());
    C.ee.erase(std::unique(C.ee.begin(), C.ee.end()), C.ee.end());
    std::sort(C.fv.begin(), C.fv.end());
    C.fv.erase(std::unique(C.fv.begin(), C.fv.end()), C.fv.end());
}

void element_aabbs(const Scene& s, const std::vector<AABB>& va, const std::vector<char>& vin,
                   std::vector<AABB>& ea, std::vector<char>& ein, std::vector<AABB>& fa,
                   std::vector<char>& fin) {
    ea.resize(s.ne); ein.assign(s.ne, 0);
    fa.resize(s.nf); fin.assign(s.nf, 0);
    for (int i = 0; i < s.ne; i++) {
        const int a = s.E[2 * i], b = s.E[2 * i + 1];
        if (!(vin[a] && vin[b])) continue;
        ein[i] = 1;
        for (int k = 0; k < 3; k++) {
            ea[i].mn[k] = std::min(va[a].mn[k], va[b].mn[k]);
            ea[i].mx[k] = std::max(va[a].mx[k], va[b].mx[k]);
        }
        ea[i].finish();
    }
    for (int i = 0; i < s.nf; i++) {
        const int a = s.F[3 * i], b = s.F[3 * i + 1], c = s.F[3 * i + 2];
        if (!(vin[a] && vin[b] && vin[c])) continue;
        fin[i] = 1;
        for (int k = 0; k < 3; k++) {
            fa[i].mn[k] = std::min(std::min(va[a].mn[k], va[b].mn[k]), va[c].mn[k]);
            fa[i].mx[k] = std::max(std::max(va[a].mx[k], va[b].mx[k]), va[c].mx[k]);
        }
        fa[i].finish();
    }
}

void export_vertex_aabbs(const std::vector<AABB>& va, const std::vector<char>& vin) {
    std::vector<double>& o = last_vertex_aabbs();
    o.assign(7 * va.size(), 0.0);
    for (size_t i = 0; i < va.size(); i++) {
        o[7 * i] = vin[i] ? 1.0 : 0.0;
        for (int k = 0; k < 3; k++) {
            o[7 * i + 1 + k] = va[i].mn[k];
            o[7 * i + 4 + k] = va[i].mx[k];
        }
    }
}

// compute_vertices_intervals recursion (rigid_body_hash_grid.cpp:188-246)
// for one body; out receives the hull of the leaf enclosures.
void vertex_intervals(const Scene& s, int b, const double* p0, const double* p1, const double dmn[3],
                      const double dmx[3], const Interval& t, int force, int depth,
                      std::vector<Interval>& out) {
    const int n = s.bv[b + 1] - s.bv[b];
    if (force <= 0 || depth > 40) {
        PoseI pose = interpolate(p0, p1, t);
        Mat3I R = rotation_matrix(pose.r);
        out.resize(3 * (size_t)n);
        bool fits = true;
        for (int i = 0; i < n; i++) {
            Vec3I w = world_vertex(R, pose.p, &s.V[3 * (s.bv[b] + i)]);
            for (int k = 0; k < 3; k++) {
                out[3 * i + k] = w[k];
                // the reference tests w.lo - r < min || w.hi + r > max
                // (rigid_body_hash_grid.cpp:203-206). The domain is not
                // grown by r (:103-104 leave it as it was), so for r > 0 a
                // vertex on the domain's edge never fits and the reference
                // splits t without end. The test leaves r out: identical
                // for r = 0, and it ends for r > 0.
                if (w[k].lo < dmn[k] || w[k].hi > dmx[k]) fits = false;
            }
        }
        if (fits || depth > 40) return;
    }
    force--;
    const double m = median(t);
    std::vector<Interval> a, c;
    vertex_intervals(s, b, p0, p1, dmn, dmx, Interval::raw(t.lo, m), force, depth + 1, a);
    vertex_intervals(s, b, p0, p1, dmn, dmx, Interval::raw(m, t.hi), force, depth + 1, c);
    out.resize(3 * (size_t)n);
    for (size_t i = 0; i < out.size(); i++) out[i] = hull(a[i], c[i]);
}

}  // namespace

std::vector<double>& last_vertex_aabbs();

void broadphase_ccd(const Scene& s, const double* p0, const double* p1, double r, Candidates& out) {
    out.ee.clear();
    out.fv.clear();
    if (s.nb <= 1) return;
    std::vector<Box> raw;
    std::vector<char> inc = included_bodies(s, p0, p1, r, &raw);
    bool any = false;
    for (int b = 0; b < s.nb; b++) any = any || inc[b];
    if (!any) return;

    // compute_scene_conservative_bbox (+ the resize quirk min -= r; min += r)
    double dmn[3] = { INFINITY, INFINITY, INFINITY }, dmx[3] = { -INFINITY, -INFINITY, -INFINITY };
    for (int b = 0; b < s.nb; b++)
        if (inc[b])
            for (int k = 0; k < 3; k++) {
                dmn[k] = std::min(dmn[k], raw[b].mn[k]);
                dmx[k] = std::max(dmx[k], raw[b].mx[k]);
            }
    const double dx = dmx[0] - dmn[0], dy = dmx[1] - dmn[1], dz = dmx[2] - dmn[2];
    const double diag = std::sqrt((dx * dx + dy * dy) + dz * dz);
    for (int k = 0; k < 3; k++) {
        dmn[k] -= 1e-8 * diag;
        dmx[k] += 1e-8 * diag;
        dmn[k] -= r;
        dmn[k] += r;
    }

    // vertex AABBs from the interval trajectories
    std::vector<AABB> va(s.nv);
    std::vector<char> vin(s.nv, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int b = 0; b < s.nb; b++) {
        if (!inc[b]) continue;
        std::vector<Interval> iv;
        vertex_intervals(s, b, p0 + 6 * b, p1 + 6 * b, dmn, dmx, Interval(0, 1), 1, 0, iv);
        for (int i = 0; i < s.bv[b + 1] - s.bv[b]; i++) {
            AABB& a = va[s.bv[b] + i];
            for (int k = 0; k < 3; k++) {
                a.mn[k] = iv[3 * i + k].lo - r;
                a.mx[k] = iv[3 * i + k].hi + r;
            }
            a.finish();
            vin[s.bv[b] + i] = 1;
        }
    }
    std::vector<AABB> ea, fa;
    std::vector<char> ein, fin;
    element_aabbs(s, va, vin, ea, ein, fa, fin);
    export_vertex_aabbs(va, vin);

    // cell size: max(mean linearised displacement, mean edge length) (:108-116)
    double disp = 0;
    std::vector<double> w0, w1;
    for (int b = 0; b < s.nb; b++) {
        world_vertices(s, b, p0 + 6 * b, w0);
        world_vertices(s, b, p1 + 6 * b, w1);
        for (size_t i = 0; i < w0.size() / 3; i++) {
            const double a = w1[3 * i] - w0[3 * i], c = w1[3 * i + 1] - w0[3 * i + 1], d = w1[3 * i + 2] - w0[3 * i + 2];
            disp += std::sqrt((a * a + c * c) + d * d);
        }
    }
    disp = s.nv ? disp / s.nv : 0;
    emit_pairs(s, va, vin, ea, ein, fa, fin, std::max(disp, s.average_edge_length), out);
}

void broadphase_static(const Scene& s, const double* poses, double r, Candidates& out) {
    out.ee.clear();
    out.fv.clear();
    if (s.nb <= 1) return;
    std::vector<char> inc = included_bodies(s, poses, poses, r, nullptr);
    std::vector<AABB> va(s.nv);
    std::vector<char> vin(s.nv, 0);
#pragma omp parallel for schedule(dynamic, 4)
    for (int b = 0; b < s.nb; b++) {
        if (!inc[b]) continue;
        std::vector<double> w;
        world_vertices(s, b, poses + 6 * b, w);
        for (int i = 0; i < s.bv[b + 1] - s.bv[b]; i++) {
            AABB& a = va[s.bv[b] + i];
            for (int k = 0; k < 3; k++) {
                a.mn[k] = w[3 * i + k] - r;
                a.mx[k] = w[3 * i + k] + r;
            }
            a.finish();
            vin[s.bv[b] + i] = 1;
        }
    }
    std::vector<AABB> ea, fa;
    std::vector<char> ein, fin;
    element_aabbs(s, va, vin, ea, ein, fa, fin);
    export_vertex_aabbs(va, vin);
    emit_pairs(s, va, vin, ea, ein, fa, fin, s.average_edge_length, out);
}

}  // namespace oracle
