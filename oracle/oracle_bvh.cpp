// ORACLE — test infrastructure only (see oracle_interval.hpp).
//
// Restates the rigid BVH broad phase (DetectionMethod::BVH, the reference's
// default detection_method, src/SimState.cpp:178):
//   detect_collision_candidates_rigid_bvh, discrete   src/ccd/rigid/broad_phase.cpp:92-121
//   detect_collision_candidates_rigid_bvh, continuous src/ccd/rigid/broad_phase.cpp:205-230
//   close_bodies_bvh                                  src/physics/rigid_body_assembler.cpp:307-356
//   sort_body_pair, vertex_aabb(s),
//   detect_body_pair_collision_candidates_bvh         src/ccd/rigid/rigid_body_bvh.hpp:14-94
//   detect_body_pair_collision_candidates_from_aabbs  src/ccd/rigid/rigid_body_bvh.cpp:5-190 (query (f, *))
//   RigidBody::init_bvh (primitive boxes)             src/physics/rigid_body.cpp:184-230
//   MeshSelector ownership (lowest face index)        src/utils/mesh_selector.cpp:47-120
//
// The BVH (SimpleBVH, external) returns every primitive of body B whose box
// overlaps the query box (inclusive min/max test); that is restated as the
// test over all of B's faces, so the candidate set is the reference's
// whatever the tree. The interval transform of body A's vertices into B's
// frame, ((V_A R_A^T) + (p_A - p_B)^T) R_B, sums its three-term products
// left to right, as the oracle restates Eigen's dynamic-size products
// (oracle_ccd.hpp). Codimensional vertices and edges (in no edge / no face)
// are not restated: the reference queries them separately and no scene
// here has them (the caller checks).
#include <algorithm>
#include <cmath>

#include "oracle_scene.hpp"

namespace oracle {

namespace {

// ipc-toolkit AABB: min / max with center and half extent; overlap in the
// center form (as the hash-grid restatement).
struct Box {
    double mn[3], mx[3], c[3], h[3];
    void finish() {
        for (int k = 0; k < 3; k++) {
            h[k] = (mx[k] - mn[k]) / 2;
            c[k] = mn[k] + h[k];
        }
    }
};
inline bool overlapping(const Box& a, const Box& b) {
    for (int k = 0; k < 3; k++)
        if (!(std::fabs(a.c[k] - b.c[k]) <= a.h[k] + b.h[k])) return false;
    return true;
}
// AABB(a, b): the union's min / max, then center and half extent.
Box merge(const Box& a, const Box& b) {
    Box r;
    for (int k = 0; k < 3; k++) {
        r.mn[k] = std::min(a.mn[k], b.mn[k]);
        r.mx[k] = std::max(a.mx[k], b.mx[k]);
    }
    r.finish();
    return r;
}
// AABB(a, b, c) = AABB(a, AABB(b, c)).
Box merge(const Box& a, const Box& b, const Box& c) { return merge(a, merge(b, c)); }

// RigidBody::compute_bounding_box (rigid_body.cpp:246-272) as the
// hash-grid restatement computes it.
void body_box(const Scene& s, int b, const double* p0, const double* p1, double mn[3], double mx[3]) {
    const bool same_rot = p0[3] == p1[3] && p0[4] == p1[4] && p0[5] == p1[5];
    if (s.is_static[b] || same_rot) {
        std::vector<double> w;
        for (int k = 0; k < 3; k++) { mn[k] = INFINITY; mx[k] = -INFINITY; }
        for (const double* p : { p0, p1 }) {
            world_vertices(s, b, p, w);
            for (size_t i = 0; i < w.size() / 3; i++)
                for (int k = 0; k < 3; k++) {
                    mn[k] = std::min(mn[k], w[3 * i + k]);
                    mx[k] = std::max(mx[k], w[3 * i + k]);
                }
        }
    } else {
        for (int k = 0; k < 3; k++) {
            mn[k] = std::min(p0[k], p1[k]) - s.r_max[b];
            mx[k] = std::max(p0[k], p1[k]) + s.r_max[b];
        }
    }
}

}  // namespace

std::vector<std::array<int, 2>> close_bodies_bvh(const Scene& s, const double* p0, const double* p1, double r) {
    std::vector<std::array<double, 6>> bx(s.nb);
    for (int b = 0; b < s.nb; b++) {
        double mn[3], mx[3];
        body_box(s, b, p0 + 6 * b, p1 + 6 * b, mn, mx);
        for (int k = 0; k < 3; k++) {
            bx[b][k] = mn[k] - r;
            bx[b][3 + k] = mx[k] + r;
        }
    }
    std::vector<std::array<int, 2>> pairs;
    for (int i = 0; i < s.nb; i++)
        for (int j = i + 1; j < s.nb; j++) {
            if (s.group[i] == s.group[j]) continue;
            bool ov = true;
            for (int k = 0; k < 3; k++) ov = ov && bx[i][k] <= bx[j][3 + k] && bx[j][k] <= bx[i][3 + k];
            if (ov) pairs.push_back({ i, j });
        }
    return pairs;
}

bool bvh_supported(const Scene& s) {
    // every vertex in an edge and every edge in a face
    std::vector<char> ve(s.nv, 0), ef(s.ne, 0);
    for (int e = 0; e < s.ne; e++) ve[s.E[2 * e]] = ve[s.E[2 * e + 1]] = 1;
    for (int f = 0; f < s.nf; f++)
        for (int j = 0; j < 3; j++) ef[s.FE[3 * f + j]] = 1;
    for (int v = 0; v < s.nv; v++)
        if (!ve[v]) return false;
    for (int e = 0; e < s.ne; e++)
        if (!ef[e]) return false;
    return true;
}

void broadphase_bvh(const Scene& s, const double* p0, const double* p1, double r, bool continuous,
                    Candidates& out) {
    out.ee.clear();
    out.fv.clear();
    if (s.nb <= 1) return;
    const double* q1 = continuous ? p1 : p0;
    const std::vector<std::array<int, 2>> pairs = close_bodies_bvh(s, p0, q1, r);
    // MeshSelector: lowest face index holding a vertex / an edge
    std::vector<int> v2f(s.nv, INT32_MAX), e2f(s.ne, INT32_MAX);
    for (int f = 0; f < s.nf; f++)
        for (int j = 0; j < 3; j++) {
            v2f[s.F[3 * f + j]] = std::min(v2f[s.F[3 * f + j]], f);
            e2f[s.FE[3 * f + j]] = std::min(e2f[s.FE[3 * f + j]], f);
        }
    std::vector<Candidates> part(pairs.size());
#pragma omp parallel for schedule(dynamic, 4)
    for (long pi = 0; pi < (long)pairs.size(); pi++) {
        int A = pairs[pi][0], B = pairs[pi][1];
        // sort_body_pair: the body with fewer BVH primitives first
        if (s.bf[A + 1] - s.bf[A] > s.bf[B + 1] - s.bf[B]) std::swap(A, B);
        PoseI pa, pb;
        if (continuous) {
            pa = interpolate(p0 + 6 * A, p1 + 6 * A, Interval::raw(0.0, 1.0));
            pb = interpolate(p0 + 6 * B, p1 + 6 * B, Interval::raw(0.0, 1.0));
        } else {
            for (int k = 0; k < 3; k++) {
                pa.p[k] = Interval(p0[6 * A + k]);
                pa.r[k] = Interval(p0[6 * A + 3 + k]);
                pb.p[k] = Interval(p0[6 * B + k]);
                pb.r[k] = Interval(p0[6 * B + 3 + k]);
            }
        }
        const Mat3I RA = rotation_matrix(pa.r), RB = rotation_matrix(pb.r);
        Vec3I d;
        for (int k = 0; k < 3; k++) d[k] = pa.p[k] - pb.p[k];
        // body A's vertices in body B's frame, boxes grown by r
        const int va0 = s.bv[A], nva = s.bv[A + 1] - va0;
        std::vector<Box> av(nva);
        for (int i = 0; i < nva; i++) {
            const double* v = &s.V[3 * (va0 + i)];
            Interval t[3];
            for (int j = 0; j < 3; j++) t[j] = ((v[0] * RA.m[j][0] + v[1] * RA.m[j][1]) + v[2] * RA.m[j][2]) + d[j];
            for (int m = 0; m < 3; m++) {
                const Interval w = (t[0] * RB.m[0][m] + t[1] * RB.m[1][m]) + t[2] * RB.m[2][m];
                av[i].mn[m] = (w - r).lo;
                av[i].mx[m] = (w + r).hi;
            }
            av[i].finish();
        }
        // body B's vertices in its own frame, boxes grown by r (doubles)
        const int vb0 = s.bv[B], nvb = s.bv[B + 1] - vb0;
        std::vector<Box> bv(nvb);
        for (int i = 0; i < nvb; i++) {
            for (int k = 0; k < 3; k++) {
                bv[i].mn[k] = s.V[3 * (vb0 + i) + k] - r;
                bv[i].mx[k] = s.V[3 * (vb0 + i) + k] + r;
            }
            bv[i].finish();
        }
        Candidates& c = part[pi];
        for (int fa = s.bf[A]; fa < s.bf[A + 1]; fa++) {
            const int* FA = &s.F[3 * fa];
            const Box fa_box = merge(av[FA[0] - va0], av[FA[1] - va0], av[FA[2] - va0]);
            double qmn[3], qmx[3];
            for (int k = 0; k < 3; k++) {
                qmn[k] = fa_box.mn[k] - r;
                qmx[k] = fa_box.mx[k] + r;
            }
            for (int fb = s.bf[B]; fb < s.bf[B + 1]; fb++) {
                const int* FB = &s.F[3 * fb];
                // the BVH's primitive box: B's face from its local vertices
                bool hit = true;
                for (int k = 0; k < 3; k++) {
                    const double x0 = s.V[3 * FB[0] + k], x1 = s.V[3 * FB[1] + k], x2 = s.V[3 * FB[2] + k];
                    const double mn = std::min(std::min(x0, x1), x2), mx = std::max(std::max(x0, x1), x2);
                    hit = hit && qmn[k] <= mx && mn <= qmx[k];
                }
                if (!hit) continue;
                const Box fb_box = merge(bv[FB[0] - vb0], bv[FB[1] - vb0], bv[FB[2] - vb0]);
                for (int j = 0; j < 3; j++) {
                    const int va = FA[j];
                    if (v2f[va] == fa && overlapping(av[va - va0], fb_box)) c.fv.push_back({ fb, va });
                    const int vb = FB[j];
                    if (v2f[vb] == fb && overlapping(fa_box, bv[vb - vb0])) c.fv.push_back({ fa, vb });
                }
                for (int ja = 0; ja < 3; ja++) {
                    const int ea = s.FE[3 * fa + ja];
                    if (e2f[ea] != fa) continue;
                    const Box ea_box = merge(av[s.E[2 * ea] - va0], av[s.E[2 * ea + 1] - va0]);
                    for (int jb = 0; jb < 3; jb++) {
                        const int eb = s.FE[3 * fb + jb];
                        if (e2f[eb] != fb) continue;
                        const Box eb_box = merge(bv[s.E[2 * eb] - vb0], bv[s.E[2 * eb + 1] - vb0]);
                        if (overlapping(ea_box, eb_box)) c.ee.push_back({ ea, eb });
                    }
                }
            }
        }
    }
    for (const Candidates& c : part) {
        out.ee.insert(out.ee.end(), c.ee.begin(), c.ee.end());
        out.fv.insert(out.fv.end(), c.fv.begin(), c.fv.end());
    }
    std::sort(out.ee.begin(), out.ee.end());
    std::sort(out.fv.begin(), out.fv.end());
}

}  // namespace oracle
