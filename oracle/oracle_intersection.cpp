// ORACLE — test infrastructure only (see oracle_interval.hpp).
//
// Restates the 3D branch of RigidBodyProblem::detect_intersections
// (src/physics/rigid_body_problem.cpp:289-360):
//   close_bodies(poses, poses, 0)            src/physics/rigid_body_assembler.cpp:268-356
//     (discrete poses: each body's box is the box of its world vertices,
//     RigidBody::compute_bounding_box, src/physics/rigid_body.cpp:246-272)
//   detect_intersection_candidates_rigid_bvh src/ccd/rigid/broad_phase.cpp:265-305
//   is_edge_intersecting_triangle            ipc-toolkit @d41d6ca (not in the
//     reference tree) over igl::predicates::orient3d
// but returns every intersecting (edge, face) pair instead of stopping at
// the first (the reference's answer is "the list is not empty"). Candidate
// pairs are every edge of one body and face of the other whose boxes of
// world vertices overlap (inclusive): a superset of the reference's BVH
// candidates, and the pairs that intersect do not depend on which
// superset is used.
//
// orient3d is the exact sign of det [a - d; b - d; c - d]: all twelve
// coordinates are scaled by the smallest power of two among them to
// integers, and the determinant is evaluated in signed multi-word integer
// arithmetic (row expansion), so no rounding occurs anywhere.
//
// Intersection rule (closed segment, closed triangle): false when both end
// points are strictly on one side of the triangle's plane, or both lie in
// it; otherwise true iff orient3d(p, q, t0, t1), orient3d(p, q, t1, t2) and
// orient3d(p, q, t2, t0) have no two opposite strict signs. Parity with
// ipc-toolkit for segments touching the triangle (a zero orientation) is
// unpinned: the reference tree holds neither the routine nor a test of it.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <vector>

#include "oracle_scene.hpp"

namespace oracle {

namespace {

// Signed integer: magnitude in 32-bit words, least significant first.
struct Big {
    bool neg = false;
    std::vector<uint32_t> m;
    void trim() {
        while (!m.empty() && m.back() == 0) m.pop_back();
        if (m.empty()) neg = false;
    }
    int sign() const { return m.empty() ? 0 : (neg ? -1 : 1); }
};

int cmp_mag(const Big& a, const Big& b) {
    if (a.m.size() != b.m.size()) return a.m.size() < b.m.size() ? -1 : 1;
    for (size_t i = a.m.size(); i-- > 0;)
        if (a.m[i] != b.m[i]) return a.m[i] < b.m[i] ? -1 : 1;
    return 0;
}

std::vector<uint32_t> add_mag(const std::vector<uint32_t>& a, const std::vector<uint32_t>& b) {
    std::vector<uint32_t> r(std::max(a.size(), b.size()) + 1, 0);
    uint64_t carry = 0;
    for (size_t i = 0; i < r.size(); i++) {
        const uint64_t s = carry + (i < a.size() ? a[i] : 0) + (i < b.size() ? b[i] : 0);
        r[i] = (uint32_t)s;
        carry = s >> 32;
    }
    return r;
}

// |a| - |b| with |a| >= |b|.
std::vector<uint32_t> sub_mag(const std::vector<uint32_t>& a, const std::vector<uint32_t>& b) {
    std::vector<uint32_t> r(a.size(), 0);
    int64_t borrow = 0;
    for (size_t i = 0; i < a.size(); i++) {
        int64_t d = (int64_t)a[i] - borrow - (i < b.size() ? (int64_t)b[i] : 0);
        borrow = d < 0;
        if (d < 0) d += (int64_t)1 << 32;
        r[i] = (uint32_t)d;
    }
    return r;
}

Big add(const Big& a, const Big& b) {
    Big r;
    if (a.neg == b.neg) {
        r.m = add_mag(a.m, b.m);
        r.neg = a.neg;
    } else if (cmp_mag(a, b) >= 0) {
        r.m = sub_mag(a.m, b.m);
        r.neg = a.neg;
    } else {
        r.m = sub_mag(b.m, a.m);
        r.neg = b.neg;
    }
    r.trim();
    return r;
}

Big negate(Big a) {
    if (!a.m.empty()) a.neg = !a.neg;
    return a;
}

Big sub(const Big& a, const Big& b) { return add(a, negate(b)); }

Big mul(const Big& a, const Big& b) {
    Big r;
    if (a.m.empty() || b.m.empty()) return r;
    r.m.assign(a.m.size() + b.m.size(), 0);
    for (size_t i = 0; i < a.m.size(); i++) {
        uint64_t carry = 0;
        for (size_t j = 0; j < b.m.size(); j++) {
            const uint64_t t = (uint64_t)a.m[i] * b.m[j] + r.m[i + j] + carry;
            r.m[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        size_t k = i + b.m.size();
        while (carry) {
            const uint64_t t = (uint64_t)r.m[k] + carry;
            r.m[k] = (uint32_t)t;
            carry = t >> 32;
            k++;
        }
    }
    r.neg = a.neg != b.neg;
    r.trim();
    return r;
}

// x = mantissa * 2^exponent with an integer mantissa below 2^53 (x != 0, finite).
void split(double x, uint64_t& mant, int& expo) {
    int e;
    const double f = std::frexp(std::fabs(x), &e);
    mant = (uint64_t)std::ldexp(f, 53);
    expo = e - 53;
}

// x * 2^-base as an integer (base <= the exponent of x).
Big scaled(double x, int base) {
    Big r;
    if (x == 0.0) return r;
    uint64_t mant;
    int expo;
    split(x, mant, expo);
    const int shift = expo - base;
    const int words = shift / 32, bits = shift % 32;
    r.m.assign((size_t)words + 3, 0);
    const unsigned __int128 v = (unsigned __int128)mant << bits;
    r.m[words] = (uint32_t)v;
    r.m[words + 1] = (uint32_t)(v >> 32);
    r.m[words + 2] = (uint32_t)(v >> 64);
    r.neg = x < 0;
    r.trim();
    return r;
}

}  // namespace

int orient3d_sign(const double* a, const double* b, const double* c, const double* d) {
    const double* pts[4] = { a, b, c, d };
    int base = 1 << 30;
    for (const double* p : pts)
        for (int k = 0; k < 3; k++)
            if (p[k] != 0.0) {
                uint64_t mant;
                int expo;
                split(p[k], mant, expo);
                base = std::min(base, expo);
            }
    if (base == (1 << 30)) return 0;  // all coordinates zero
    Big X[4][3];
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 3; k++) X[i][k] = scaled(pts[i][k], base);
    Big r[3][3];  // rows a - d, b - d, c - d
    for (int i = 0; i < 3; i++)
        for (int k = 0; k < 3; k++) r[i][k] = sub(X[i][k], X[3][k]);
    // det = r00 (r11 r22 - r12 r21) - r01 (r10 r22 - r12 r20) + r02 (r10 r21 - r11 r20)
    const Big m0 = sub(mul(r[1][1], r[2][2]), mul(r[1][2], r[2][1]));
    const Big m1 = sub(mul(r[1][0], r[2][2]), mul(r[1][2], r[2][0]));
    const Big m2 = sub(mul(r[1][0], r[2][1]), mul(r[1][1], r[2][0]));
    const Big det = add(sub(mul(r[0][0], m0), mul(r[0][1], m1)), mul(r[0][2], m2));
    return det.sign();
}

bool edge_triangle_intersect(const double* p, const double* q, const double* t0, const double* t1,
                             const double* t2) {
    const int o1 = orient3d_sign(t0, t1, t2, p), o2 = orient3d_sign(t0, t1, t2, q);
    if (o1 != 0 && o1 == o2) return false;
    if (o1 == 0 && o2 == 0) return false;
    const int o3 = orient3d_sign(p, q, t0, t1), o4 = orient3d_sign(p, q, t1, t2), o5 = orient3d_sign(p, q, t2, t0);
    const bool pos = o3 > 0 || o4 > 0 || o5 > 0, neg = o3 < 0 || o4 < 0 || o5 < 0;
    return !(pos && neg);
}

namespace {
struct Box3 {
    double mn[3], mx[3];
};
inline bool meet(const Box3& a, const Box3& b) {
    for (int k = 0; k < 3; k++)
        if (!(a.mn[k] <= b.mx[k] && b.mn[k] <= a.mx[k])) return false;
    return true;
}
Box3 box_of(std::initializer_list<const double*> pts) {
    Box3 b;
    for (int k = 0; k < 3; k++) { b.mn[k] = INFINITY; b.mx[k] = -INFINITY; }
    for (const double* p : pts)
        for (int k = 0; k < 3; k++) {
            b.mn[k] = std::min(b.mn[k], p[k]);
            b.mx[k] = std::max(b.mx[k], p[k]);
        }
    return b;
}
}  // namespace

// Sorted (edge, face) pairs that intersect at `poses`.
std::vector<std::array<int, 2>> detect_intersections(const Scene& s, const double* poses) {
    std::vector<double> W(3 * (size_t)s.nv);
    std::vector<double> w;
    for (int b = 0; b < s.nb; b++) {
        world_vertices(s, b, poses + 6 * b, w);
        std::copy(w.begin(), w.end(), W.begin() + 3 * (size_t)s.bv[b]);
    }
    std::vector<Box3> body(s.nb);
    for (int b = 0; b < s.nb; b++) {
        Box3& x = body[b];
        for (int k = 0; k < 3; k++) { x.mn[k] = INFINITY; x.mx[k] = -INFINITY; }
        for (int i = s.bv[b]; i < s.bv[b + 1]; i++)
            for (int k = 0; k < 3; k++) {
                x.mn[k] = std::min(x.mn[k], W[3 * i + k]);
                x.mx[k] = std::max(x.mx[k], W[3 * i + k]);
            }
    }
    std::vector<std::array<int, 2>> out;
    for (int i = 0; i < s.nb; i++)
        for (int j = i + 1; j < s.nb; j++) {
            if (s.group[i] == s.group[j] || !meet(body[i], body[j])) continue;
            for (int dir = 0; dir < 2; dir++) {
                const int X = dir ? j : i, Y = dir ? i : j;
                for (int e = s.be[X]; e < s.be[X + 1]; e++) {
                    const double* p = &W[3 * s.E[2 * e]];
                    const double* q = &W[3 * s.E[2 * e + 1]];
                    const Box3 eb = box_of({ p, q });
                    for (int f = s.bf[Y]; f < s.bf[Y + 1]; f++) {
                        const double* t0 = &W[3 * s.F[3 * f]];
                        const double* t1 = &W[3 * s.F[3 * f + 1]];
                        const double* t2 = &W[3 * s.F[3 * f + 2]];
                        if (!meet(eb, box_of({ t0, t1, t2 }))) continue;
                        if (edge_triangle_intersect(p, q, t0, t1, t2)) out.push_back({ e, f });
                    }
                }
            }
        }
    std::sort(out.begin(), out.end());
    return out;
}

}  // namespace oracle
