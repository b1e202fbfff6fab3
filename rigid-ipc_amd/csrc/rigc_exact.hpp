// Exact orientation sign and the edge-triangle intersection predicate.
//
//   is_edge_intersecting_triangle (ipc-toolkit @d41d6ca, ipc/utils/intersection.cpp,
//     not in the reference tree; called by RigidBodyProblem::detect_intersections,
//     src/physics/rigid_body_problem.cpp:343-355) over igl::predicates::orient3d.
//
// orient3d(a, b, c, d) is the sign of det [a - d; b - d; c - d], expanded
// along the z column. A floating-point evaluation decides it when its
// magnitude exceeds an error bound; otherwise the determinant is evaluated
// exactly in expansion arithmetic (sums of non-overlapping doubles,
// J. R. Shewchuk, "Adaptive Precision Floating-Point Arithmetic and Fast
// Robust Geometric Predicates", 1997: the two-sum and two-product
// error-free transforms, growing an expansion by one double, scaling an
// expansion by one double), whose sign is that of its largest component.
//
// Error bound of the filter: with unit roundoff u = 2^-53, each difference
// carries a relative error u, each 2x2 minor 4u of the sum of its product
// magnitudes, each column term 6u of its share of the permanent P (the same
// expression with every product and difference replaced by its magnitude),
// and the two final sums 2u of P: |det_fp - det| <= 8u P + O(u^2) P. The
// filter decides when |det_fp| > 2e-15 P (about 18u P). Below 1e-250 P is
// left to the exact path, which assumes no product underflows (coordinate
// differences above about 1e-100 or exactly zero).
//
// Intersection rule (closed segment against closed triangle): false when
// both end points are strictly on one side of the triangle's plane or both
// lie in it (a segment in the plane of the triangle is not counted, as in
// the restated routine); otherwise the segment meets the plane in one
// point, which is in the closed triangle iff orient3d(p, q, t0, t1),
// orient3d(p, q, t1, t2) and orient3d(p, q, t2, t0) have no two opposite
// strict signs.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace rigc {

#ifndef RIGC_HD
#define RIGC_HD __host__ __device__ inline
#endif

namespace exact {

// s + e == a + b exactly (round to nearest, no contraction).
RIGC_HD void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bv = s - a;
    const double av = s - bv;
    e = (a - av) + (b - bv);
}
// p + e == a * b exactly (fma gives the rounding error of the product).
RIGC_HD void two_prod(double a, double b, double& p, double& e) {
    p = a * b;
    e = fma(a, b, -p);
}

// h = e + b. e: non-overlapping, increasing magnitude (zeros allowed); h the
// same without zeros (a single 0 when the sum is zero). h may alias e.
// Returns the length of h (<= n + 1).
RIGC_HD int grow(const double* e, int n, double b, double* h) {
    double q = b;
    int k = 0;
    for (int i = 0; i < n; i++) {
        double s, t;
        two_sum(q, e[i], s, t);
        if (t != 0.0) h[k++] = t;
        q = s;
    }
    if (q != 0.0 || k == 0) h[k++] = q;
    return k;
}

// h += f (h holds n components and room for n + m). Returns the new length.
RIGC_HD int add(double* h, int n, const double* f, int m) {
    for (int j = 0; j < m; j++) n = grow(h, n, f[j], h);
    return n;
}

// h = b * e (length <= 2 n), same conventions as grow; h must not alias e.
RIGC_HD int scale(const double* e, int n, double b, double* h) {
    double q, t;
    two_prod(e[0], b, q, t);
    int k = 0;
    if (t != 0.0) h[k++] = t;
    for (int i = 1; i < n; i++) {
        double p, pe, s, se, r, re;
        two_prod(e[i], b, p, pe);
        two_sum(q, pe, s, se);
        if (se != 0.0) h[k++] = se;
        two_sum(p, s, r, re);
        if (re != 0.0) h[k++] = re;
        q = r;
    }
    if (q != 0.0 || k == 0) h[k++] = q;
    return k;
}

// h = e * f with e and f of two components each (length <= 8).
RIGC_HD int mul2(const double* e, const double* f, double* h) {
    double t[4];
    int n = scale(e, 2, f[0], h);
    const int m = scale(e, 2, f[1], t);
    return add(h, n, t, m);
}

// h = x * g, x of two components, g of n <= 16 (length <= 4 n).
RIGC_HD int mul_2n(const double* x, const double* g, int n, double* h) {
    double t[32];
    int k = scale(g, n, x[0], h);
    const int m = scale(g, n, x[1], t);
    return add(h, k, t, m);
}

// Exact sign of det [a - d; b - d; c - d].
RIGC_HD int orient3d_exact(const double* a, const double* b, const double* c, const double* d) {
    double ad[3][2], bd[3][2], cd[3][2];  // exact differences, (error, rounded)
    for (int k = 0; k < 3; k++) {
        two_sum(a[k], -d[k], ad[k][1], ad[k][0]);
        two_sum(b[k], -d[k], bd[k][1], bd[k][0]);
        two_sum(c[k], -d[k], cd[k][1], cd[k][0]);
    }
    // det = adz (bdx cdy - cdx bdy) + bdz (cdx ady - adx cdy) + cdz (adx bdy - bdx ady)
    const double* col[3][5] = { { ad[2], bd[0], cd[1], cd[0], bd[1] },
                                { bd[2], cd[0], ad[1], ad[0], cd[1] },
                                { cd[2], ad[0], bd[1], bd[0], ad[1] } };
    double det[192];
    int nd = 0;
    for (int r = 0; r < 3; r++) {
        double p1[8], p2[8], minor[16], term[64];
        const int n1 = mul2(col[r][1], col[r][2], p1);
        const int n2 = mul2(col[r][3], col[r][4], p2);
        for (int i = 0; i < n2; i++) p2[i] = -p2[i];
        for (int i = 0; i < n1; i++) minor[i] = p1[i];
        const int nm = add(minor, n1, p2, n2);
        const int nt = mul_2n(col[r][0], minor, nm, term);
        if (r == 0) {
            for (int i = 0; i < nt; i++) det[i] = term[i];
            nd = nt;
        } else {
            nd = add(det, nd, term, nt);
        }
    }
    const double top = det[nd - 1];
    return (top > 0.0) - (top < 0.0);
}

// Filtered sign: +1 / -1 when the floating-point determinant decides it,
// 2 when it does not (exact zeros included).
RIGC_HD int orient3d_filtered(const double* a, const double* b, const double* c, const double* d) {
    const double adx = a[0] - d[0], ady = a[1] - d[1], adz = a[2] - d[2];
    const double bdx = b[0] - d[0], bdy = b[1] - d[1], bdz = b[2] - d[2];
    const double cdx = c[0] - d[0], cdy = c[1] - d[1], cdz = c[2] - d[2];
    const double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    const double cdxady = cdx * ady, adxcdy = adx * cdy;
    const double adxbdy = adx * bdy, bdxady = bdx * ady;
    const double det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
    const double perm = (fabs(bdxcdy) + fabs(cdxbdy)) * fabs(adz) + (fabs(cdxady) + fabs(adxcdy)) * fabs(bdz)
        + (fabs(adxbdy) + fabs(bdxady)) * fabs(cdz);
    if (!(perm >= 1e-250)) return 2;
    const double bound = 2e-15 * perm;
    if (det > bound) return 1;
    if (-det > bound) return -1;
    return 2;
}

RIGC_HD int orient3d(const double* a, const double* b, const double* c, const double* d) {
    const int s = orient3d_filtered(a, b, c, d);
    return s != 2 ? s : orient3d_exact(a, b, c, d);
}

// The intersection rule with exact signs.
RIGC_HD bool edge_triangle(const double* p, const double* q, const double* t0, const double* t1, const double* t2) {
    const int o1 = orient3d(t0, t1, t2, p), o2 = orient3d(t0, t1, t2, q);
    if (o1 != 0 && o1 == o2) return false;
    if (o1 == 0 && o2 == 0) return false;
    const int o3 = orient3d(p, q, t0, t1), o4 = orient3d(p, q, t1, t2), o5 = orient3d(p, q, t2, t0);
    const bool pos = o3 > 0 || o4 > 0 || o5 > 0, neg = o3 < 0 || o4 < 0 || o5 < 0;
    return !(pos && neg);
}

// The rule with filtered signs only: 1 intersecting, 0 not, -1 undecided.
RIGC_HD int edge_triangle_filtered(const double* p, const double* q, const double* t0, const double* t1,
                                   const double* t2) {
    const int o1 = orient3d_filtered(t0, t1, t2, p), o2 = orient3d_filtered(t0, t1, t2, q);
    if (o1 != 2 && o1 == o2) return 0;
    if (o1 == 2 || o2 == 2) return -1;
    const int o[3] = { orient3d_filtered(p, q, t0, t1), orient3d_filtered(p, q, t1, t2),
                       orient3d_filtered(p, q, t2, t0) };
    bool pos = false, neg = false, open = false;
    for (int k = 0; k < 3; k++) {
        pos = pos || o[k] == 1;
        neg = neg || o[k] == -1;
        open = open || o[k] == 2;
    }
    if (pos && neg) return 0;
    return open ? -1 : 1;
}

}  // namespace exact
}  // namespace rigc
