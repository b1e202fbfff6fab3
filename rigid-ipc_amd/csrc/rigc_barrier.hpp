// Contact-potential math shared by host and device code.
//
// Distance-type classification and squared distances of the constraint
// set, the IPC log barrier on squared distances, the edge-edge mollifier,
// friction tangent frames and the smoothed friction kernel, and the
// rotation-vector derivatives of a body's rotation matrix. These are the
// pieces of ipc-toolkit @d41d6ca that DistanceBarrierRBProblem calls
// (src/problems/distance_barrier_rb_problem.cpp:1299-1572,
// src/opt/distance_barrier_constraint.cpp:267-323) and of
// RigidBody::world_vertices_diff (src/physics/rigid_body.tpp:26-78,
// src/utils/sinc.cpp:103-142).
//
// Values are evaluated with the same operation order as the CPU
// restatement (no FMA contraction), so distance types, distances and
// activity decisions agree bit for bit. Derivatives are closed forms: a
// squared distance is written as P / Q over at most three difference
// vectors y_a = x_i - x_j,
//   point-point   P = |a|^2,                 Q = 1
//   point-line    P = |a x b|^2,             Q = |b - a|^2
//   point-plane / line-line
//                 P = (w . (u x v))^2,       Q = |u x v|^2
// its gradient and Hessian are formed in difference space and mapped to
// the constraint's vertices through the +-1 incidence sig[slot][a].
#pragma once

#include "rigc_geom.hpp"

namespace rigc {

enum { CON_VV = 0, CON_EV = 1, CON_EE = 2, CON_FV = 3 };
enum { PE_E0 = 0, PE_E1, PE_E };
enum { EE_A0B0 = 0, EE_A0B1, EE_A1B0, EE_A1B1, EE_AB0, EE_AB1, EE_A0B, EE_A1B, EE_AB };
enum { PT_T0 = 0, PT_T1, PT_T2, PT_E0, PT_E1, PT_E2, PT_T };

RIGC_HD double v3_dot(const double* a, const double* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
RIGC_HD void v3_sub(const double* a, const double* b, double* r) {
    r[0] = a[0] - b[0];
    r[1] = a[1] - b[1];
    r[2] = a[2] - b[2];
}
RIGC_HD void v3_cross(const double* a, const double* b, double* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}
// [x]_x (k, l): the cross-product matrix, [x]_x y = x cross y.
RIGC_HD double v3_skew(const double* x, int k, int l) {
    if (k == l) return 0.0;
    const int m = 3 - k - l;
    // +x_m for (k,l) in {(1,0),(2,1),(0,2)}, -x_m for the transposed slots
    const bool pos = (k == (l + 1) % 3);
    return pos ? x[m] : -x[m];
}
// [e_c]_x (i, j) for the unit vector e_c.
RIGC_HD double v3_skew_unit(int c, int i, int j) {
    if (i == j || 3 - i - j != c) return 0.0;
    return (i == (j + 1) % 3) ? 1.0 : -1.0;
}

// ------------------------------------------------------------ classification
RIGC_HD int point_edge_type(const double* p, const double* e0, const double* e1) {
    double e[3], pe[3];
    v3_sub(e1, e0, e);
    v3_sub(p, e0, pe);
    const double l2 = v3_dot(e, e);
    if (l2 == 0) return PE_E0;
    const double ratio = v3_dot(e, pe) / l2;
    if (ratio < 0) return PE_E0;
    if (ratio > 1) return PE_E1;
    return PE_E;
}

RIGC_HD double point_segment_sq(const double* p, const double* a, const double* b) {
    double ab[3], ap[3], c[3];
    v3_sub(b, a, ab);
    v3_sub(p, a, ap);
    const double l2 = v3_dot(ab, ab);
    double t = l2 > 0 ? v3_dot(ap, ab) / l2 : 0;
    t = t < 0.0 ? 0.0 : t;
    t = t > 1.0 ? 1.0 : t;
    for (int k = 0; k < 3; k++) c[k] = a[k] + t * ab[k] - p[k];
    return v3_dot(c, c);
}

// Closest-feature pair of two segments (Sunday's segment-segment
// parametrisation; nearly parallel segments fall back to the closest
// endpoint-segment pair).
RIGC_HD int edge_edge_type(const double* ea0, const double* ea1, const double* eb0, const double* eb1) {
    double u[3], v[3], w[3], uv[3];
    v3_sub(ea1, ea0, u);
    v3_sub(eb1, eb0, v);
    v3_sub(ea0, eb0, w);
    const double a = v3_dot(u, u), b = v3_dot(u, v), c = v3_dot(v, v), d = v3_dot(u, w), e = v3_dot(v, w);
    const double D = a * c - b * b;
    if (a == 0.0 && c == 0.0) return EE_A0B0;
    if (a == 0.0) return EE_A0B;
    if (c == 0.0) return EE_AB0;
    v3_cross(u, v, uv);
    const double ac = a * c;
    if (v3_dot(uv, uv) < 1e-20 * (1.0 > ac ? 1.0 : ac)) {
        const double dd[4] = { point_segment_sq(ea0, eb0, eb1), point_segment_sq(ea1, eb0, eb1),
                               point_segment_sq(eb0, ea0, ea1), point_segment_sq(eb1, ea0, ea1) };
        int k = 0;
        for (int i = 1; i < 4; i++)
            if (dd[i] < dd[k]) k = i;
        if (k < 2) {
            const int t = point_edge_type(k == 0 ? ea0 : ea1, eb0, eb1);
            if (t == PE_E) return k == 0 ? EE_A0B : EE_A1B;
            return k == 0 ? (t == PE_E0 ? EE_A0B0 : EE_A0B1) : (t == PE_E0 ? EE_A1B0 : EE_A1B1);
        }
        const int t = point_edge_type(k == 2 ? eb0 : eb1, ea0, ea1);
        if (t == PE_E) return k == 2 ? EE_AB0 : EE_AB1;
        return k == 2 ? (t == PE_E0 ? EE_A0B0 : EE_A1B0) : (t == PE_E0 ? EE_A0B1 : EE_A1B1);
    }
    int def = EE_AB;
    const double sN = b * e - c * d;
    double tN, tD;
    if (sN <= 0.0) {
        tN = e;
        tD = c;
        def = EE_A0B;
    } else if (sN >= D) {
        tN = e + b;
        tD = c;
        def = EE_A1B;
    } else {
        tN = a * e - b * d;
        tD = D;
    }
    if (tN <= 0.0) {
        if (-d <= 0.0) return EE_A0B0;
        if (-d >= a) return EE_A1B0;
        return EE_AB0;
    }
    if (tN >= tD) {
        if ((-d + b) <= 0.0) return EE_A0B1;
        if ((-d + b) >= a) return EE_A1B1;
        return EE_AB1;
    }
    return def;
}

RIGC_HD void solve_2x2(double m00, double m01, double m11, double b0, double b1, double& x0, double& x1) {
    const double det = m00 * m11 - m01 * m01;
    x0 = (m11 * b0 - m01 * b1) / det;
    x1 = (m00 * b1 - m01 * b0) / det;
}

// Symmetric 2x2 solve in the manner of an LDL^T with diagonal pivoting:
// the larger diagonal entry first, D pseudo-inverted (a zero pivot gives a
// zero component), as ipc-toolkit's closest-point coordinates (Eigen LDLT).
RIGC_HD void solve_2x2_ldlt(double a00, double a01, double a11, double b0, double b1, double& x0, double& x1) {
    const bool swap = fabs(a11) > fabs(a00);
    if (swap) {
        const double t = a00; a00 = a11; a11 = t;
        const double u = b0; b0 = b1; b1 = u;
    }
    const double d0 = a00;
    const double l10 = d0 != 0.0 ? a01 / d0 : a01;
    const double d1 = a11 - l10 * (d0 * l10);
    const double y0 = b0, y1 = b1 - l10 * y0;
    const double tiny = 2.2250738585072014e-308;
    const double z0 = fabs(d0) > tiny ? y0 / d0 : 0.0;
    const double z1 = fabs(d1) > tiny ? y1 / d1 : 0.0;
    const double w1 = z1, w0 = z0 - l10 * w1;
    x0 = swap ? w1 : w0;
    x1 = swap ? w0 : w1;
}

RIGC_HD int point_triangle_type(const double* p, const double* t0, const double* t1, const double* t2) {
    double e01[3], e02[3], n[3];
    v3_sub(t1, t0, e01);
    v3_sub(t2, t0, e02);
    v3_cross(e01, e02, n);
    const double* T[3] = { t0, t1, t2 };
    double par[3][2];
    for (int k = 0; k < 3; k++) {
        double r0[3], r1[3], d[3];
        v3_sub(T[(k + 1) % 3], T[k], r0);
        v3_cross(r0, n, r1);  // in-plane normal of edge k, pointing out of the triangle
        v3_sub(p, T[k], d);
        solve_2x2(v3_dot(r0, r0), v3_dot(r0, r1), v3_dot(r1, r1), v3_dot(r0, d), v3_dot(r1, d), par[k][0], par[k][1]);
        if (par[k][0] > 0.0 && par[k][0] < 1.0 && par[k][1] >= 0.0) return PT_E0 + k;
    }
    if (par[0][0] <= 0.0 && par[2][0] >= 1.0) return PT_T0;
    if (par[1][0] <= 0.0 && par[0][0] >= 1.0) return PT_T1;
    if (par[2][0] <= 0.0 && par[1][0] >= 1.0) return PT_T2;
    return PT_T;
}

// ------------------------------------------------------------ distance forms
// Squared distance of one closest-feature pair as P / Q over difference
// vectors y[a], with gradients in difference space (index 3 a + k).
struct DistForm {
    int kind;              // 0 point-point, 1 point-line, 2 point-plane / line-line
    signed char sig[4][3]; // constraint slot -> coefficient of difference vector a
    double y[3][3];
    double c[3];           // point-line: a x b; planes: u x v
    double s;              // planes: w . (u x v)
    double yy[3];          // point-line: |a|^2, |b|^2, a.b; planes: |u|^2, |v|^2, u.v
    double P, Q, iQ, d;
    double gP[9], gQ[9], gd[9];
};

RIGC_HD void form_clear(DistForm& f) {
    for (int i = 0; i < 4; i++)
        for (int a = 0; a < 3; a++) f.sig[i][a] = 0;
    for (int i = 0; i < 9; i++) f.gP[i] = f.gQ[i] = f.gd[i] = 0.0;
    for (int a = 0; a < 3; a++)
        for (int k = 0; k < 3; k++) f.y[a][k] = 0.0;
    f.c[0] = f.c[1] = f.c[2] = 0.0;
    f.s = 0.0;
}

RIGC_HD void form_finish(DistForm& f) {
    for (int i = 0; i < 9; i++) f.gd[i] = (f.gP[i] - f.d * f.gQ[i]) * f.iQ;
}

RIGC_HD void form_point_point(DistForm& f, const double* X, int p, int q) {
    form_clear(f);
    f.kind = 0;
    v3_sub(X + 3 * p, X + 3 * q, f.y[0]);
    f.sig[p][0] = 1;
    f.sig[q][0] = -1;
    f.P = v3_dot(f.y[0], f.y[0]);
    f.Q = 1.0;
    f.iQ = 1.0;
    f.d = f.P;
    for (int k = 0; k < 3; k++) f.gP[k] = 2.0 * f.y[0][k];
    form_finish(f);
}

// |(e0 - p) x (e1 - p)|^2 / |e1 - e0|^2
RIGC_HD void form_point_line(DistForm& f, const double* X, int p, int e0, int e1) {
    form_clear(f);
    f.kind = 1;
    double* a = f.y[0];
    double* b = f.y[1];
    v3_sub(X + 3 * e0, X + 3 * p, a);
    v3_sub(X + 3 * e1, X + 3 * p, b);
    f.sig[p][0] = -1;
    f.sig[p][1] = -1;
    f.sig[e0][0] = 1;
    f.sig[e1][1] = 1;
    v3_cross(a, b, f.c);
    double e[3];
    v3_sub(X + 3 * e1, X + 3 * e0, e);
    f.P = v3_dot(f.c, f.c);
    f.Q = v3_dot(e, e);
    f.iQ = 1.0 / f.Q;
    f.d = f.P * f.iQ;
    f.yy[0] = v3_dot(a, a);
    f.yy[1] = v3_dot(b, b);
    f.yy[2] = v3_dot(a, b);
    double t[3];
    v3_cross(b, f.c, t);
    for (int k = 0; k < 3; k++) f.gP[k] = 2.0 * t[k];
    v3_cross(f.c, a, t);
    for (int k = 0; k < 3; k++) f.gP[3 + k] = 2.0 * t[k];
    for (int k = 0; k < 3; k++) {
        f.gQ[k] = -2.0 * e[k];
        f.gQ[3 + k] = 2.0 * e[k];
    }
    form_finish(f);
}

// (w . (u x v))^2 / |u x v|^2 with w, u, v given by the caller's incidence.
RIGC_HD void form_plane_common(DistForm& f) {
    f.kind = 2;
    const double* w = f.y[0];
    const double* u = f.y[1];
    const double* v = f.y[2];
    v3_cross(u, v, f.c);
    f.s = v3_dot(w, f.c);
    f.P = f.s * f.s;
    f.Q = v3_dot(f.c, f.c);
    f.iQ = 1.0 / f.Q;
    f.d = f.P * f.iQ;
    f.yy[0] = v3_dot(u, u);
    f.yy[1] = v3_dot(v, v);
    f.yy[2] = v3_dot(u, v);
    double t[3];
    // grad s = (u x v, v x w, w x u); grad P = 2 s grad s
    for (int k = 0; k < 3; k++) f.gP[k] = 2.0 * f.s * f.c[k];
    v3_cross(v, w, t);
    for (int k = 0; k < 3; k++) f.gP[3 + k] = 2.0 * f.s * t[k];
    v3_cross(w, u, t);
    for (int k = 0; k < 3; k++) f.gP[6 + k] = 2.0 * f.s * t[k];
    // grad |u x v|^2 = (0, 2 v x n, 2 n x u)
    v3_cross(v, f.c, t);
    for (int k = 0; k < 3; k++) f.gQ[3 + k] = 2.0 * t[k];
    v3_cross(f.c, u, t);
    for (int k = 0; k < 3; k++) f.gQ[6 + k] = 2.0 * t[k];
    form_finish(f);
}

// point p against the plane of (t0, t1, t2): w = p - t0, u = t1 - t0, v = t2 - t0
RIGC_HD void form_point_plane(DistForm& f, const double* X, int p, int t0, int t1, int t2) {
    form_clear(f);
    v3_sub(X + 3 * t1, X + 3 * t0, f.y[1]);
    v3_sub(X + 3 * t2, X + 3 * t0, f.y[2]);
    v3_sub(X + 3 * p, X + 3 * t0, f.y[0]);
    f.sig[p][0] = 1;
    f.sig[t0][0] = -1;
    f.sig[t0][1] = -1;
    f.sig[t0][2] = -1;
    f.sig[t1][1] = 1;
    f.sig[t2][2] = 1;
    form_plane_common(f);
}

// lines (a0, a1), (b0, b1): w = b0 - a0, u = a1 - a0, v = b1 - b0
RIGC_HD void form_line_line(DistForm& f, const double* X, int a0, int a1, int b0, int b1) {
    form_clear(f);
    v3_sub(X + 3 * a1, X + 3 * a0, f.y[1]);
    v3_sub(X + 3 * b1, X + 3 * b0, f.y[2]);
    v3_sub(X + 3 * b0, X + 3 * a0, f.y[0]);
    f.sig[a0][0] = -1;
    f.sig[a0][1] = -1;
    f.sig[a1][1] = 1;
    f.sig[b0][0] = 1;
    f.sig[b0][2] = -1;
    f.sig[b1][2] = 1;
    form_plane_common(f);
}

// Hessian entry of N = |x cross y|^2 (blocks 0 = x, 1 = y), with
// xx = |x|^2, yy = |y|^2, xy = x.y and c = x cross y.
RIGC_HD double cross_sq_hess(const double* x, const double* y, const double* c, double xx, double yy, double xy,
                             int A, int B, int k, int l) {
    const double dl = (k == l) ? 1.0 : 0.0;
    if (A == 0 && B == 0) return 2.0 * (yy * dl - y[k] * y[l]);
    if (A == 1 && B == 1) return 2.0 * (xx * dl - x[k] * x[l]);
    if (A == 0) return 2.0 * (x[k] * y[l] - xy * dl) - 2.0 * v3_skew(c, k, l);
    return 2.0 * (y[k] * x[l] - xy * dl) + 2.0 * v3_skew(c, k, l);
}

// Difference-space Hessian entry of d = P / Q at (a, k), (b, l):
// (H_P - gd gQ^T - gQ gd^T - d H_Q) / Q.
RIGC_HD double form_hess_diff(const DistForm& f, int A, int k, int B, int l) {
    const double dl = (k == l) ? 1.0 : 0.0;
    double HP = 0.0, HQ = 0.0;
    if (f.kind == 0) {
        HP = 2.0 * dl;
    } else if (f.kind == 1) {
        HP = cross_sq_hess(f.y[0], f.y[1], f.c, f.yy[0], f.yy[1], f.yy[2], A, B, k, l);
        HQ = (A == B ? 2.0 : -2.0) * dl;
    } else {
        // grad s entries: gP = 2 s grad s, recover grad s from the vectors
        const double* w = f.y[0];
        const double* u = f.y[1];
        const double* v = f.y[2];
        double gsA, gsB;
        {
            double t[3];
            if (A == 0) { gsA = f.c[k]; } else if (A == 1) { v3_cross(v, w, t); gsA = t[k]; } else { v3_cross(w, u, t); gsA = t[k]; }
            if (B == 0) { gsB = f.c[l]; } else if (B == 1) { v3_cross(v, w, t); gsB = t[l]; } else { v3_cross(w, u, t); gsB = t[l]; }
        }
        double Hs = 0.0;
        if (A == 0 && B == 1) Hs = -v3_skew(v, k, l);
        else if (A == 1 && B == 0) Hs = v3_skew(v, k, l);
        else if (A == 0 && B == 2) Hs = v3_skew(u, k, l);
        else if (A == 2 && B == 0) Hs = -v3_skew(u, k, l);
        else if (A == 1 && B == 2) Hs = -v3_skew(w, k, l);
        else if (A == 2 && B == 1) Hs = v3_skew(w, k, l);
        HP = 2.0 * gsA * gsB + 2.0 * f.s * Hs;
        if (A > 0 && B > 0) HQ = cross_sq_hess(u, v, f.c, f.yy[0], f.yy[1], f.yy[2], A - 1, B - 1, k, l);
    }
    const int ia = 3 * A + k, ib = 3 * B + l;
    return (HP - f.gd[ia] * f.gQ[ib] - f.gQ[ia] * f.gd[ib] - f.d * HQ) * f.iQ;
}

// Vertex-space gradient of d: slot i, coordinate k.
RIGC_HD double form_grad_vertex(const DistForm& f, int i, int k) {
    double g = 0.0;
    for (int a = 0; a < 3; a++)
        if (f.sig[i][a]) g += f.sig[i][a] * f.gd[3 * a + k];
    return g;
}

// Vertex-space Hessian of d: (slot i, coordinate k), (slot j, coordinate l).
RIGC_HD double form_hess_vertex(const DistForm& f, int i, int k, int j, int l) {
    double h = 0.0;
    for (int a = 0; a < 3; a++) {
        if (!f.sig[i][a]) continue;
        for (int b = 0; b < 3; b++) {
            if (!f.sig[j][b]) continue;
            h += (f.sig[i][a] * f.sig[j][b]) * form_hess_diff(f, a, k, b, l);
        }
    }
    return h;
}

// Constraint vertices X (4 slots x 3): the distance form of the closest
// feature pair at X, with the type recomputed as ipc-toolkit's
// compute_distance does.
RIGC_HD void constraint_form(int type, const double* X, DistForm& f) {
    const double* x0 = X;
    const double* x1 = X + 3;
    const double* x2 = X + 6;
    const double* x3 = X + 9;
    switch (type) {
    case CON_VV: form_point_point(f, X, 0, 1); return;
    case CON_EV:
        switch (point_edge_type(x0, x1, x2)) {
        case PE_E0: form_point_point(f, X, 0, 1); return;
        case PE_E1: form_point_point(f, X, 0, 2); return;
        default: form_point_line(f, X, 0, 1, 2); return;
        }
    case CON_EE:
        switch (edge_edge_type(x0, x1, x2, x3)) {
        case EE_A0B0: form_point_point(f, X, 0, 2); return;
        case EE_A0B1: form_point_point(f, X, 0, 3); return;
        case EE_A1B0: form_point_point(f, X, 1, 2); return;
        case EE_A1B1: form_point_point(f, X, 1, 3); return;
        case EE_AB0: form_point_line(f, X, 2, 0, 1); return;
        case EE_AB1: form_point_line(f, X, 3, 0, 1); return;
        case EE_A0B: form_point_line(f, X, 0, 2, 3); return;
        case EE_A1B: form_point_line(f, X, 1, 2, 3); return;
        default: form_line_line(f, X, 0, 1, 2, 3); return;
        }
    default:
        switch (point_triangle_type(x0, x1, x2, x3)) {
        case PT_T0: form_point_point(f, X, 0, 1); return;
        case PT_T1: form_point_point(f, X, 0, 2); return;
        case PT_T2: form_point_point(f, X, 0, 3); return;
        case PT_E0: form_point_line(f, X, 0, 1, 2); return;
        case PT_E1: form_point_line(f, X, 0, 2, 3); return;
        case PT_E2: form_point_line(f, X, 0, 3, 1); return;
        default: form_point_plane(f, X, 0, 1, 2, 3); return;
        }
    }
}

// ------------------------------------------------------------ barrier
// IPC log barrier b(d; dhat) = -(d - dhat)^2 ln(d / dhat) on squared
// distances, and its first two derivatives in d.
RIGC_HD double ipc_barrier(double d, double dhat) {
    if (d <= 0.0) return HUGE_VAL;
    if (d >= dhat) return 0.0;
    return -(d - dhat) * (d - dhat) * log(d / dhat);
}
RIGC_HD double ipc_barrier_gradient(double d, double dhat) {
    if (d <= 0.0 || d >= dhat) return 0.0;
    return (dhat - d) * (2 * log(d / dhat) - dhat / d + 1);
}
RIGC_HD double ipc_barrier_hessian(double d, double dhat) {
    if (d <= 0.0 || d >= dhat) return 0.0;
    return (dhat / d + 2) * dhat / d - 2 * log(d / dhat) - 3;
}

// Edge-edge mollifier on x = |ea x eb|^2 (ea = X1 - X0, eb = X3 - X2).
struct Mollifier {
    double ea[3], eb[3], n[3];
    double ee_a, ee_b, ee_ab;
    double x, m, m1, m2;  // value and derivatives in x
    double gx[6];         // gradient of x in (ea, eb)
};

RIGC_HD void mollifier_eval(const double* X, double eps_x, Mollifier& M) {
    v3_sub(X + 3, X, M.ea);
    v3_sub(X + 9, X + 6, M.eb);
    v3_cross(M.ea, M.eb, M.n);
    M.x = v3_dot(M.n, M.n);
    M.ee_a = v3_dot(M.ea, M.ea);
    M.ee_b = v3_dot(M.eb, M.eb);
    M.ee_ab = v3_dot(M.ea, M.eb);
    double t[3];
    v3_cross(M.eb, M.n, t);
    for (int k = 0; k < 3; k++) M.gx[k] = 2.0 * t[k];
    v3_cross(M.n, M.ea, t);
    for (int k = 0; k < 3; k++) M.gx[3 + k] = 2.0 * t[k];
    if (M.x >= eps_x) {
        M.m = 1.0;
        M.m1 = M.m2 = 0.0;
    } else {
        M.m = (-M.x / eps_x + 2.0) * M.x / eps_x;
        M.m1 = 2.0 / eps_x - 2.0 * M.x / (eps_x * eps_x);
        M.m2 = -2.0 / (eps_x * eps_x);
    }
}

// slot -> sign of (ea, eb): X0 = -ea, X1 = +ea, X2 = -eb, X3 = +eb
RIGC_HD double moll_grad_vertex(const Mollifier& M, int i, int k) {
    const double s = (i == 0 || i == 2) ? -1.0 : 1.0;
    return s * M.gx[3 * (i >> 1) + k];
}
RIGC_HD double moll_hess_vertex(const Mollifier& M, int i, int k, int j, int l) {
    const double si = (i == 0 || i == 2) ? -1.0 : 1.0;
    const double sj = (j == 0 || j == 2) ? -1.0 : 1.0;
    return si * sj * cross_sq_hess(M.ea, M.eb, M.n, M.ee_a, M.ee_b, M.ee_ab, i >> 1, j >> 1, k, l);
}

// ------------------------------------------------------------ friction
RIGC_HD double friction_f0(double x, double e) { return x >= e ? x : x * x * (-x / 3.0 / e + 1.0) / e + e / 3.0; }
RIGC_HD double friction_f1_over_x(double x, double e) { return x >= e ? 1 / x : (-x / e + 2.0) / e; }
RIGC_HD double friction_f2_term(double x, double e) { return x >= e ? -1 / (x * x * x) : -1 / (x * e * e); }

RIGC_HD void v3_normalize(double* v) {
    const double n = sqrt(v3_dot(v, v));
    for (int k = 0; k < 3; k++) v[k] /= n;
}

// Tangent basis (columns) and closest-point parameters of a constraint at
// X (ipc-toolkit's *_tangent_basis and closest-point functions).
RIGC_HD void friction_frame(int type, const double* X, double coords[2], double basis[3][2]) {
    const double* x0 = X;
    const double* x1 = X + 3;
    const double* x2 = X + 6;
    const double* x3 = X + 9;
    double c0[3], c1[3], t[3], u[3], n[3];
    switch (type) {
    case CON_VV: {
        v3_sub(x1, x0, t);
        const double ex[3] = { 1, 0, 0 }, ey[3] = { 0, 1, 0 };
        double cx[3], cy[3];
        v3_cross(ex, t, cx);
        v3_cross(ey, t, cy);
        const double* cc = v3_dot(cx, cx) > v3_dot(cy, cy) ? cx : cy;
        for (int k = 0; k < 3; k++) c0[k] = cc[k];
        v3_cross(t, cc, c1);
        coords[0] = coords[1] = 0;
        break;
    }
    case CON_EV: {
        v3_sub(x2, x1, t);
        v3_sub(x0, x1, u);
        coords[0] = v3_dot(u, t) / v3_dot(t, t);
        coords[1] = 0;
        for (int k = 0; k < 3; k++) c0[k] = t[k];
        v3_cross(t, u, c1);
        break;
    }
    case CON_EE: {
        double ea[3], eb[3], ba[3];
        v3_sub(x1, x0, ea);
        v3_sub(x3, x2, eb);
        v3_sub(x0, x2, ba);
        solve_2x2_ldlt(v3_dot(ea, ea), -v3_dot(ea, eb), v3_dot(eb, eb), -v3_dot(ba, ea), v3_dot(ba, eb),
                       coords[0], coords[1]);
        v3_cross(ea, eb, n);
        if (v3_dot(n, n) == 0) {
            const double ez[3] = { 0, 0, 1 };
            v3_cross(ea, ez, n);
        }
        for (int k = 0; k < 3; k++) c0[k] = ea[k];
        v3_cross(n, ea, c1);
        break;
    }
    default: {
        double e0[3], e1[3], d[3];
        v3_sub(x2, x1, e0);
        v3_sub(x3, x1, e1);
        v3_sub(x0, x1, d);
        solve_2x2_ldlt(v3_dot(e0, e0), v3_dot(e0, e1), v3_dot(e1, e1), v3_dot(e0, d), v3_dot(e1, d), coords[0],
                       coords[1]);
        v3_cross(e0, e1, n);
        for (int k = 0; k < 3; k++) c0[k] = e0[k];
        v3_cross(n, e0, c1);
        break;
    }
    }
    v3_normalize(c0);
    v3_normalize(c1);
    for (int k = 0; k < 3; k++) {
        basis[k][0] = c0[k];
        basis[k][1] = c1[k];
    }
}

// Relative-displacement weights of the constraint's vertices at the
// closest point given by coords.
RIGC_HD void friction_weights(int type, const double* coords, double w[4]) {
    switch (type) {
    case CON_VV: w[0] = 1; w[1] = -1; w[2] = w[3] = 0; break;
    case CON_EV: w[0] = 1; w[1] = -(1 - coords[0]); w[2] = -coords[0]; w[3] = 0; break;
    case CON_EE:
        w[0] = 1 - coords[0]; w[1] = coords[0]; w[2] = -(1 - coords[1]); w[3] = -coords[1];
        break;
    default: w[0] = 1; w[1] = -(1 - coords[0] - coords[1]); w[2] = -coords[0]; w[3] = -coords[1]; break;
    }
}

// ------------------------------------------------------------ body frames
// cos with the same three-part pi/2 reduction as geom_sin (no shortcut).
RIGC_HD double geom_cos_full(double x) {
    const double k = rint(x * 6.36619772367581382433e-01);
    const double r = ((x - k * 1.57079632673412561417e+00) - k * 6.07710050630396597660e-11)
        - k * 2.02226624879595063154e-21;
    switch (static_cast<long long>(k) & 3) {
    case 0: return geom_cos_poly(r);
    case 1: return -geom_sin_poly(r);
    case 2: return -geom_cos_poly(r);
    default: return geom_sin_poly(r);
    }
}

// sinc'(x) / x  (src/utils/sinc.cpp:103-120)
RIGC_HD double dsinc_over_x(double x) {
    const double x2 = x * x;
    if (fabs(x) > 1e-4) return (x * geom_cos_full(x) - geom_sin(x)) / (x2 * x);
    return x2 * (-x2 / 840.0 + 1.0 / 30.0) - 1.0 / 3.0;
}
// (sinc''(x) / x^2 - sinc'(x) / x^3)  (src/utils/sinc.cpp:122-142)
RIGC_HD double ddsinc_term(double x) {
    const double x2 = x * x, x4 = x2 * x2;
    if (fabs(x) > 0.1) return ((3 - x2) * geom_sin(x) - 3 * x * geom_cos_full(x)) / (x4 * x);
    return x4 / 7560.0 - x2 / 210.0 + 1.0 / 15.0;
}

// Derivatives of R(r) = I + a [r]x + b [r]x^2 (a = sinc|r|,
// b = sinc(|r|/2)^2 / 2) with respect to the rotation vector:
// dR[c] = dR/dr_c, d2R[cd] = d2R/dr_c dr_d for (c,d) in
// (0,0) (0,1) (0,2) (1,1) (1,2) (2,2). Row-major 3x3 each.
struct BodyFrame {
    double dR[3][9];
    double d2R[6][9];
};

RIGC_HD int sym_index(int c, int d) {
    if (c > d) { const int t = c; c = d; d = t; }
    return c == 0 ? d : (c == 1 ? 2 + d : 5);
}

// The scalars and small tensors every entry of dR / d2R is built from.
struct FrameTerms {
    double r[3], a, b, rr, ga[3], gb[3], Ha[3][3], Hb[3][3];
};

RIGC_HD void frame_terms(const double* r, FrameTerms& T) {
    for (int c = 0; c < 3; c++) T.r[c] = r[c];
    const double th = sqrt((r[0] * r[0] + r[1] * r[1]) + r[2] * r[2]);
    const double a = geom_sinc(th), ds = dsinc_over_x(th), dd = ddsinc_term(th);
    const double rh[3] = { r[0] / 2.0, r[1] / 2.0, r[2] / 2.0 };
    const double thh = sqrt((rh[0] * rh[0] + rh[1] * rh[1]) + rh[2] * rh[2]);
    const double h = geom_sinc(thh), dsh = dsinc_over_x(thh), ddh = ddsinc_term(thh);
    double gh[3];
    for (int c = 0; c < 3; c++) {
        T.ga[c] = ds * r[c];
        gh[c] = dsh * rh[c] * 0.5;
    }
    for (int c = 0; c < 3; c++) T.gb[c] = h * gh[c];
    for (int c = 0; c < 3; c++)
        for (int d = 0; d < 3; d++) {
            const double dl = (c == d) ? 1.0 : 0.0;
            T.Ha[c][d] = dd * r[c] * r[d] + ds * dl;
            const double Hh = (ddh * rh[c] * rh[d] + dsh * dl) * 0.25;
            T.Hb[c][d] = gh[c] * gh[d] + h * Hh;
        }
    T.a = a;
    T.b = (0.5 * h) * h;
    T.rr = (r[0] * r[0] + r[1] * r[1]) + r[2] * r[2];
}

RIGC_HD double skew_entry(const double* r, int i, int j) {
    const double K[3][3] = { { 0.0, -r[2], r[1] }, { r[2], 0.0, -r[0] }, { -r[1], r[0], 0.0 } };
    return K[i][j];
}

// dR/dr_c, entry (i, j)
RIGC_HD double frame_dR(const FrameTerms& T, int c, int i, int j) {
    const double* r = T.r;
    const double dl = (i == j) ? 1.0 : 0.0;
    const double K2 = r[i] * r[j] - T.rr * dl;
    const double Ec = v3_skew_unit(c, i, j);
    // E_c K + K E_c = r e_c^T + e_c r^T - 2 r_c I
    const double EK = (j == c ? r[i] : 0.0) + (i == c ? r[j] : 0.0) - 2.0 * r[c] * dl;
    return T.ga[c] * skew_entry(r, i, j) + T.a * Ec + T.gb[c] * K2 + T.b * EK;
}

// d2R/dr_c dr_d, entry (i, j)
RIGC_HD double frame_d2R(const FrameTerms& T, int c, int d, int i, int j) {
    const double* r = T.r;
    const double dcd = (c == d) ? 1.0 : 0.0;
    const double dl = (i == j) ? 1.0 : 0.0;
    const double K2 = r[i] * r[j] - T.rr * dl;
    const double Ec = v3_skew_unit(c, i, j), Ed = v3_skew_unit(d, i, j);
    const double EcK = (j == c ? r[i] : 0.0) + (i == c ? r[j] : 0.0) - 2.0 * r[c] * dl;
    const double EdK = (j == d ? r[i] : 0.0) + (i == d ? r[j] : 0.0) - 2.0 * r[d] * dl;
    // E_c E_d + E_d E_c = e_c e_d^T + e_d e_c^T - 2 delta_cd I
    const double EE = (i == c && j == d ? 1.0 : 0.0) + (i == d && j == c ? 1.0 : 0.0) - 2.0 * dcd * dl;
    return T.Ha[c][d] * skew_entry(r, i, j) + T.ga[c] * Ed + T.ga[d] * Ec + T.Hb[c][d] * K2 + T.gb[c] * EdK +
           T.gb[d] * EcK + T.b * EE;
}

// Entry e of the frame: e < 27 is dR[e / 9][e % 9], else d2R[(e - 27) / 9][(e - 27) % 9]
// (symmetric pair index s = sym_index(c, d)).
RIGC_HD double frame_entry(const FrameTerms& T, int e) {
    if (e < 27) return frame_dR(T, e / 9, (e % 9) / 3, e % 3);
    const int s = (e - 27) / 9, k = (e - 27) % 9;
    // s -> (c, d): 0 (0,0), 1 (0,1), 2 (0,2), 3 (1,1), 4 (1,2), 5 (2,2)
    const int c = s < 3 ? 0 : (s < 5 ? 1 : 2), d = s < 3 ? s : (s < 5 ? s - 2 : 2);
    return frame_d2R(T, c, d, k / 3, k % 3);
}

RIGC_HD void body_frame(const double* r, BodyFrame& F) {
    FrameTerms T;
    frame_terms(r, T);
    for (int c = 0; c < 3; c++)
        for (int k = 0; k < 9; k++) F.dR[c][k] = frame_dR(T, c, k / 3, k % 3);
    for (int c = 0; c < 3; c++)
        for (int d = c; d < 3; d++)
            for (int k = 0; k < 9; k++) F.d2R[sym_index(c, d)][k] = frame_d2R(T, c, d, k / 3, k % 3);
}

}  // namespace rigc

namespace rigc {

// ------------------------------------------------------------ constraint potential
// m * [mollifier(x)] * b(d - dmin^2; 2 dmin dhat + dhat^2) of one
// constraint in its vertices' coordinates (slot i, coordinate k).
struct BarrierLocal {
    DistForm f;
    Mollifier M;
    bool moll;
    double mult, b, b1, b2;
};

RIGC_HD void barrier_local_setup(BarrierLocal& L, int type, const double* X, double eps_x, long long multiplicity,
                                 double dhat, double dmin) {
    constraint_form(type, X, L.f);
    const double dh = 2 * dmin * dhat + dhat * dhat, dm2 = dmin * dmin;
    L.b = ipc_barrier(L.f.d - dm2, dh);
    L.b1 = ipc_barrier_gradient(L.f.d - dm2, dh);
    L.b2 = ipc_barrier_hessian(L.f.d - dm2, dh);
    L.mult = (double)multiplicity;
    L.moll = (type == CON_EE && eps_x >= 0);
    if (L.moll) mollifier_eval(X, eps_x, L.M);
}

RIGC_HD double barrier_local_energy(const BarrierLocal& L) {
    return L.moll ? L.mult * (L.M.m * L.b) : L.mult * L.b;
}

RIGC_HD double barrier_local_grad(const BarrierLocal& L, int i, int k) {
    const double gd = form_grad_vertex(L.f, i, k);
    if (!L.moll) return L.mult * (L.b1 * gd);
    const double gx = moll_grad_vertex(L.M, i, k);
    return L.mult * (L.M.m1 * gx * L.b + L.M.m * L.b1 * gd);
}

RIGC_HD double barrier_local_hess(const BarrierLocal& L, int i, int k, int j, int l) {
    const double gdi = form_grad_vertex(L.f, i, k), gdj = form_grad_vertex(L.f, j, l);
    const double Hd = form_hess_vertex(L.f, i, k, j, l);
    if (!L.moll) return L.mult * (L.b2 * gdi * gdj + L.b1 * Hd);
    const double gxi = moll_grad_vertex(L.M, i, k), gxj = moll_grad_vertex(L.M, j, l);
    const double Hx = moll_hess_vertex(L.M, i, k, j, l);
    const double m = L.M.m, m1 = L.M.m1, m2 = L.M.m2;
    return L.mult * (m2 * L.b * gxi * gxj + m1 * L.b * Hx + m1 * L.b1 * (gxi * gdj + gdi * gxj)
                     + m * L.b2 * gdi * gdj + m * L.b1 * Hd);
}

// ------------------------------------------------------------ friction potential
// m mu lambda f0(|T^T u|; epsv h) with u the displacement of the lagged
// closest point: T (slot i, coord k, column a) = w_i basis[k][a].
struct FrictionLocal {
    double T[12][2];
    double u[2], nu, scale, f1x, M[2][2];
    double energy;
};

// U: per-slot displacement (4 x 3) of the constraint's vertices.
RIGC_HD void friction_local_setup(FrictionLocal& L, int type, int nv, const double* coords, const double basis[3][2],
                                  double normal_force, double mu, long long multiplicity, const double* U,
                                  double epsv_h) {
    double w[4];
    friction_weights(type, coords, w);
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 3; k++)
            for (int a = 0; a < 2; a++) L.T[3 * i + k][a] = i < nv ? w[i] * basis[k][a] : 0.0;
    L.u[0] = L.u[1] = 0.0;
    for (int i = 0; i < nv; i++)
        for (int k = 0; k < 3; k++) {
            L.u[0] += L.T[3 * i + k][0] * U[3 * i + k];
            L.u[1] += L.T[3 * i + k][1] * U[3 * i + k];
        }
    L.nu = sqrt(L.u[0] * L.u[0] + L.u[1] * L.u[1]);
    L.scale = (double)multiplicity * mu * normal_force;
    L.energy = L.scale * friction_f0(L.nu, epsv_h);
    L.f1x = friction_f1_over_x(L.nu, epsv_h);
    L.M[0][0] = L.M[1][1] = L.f1x;
    L.M[0][1] = L.M[1][0] = 0.0;
    if (L.nu > 0) {
        const double f2 = friction_f2_term(L.nu, epsv_h);
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) L.M[a][b] += f2 * L.u[a] * L.u[b];
    }
}

RIGC_HD double friction_local_grad(const FrictionLocal& L, int r) {
    return L.scale * L.f1x * (L.T[r][0] * L.u[0] + L.T[r][1] * L.u[1]);
}

RIGC_HD double friction_local_hess(const FrictionLocal& L, int r, int q) {
    double sum = 0;
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) sum += L.T[r][a] * L.M[a][b] * L.T[q][b];
    return L.scale * sum;
}

}  // namespace rigc
