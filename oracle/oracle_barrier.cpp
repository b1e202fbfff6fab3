// ORACLE — test infrastructure only (see oracle_interval.hpp and oracle_barrier.hpp).
#include "oracle_barrier.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <utility>

namespace oracle {

// ------------------------------------------------------------------------
// Second-order forward-mode dual numbers over N variables.
template <int N> struct D2 {
    double v = 0;
    double g[N];
    double h[N][N];
    D2() { std::memset(g, 0, sizeof(g)); std::memset(h, 0, sizeof(h)); }
    explicit D2(double c) : D2() { v = c; }
    static D2 var(double c, int i) { D2 r(c); r.g[i] = 1; return r; }
};
template <int N> D2<N> operator+(const D2<N>& a, const D2<N>& b) {
    D2<N> r; r.v = a.v + b.v;
    for (int i = 0; i < N; i++) { r.g[i] = a.g[i] + b.g[i]; for (int j = 0; j < N; j++) r.h[i][j] = a.h[i][j] + b.h[i][j]; }
    return r;
}
template <int N> D2<N> operator-(const D2<N>& a, const D2<N>& b) {
    D2<N> r; r.v = a.v - b.v;
    for (int i = 0; i < N; i++) { r.g[i] = a.g[i] - b.g[i]; for (int j = 0; j < N; j++) r.h[i][j] = a.h[i][j] - b.h[i][j]; }
    return r;
}
template <int N> D2<N> operator-(const D2<N>& a) { return D2<N>(0.0) - a; }
template <int N> D2<N> operator*(const D2<N>& a, const D2<N>& b) {
    D2<N> r; r.v = a.v * b.v;
    for (int i = 0; i < N; i++) {
        r.g[i] = a.g[i] * b.v + a.v * b.g[i];
        for (int j = 0; j < N; j++)
            r.h[i][j] = a.h[i][j] * b.v + a.v * b.h[i][j] + a.g[i] * b.g[j] + b.g[i] * a.g[j];
    }
    return r;
}
template <int N> D2<N> operator*(double c, const D2<N>& a) {
    D2<N> r; r.v = c * a.v;
    for (int i = 0; i < N; i++) { r.g[i] = c * a.g[i]; for (int j = 0; j < N; j++) r.h[i][j] = c * a.h[i][j]; }
    return r;
}
// f(a) given f, f', f'' at a.v
template <int N> D2<N> chain(const D2<N>& a, double f, double df, double ddf) {
    D2<N> r; r.v = f;
    for (int i = 0; i < N; i++) {
        r.g[i] = df * a.g[i];
        for (int j = 0; j < N; j++) r.h[i][j] = ddf * a.g[i] * a.g[j] + df * a.h[i][j];
    }
    return r;
}
template <int N> D2<N> operator/(const D2<N>& a, const D2<N>& b) {
    const double x = b.v;
    return a * chain(b, 1.0 / x, -1.0 / (x * x), 2.0 / (x * x * x));
}

template <int N> struct V3 { D2<N> c[3]; };
template <int N> V3<N> vsub(const V3<N>& a, const V3<N>& b) { V3<N> r; for (int k = 0; k < 3; k++) r.c[k] = a.c[k] - b.c[k]; return r; }
template <int N> D2<N> vdot(const V3<N>& a, const V3<N>& b) { return (a.c[0] * b.c[0] + a.c[1] * b.c[1]) + a.c[2] * b.c[2]; }
template <int N> V3<N> vcross(const V3<N>& a, const V3<N>& b) {
    V3<N> r;
    r.c[0] = a.c[1] * b.c[2] - a.c[2] * b.c[1];
    r.c[1] = a.c[2] * b.c[0] - a.c[0] * b.c[2];
    r.c[2] = a.c[0] * b.c[1] - a.c[1] * b.c[0];
    return r;
}

using D12 = D2<12>;
using V12 = V3<12>;

// ------------------------------------------------------------------------
// Distances (squared) as formulas of the vertices (ipc-toolkit restated).
template <int N> D2<N> point_point(const V3<N>& p, const V3<N>& q) { V3<N> d = vsub(p, q); return vdot(d, d); }
template <int N> D2<N> point_line(const V3<N>& p, const V3<N>& e0, const V3<N>& e1) {
    V3<N> c = vcross(vsub(e0, p), vsub(e1, p));
    V3<N> e = vsub(e1, e0);
    return vdot(c, c) / vdot(e, e);
}
template <int N> D2<N> point_plane(const V3<N>& p, const V3<N>& t0, const V3<N>& t1, const V3<N>& t2) {
    V3<N> n = vcross(vsub(t1, t0), vsub(t2, t0));
    D2<N> s = vdot(vsub(p, t0), n);
    return (s * s) / vdot(n, n);
}
template <int N> D2<N> line_line(const V3<N>& a0, const V3<N>& a1, const V3<N>& b0, const V3<N>& b1) {
    V3<N> n = vcross(vsub(a1, a0), vsub(b1, b0));
    D2<N> s = vdot(vsub(b0, a0), n);
    return (s * s) / vdot(n, n);
}

namespace {

double dot3(const double* a, const double* b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
void sub3(const double* a, const double* b, double* r) { for (int k = 0; k < 3; k++) r[k] = a[k] - b[k]; }
void cross3(const double* a, const double* b, double* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

enum PEType { P_E0, P_E1, P_E };
enum EEType { EA0_EB0, EA0_EB1, EA1_EB0, EA1_EB1, EA_EB0, EA_EB1, EA0_EB, EA1_EB, EA_EB };
enum PTType { P_T0, P_T1, P_T2, P_TE0, P_TE1, P_TE2, P_T };

PEType point_edge_type(const double* p, const double* e0, const double* e1) {
    double e[3], pe[3];
    sub3(e1, e0, e); sub3(p, e0, pe);
    const double l2 = dot3(e, e);
    if (l2 == 0) return P_E0;
    const double ratio = dot3(e, pe) / l2;
    if (ratio < 0) return P_E0;
    if (ratio > 1) return P_E1;
    return P_E;
}

double point_edge_sq(const double* p, const double* a, const double* b) {
    double ab[3], ap[3], c[3];
    sub3(b, a, ab); sub3(p, a, ap);
    const double l2 = dot3(ab, ab);
    double t = l2 > 0 ? dot3(ap, ab) / l2 : 0;
    t = std::min(1.0, std::max(0.0, t));
    for (int k = 0; k < 3; k++) c[k] = a[k] + t * ab[k] - p[k];
    return dot3(c, c);
}

// Segment-segment closest-feature classification (Sunday / Eberly).
EEType edge_edge_type(const double* ea0, const double* ea1, const double* eb0, const double* eb1) {
    double u[3], v[3], w[3], uv[3];
    sub3(ea1, ea0, u); sub3(eb1, eb0, v); sub3(ea0, eb0, w);
    const double a = dot3(u, u), b = dot3(u, v), c = dot3(v, v), d = dot3(u, w), e = dot3(v, w);
    const double D = a * c - b * b;
    if (a == 0.0 && c == 0.0) return EA0_EB0;
    if (a == 0.0) return EA0_EB;
    if (c == 0.0) return EA_EB0;
    cross3(u, v, uv);
    if (dot3(uv, uv) < 1e-20 * std::max(1.0, a * c)) {
        // parallel: the closest pair involves an endpoint
        const double dd[4] = { point_edge_sq(ea0, eb0, eb1), point_edge_sq(ea1, eb0, eb1),
                               point_edge_sq(eb0, ea0, ea1), point_edge_sq(eb1, ea0, ea1) };
        int k = 0;
        for (int i = 1; i < 4; i++) if (dd[i] < dd[k]) k = i;
        const double* pts[4] = { ea0, ea1, eb0, eb1 };
        if (k < 2) {
            const PEType t = point_edge_type(pts[k], eb0, eb1);
            if (t == P_E) return k == 0 ? EA0_EB : EA1_EB;
            return k == 0 ? (t == P_E0 ? EA0_EB0 : EA0_EB1) : (t == P_E0 ? EA1_EB0 : EA1_EB1);
        }
        const PEType t = point_edge_type(pts[k], ea0, ea1);
        if (t == P_E) return k == 2 ? EA_EB0 : EA_EB1;
        return k == 2 ? (t == P_E0 ? EA0_EB0 : EA1_EB0) : (t == P_E0 ? EA0_EB1 : EA1_EB1);
    }
    EEType def = EA_EB;
    const double sN = b * e - c * d;
    double tN, tD;
    if (sN <= 0.0) { tN = e; tD = c; def = EA0_EB; }
    else if (sN >= D) { tN = e + b; tD = c; def = EA1_EB; }
    else { tN = a * e - b * d; tD = D; }
    if (tN <= 0.0) {
        if (-d <= 0.0) return EA0_EB0;
        if (-d >= a) return EA1_EB0;
        return EA_EB0;
    }
    if (tN >= tD) {
        if ((-d + b) <= 0.0) return EA0_EB1;
        if ((-d + b) >= a) return EA1_EB1;
        return EA_EB1;
    }
    return def;
}

// Symmetric 2x2 solve as Eigen's LDLT (3.3, ldlt().solve) does it: the larger
// diagonal entry is pivoted first, a zero pivot keeps its column
// unscaled, and D is pseudo-inverted (components of zero pivots are 0).
// ipc-toolkit's closest-point coordinates use this solve.
void solve2_ldlt(double a00, double a01, double a11, double b0, double b1, double& x0, double& x1) {
    const bool swap = std::fabs(a11) > std::fabs(a00);
    if (swap) { std::swap(a00, a11); std::swap(b0, b1); }
    const double d0 = a00;
    const double l10 = d0 != 0.0 ? a01 / d0 : a01;
    const double d1 = a11 - l10 * (d0 * l10);
    const double y0 = b0, y1 = b1 - l10 * y0;
    const double tiny = 2.2250738585072014e-308;  // numeric_limits<double>::min()
    const double z0 = std::fabs(d0) > tiny ? y0 / d0 : 0.0;
    const double z1 = std::fabs(d1) > tiny ? y1 / d1 : 0.0;
    const double w1 = z1, w0 = z0 - l10 * w1;
    x0 = swap ? w1 : w0;
    x1 = swap ? w0 : w1;
}

// 2x2 solve of the projection onto (r0, r1).
void solve2(double m00, double m01, double m11, double b0, double b1, double& x0, double& x1) {
    const double det = m00 * m11 - m01 * m01;
    x0 = (m11 * b0 - m01 * b1) / det;
    x1 = (m00 * b1 - m01 * b0) / det;
}

PTType point_triangle_type(const double* p, const double* t0, const double* t1, const double* t2) {
    double e01[3], e02[3], n[3];
    sub3(t1, t0, e01); sub3(t2, t0, e02);
    cross3(e01, e02, n);
    const double* T[3] = { t0, t1, t2 };
    double par[3][2];
    for (int k = 0; k < 3; k++) {
        double r0[3], r1[3], d[3];
        sub3(T[(k + 1) % 3], T[k], r0);
        cross3(r0, n, r1);  // in-plane normal of edge k pointing out of the triangle
        sub3(p, T[k], d);
        solve2(dot3(r0, r0), dot3(r0, r1), dot3(r1, r1), dot3(r0, d), dot3(r1, d), par[k][0], par[k][1]);
        if (par[k][0] > 0.0 && par[k][0] < 1.0 && par[k][1] >= 0.0) return PTType(P_TE0 + k);
    }
    if (par[0][0] <= 0.0 && par[2][0] >= 1.0) return P_T0;
    if (par[1][0] <= 0.0 && par[0][0] >= 1.0) return P_T1;
    if (par[2][0] <= 0.0 && par[1][0] >= 1.0) return P_T2;
    return P_T;
}

// IPC log barrier on squared distances (ipc::barrier and derivatives)
double barrier(double d, double dhat) {
    if (d <= 0.0) return INFINITY;
    if (d >= dhat) return 0.0;
    return -(d - dhat) * (d - dhat) * std::log(d / dhat);
}
double barrier_gradient(double d, double dhat) {
    if (d <= 0.0 || d >= dhat) return 0.0;
    return (dhat - d) * (2 * std::log(d / dhat) - dhat / d + 1);
}
double barrier_hessian(double d, double dhat) {
    if (d <= 0.0 || d >= dhat) return 0.0;
    return (dhat / d + 2) * dhat / d - 2 * std::log(d / dhat) - 3;
}

}  // namespace

// Distance (squared) of a constraint with its type recomputed at V, as a
// dual number over the constraint's (up to 4) vertices (N = 12), or its
// value only (N = 1, no variables seeded).
template <int N>
static D2<N> constraint_distance_p(const Constraint& c, const double* X /*4x3*/, const V3<N>* P);
template <int N>
static D2<N> constraint_distance_t(const Constraint& c, const double* X /*4x3*/) {
    V3<N> P[4];
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 3; k++)
            P[i].c[k] = (N == 12) ? D2<N>::var(X[3 * i + k], (3 * i + k) % N) : D2<N>(X[3 * i + k]);
    return constraint_distance_p<N>(c, X, P);
}
// The same with the vertices given as dual numbers P (types classified at
// their values X).
template <int N>
static D2<N> constraint_distance_p(const Constraint& c, const double* X /*4x3*/, const V3<N>* P) {
    const double* x0 = X;
    const double* x1 = X + 3;
    const double* x2 = X + 6;
    const double* x3 = X + 9;
    switch (c.type) {
    case VV: return point_point(P[0], P[1]);
    case EV: {
        switch (point_edge_type(x0, x1, x2)) {
        case P_E0: return point_point(P[0], P[1]);
        case P_E1: return point_point(P[0], P[2]);
        default: return point_line(P[0], P[1], P[2]);
        }
    }
    case EE: {
        switch (edge_edge_type(x0, x1, x2, x3)) {
        case EA0_EB0: return point_point(P[0], P[2]);
        case EA0_EB1: return point_point(P[0], P[3]);
        case EA1_EB0: return point_point(P[1], P[2]);
        case EA1_EB1: return point_point(P[1], P[3]);
        case EA_EB0: return point_line(P[2], P[0], P[1]);
        case EA_EB1: return point_line(P[3], P[0], P[1]);
        case EA0_EB: return point_line(P[0], P[2], P[3]);
        case EA1_EB: return point_line(P[1], P[2], P[3]);
        default: return line_line(P[0], P[1], P[2], P[3]);
        }
    }
    default: {
        switch (point_triangle_type(x0, x1, x2, x3)) {
        case P_T0: return point_point(P[0], P[1]);
        case P_T1: return point_point(P[0], P[2]);
        case P_T2: return point_point(P[0], P[3]);
        case P_TE0: return point_line(P[0], P[1], P[2]);
        case P_TE1: return point_line(P[0], P[2], P[3]);
        case P_TE2: return point_line(P[0], P[3], P[1]);
        default: return point_plane(P[0], P[1], P[2], P[3]);
        }
    }
    }
}

static D12 constraint_distance(const Constraint& c, const double* X) { return constraint_distance_t<12>(c, X); }
static double constraint_distance_value(const Constraint& c, const double* X) { return constraint_distance_t<1>(c, X).v; }

// Squared distance of an element pair (type EE or FV, vertices packed as in
// a constraint record) with its distance type classified at X.
double element_distance_sq(int type, const double* X) {
    Constraint c;
    c.type = type;
    return constraint_distance_value(c, X);
}

// edge_edge_mollifier(ea0, ea1, eb0, eb1, eps_x) value.
static double mollifier_value(const double* X, double eps_x) {
    double ea[3], eb[3], n[3];
    for (int k = 0; k < 3; k++) { ea[k] = X[3 + k] - X[k]; eb[k] = X[9 + k] - X[6 + k]; }
    n[0] = ea[1] * eb[2] - ea[2] * eb[1];
    n[1] = ea[2] * eb[0] - ea[0] * eb[2];
    n[2] = ea[0] * eb[1] - ea[1] * eb[0];
    const double x = (n[0] * n[0] + n[1] * n[1]) + n[2] * n[2];
    if (x >= eps_x) return 1.0;
    return (-x / eps_x + 2.0) * x / eps_x;
}

// edge_edge_mollifier(ea0, ea1, eb0, eb1, eps_x) as a dual number.
static D12 mollifier(const double* X, double eps_x) {
    V12 P[4];
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 3; k++) P[i].c[k] = D12::var(X[3 * i + k], 3 * i + k);
    V12 n = vcross(vsub(P[1], P[0]), vsub(P[3], P[2]));
    D12 x = vdot(n, n);
    if (x.v >= eps_x) return D12(1.0);
    const double m = (-x.v / eps_x + 2.0) * x.v / eps_x;
    const double dm = 2.0 / eps_x - 2.0 * x.v / (eps_x * eps_x);
    const double ddm = -2.0 / (eps_x * eps_x);
    return chain(x, m, dm, ddm);
}

// ------------------------------------------------------------------------
// Vertex position and derivatives w.r.t. body dofs via duals over the
// rotation vector (world_vertices_diff, rigid_body.tpp:26-78).
namespace {

double cos_rn_full(double x) {
    // cos(x) = sin(x + pi/2) with the same three-part reduction
    const double INV_PIO2 = 6.36619772367581382433e-01;
    const double P1 = 1.57079632673412561417e+00, P2 = 6.07710050630396597660e-11, P3 = 2.02226624879595063154e-21;
    const double k = std::nearbyint(x * INV_PIO2);
    const double r = ((x - k * P1) - k * P2) - k * P3;
    const double z2 = r * r;
    double p = 1.0 / 20922789888000.0;
    p = p * z2 - 1.0 / 87178291200.0; p = p * z2 + 1.0 / 479001600.0; p = p * z2 - 1.0 / 3628800.0;
    p = p * z2 + 1.0 / 40320.0; p = p * z2 - 1.0 / 720.0; p = p * z2 + 1.0 / 24.0; p = p * z2 - 0.5;
    const double c = p * z2 + 1.0;
    double q = 1.0 / 355687428096000.0;
    q = q * z2 - 1.0 / 1307674368000.0; q = q * z2 + 1.0 / 6227020800.0; q = q * z2 - 1.0 / 39916800.0;
    q = q * z2 + 1.0 / 362880.0; q = q * z2 - 1.0 / 5040.0; q = q * z2 + 1.0 / 120.0; q = q * z2 - 1.0 / 6.0;
    const double s = r + (r * z2) * q;
    switch (static_cast<long>(k) & 3) {
    case 0: return c;
    case 1: return -s;
    case 2: return -c;
    default: return s;
    }
}

// sinc.cpp:103-142
double dsinc_over_x(double x) {
    const double x2 = x * x;
    if (std::fabs(x) > 1e-4) return (x * cos_rn_full(x) - sin_rn_full(x)) / (x2 * x);
    return x2 * (-x2 / 840.0 + 1.0 / 30.0) - 1.0 / 3.0;
}
double ddsinc_over_x2_minus_dsinc_over_x3(double x) {
    const double x2 = x * x, x4 = x2 * x2;
    if (std::fabs(x) > 0.1) return ((3 - x2) * sin_rn_full(x) - 3 * x * cos_rn_full(x)) / (x4 * x);
    return x4 / 7560.0 - x2 / 210.0 + 1.0 / 15.0;
}

using D3 = D2<3>;

// sinc_normx for duals (sinc.hpp:64-93)
D3 sinc_normx(const D3 x[3]) {
    const double xv[3] = { x[0].v, x[1].v, x[2].v };
    const double n = std::sqrt((xv[0] * xv[0] + xv[1] * xv[1]) + xv[2] * xv[2]);
    const double val = sinc(n);
    const double ds = dsinc_over_x(n), dd = ddsinc_over_x2_minus_dsinc_over_x3(n);
    double grad[3], hess[3][3];
    for (int k = 0; k < 3; k++) {
        grad[k] = ds * xv[k];
        for (int l = 0; l < 3; l++) hess[k][l] = dd * xv[k] * xv[l] + (k == l ? ds : 0.0);
    }
    D3 r(val);
    for (int k = 0; k < 3; k++)
        for (int i = 0; i < 3; i++) r.g[i] += grad[k] * x[k].g[i];
    for (int k = 0; k < 3; k++) {
        for (int l = 0; l < 3; l++)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) r.h[i][j] += hess[k][l] * x[l].g[i] * x[k].g[j];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) r.h[i][j] += grad[k] * x[k].h[i][j];
    }
    return r;
}

}  // namespace

void vertex_derivatives(const Scene& s, const double* pose, int vertex, double w[3], double jac[18], double hess[27]) {
    point_derivatives(pose, &s.V[3 * vertex], w, jac, hess);
}

void point_derivatives(const double* pose, const double* v, double w[3], double jac[18], double hess[27]) {
    D3 r[3], rh[3];
    for (int k = 0; k < 3; k++) {
        r[k] = D3::var(pose[3 + k], k);
        rh[k] = r[k] / D3(2.0);
    }
    const D3 sa = sinc_normx(r), sh = sinc_normx(rh);
    const D3 z(0.0);
    const D3 K[3][3] = { { z, -r[2], r[1] }, { r[2], z, -r[0] }, { -r[1], r[0], z } };
    const D3 c2 = (0.5 * sh) * sh;
    D3 R[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const D3 k2 = K[i][0] * K[0][j] + (K[i][1] * K[1][j] + K[i][2] * K[2][j]);
            R[i][j] = sa * K[i][j] + c2 * k2;
        }
    for (int i = 0; i < 3; i++) R[i][i] = R[i][i] + D3(1.0);
    for (int j = 0; j < 3; j++) {
        const D3 wj = (v[0] * R[j][0] + v[1] * R[j][1]) + v[2] * R[j][2];
        w[j] = wj.v + pose[j];
        for (int c = 0; c < 6; c++) jac[6 * j + c] = c < 3 ? (c == j ? 1.0 : 0.0) : wj.g[c - 3];
        for (int k = 0; k < 3; k++)
            for (int l = 0; l < 3; l++) hess[9 * j + 3 * k + l] = wj.h[k][l];
    }
}

// ------------------------------------------------------------------------
bool& psd_projection_enabled() {
    static bool on = true;
    return on;
}

void project_to_psd(double* A, int n) {
    if (!psd_projection_enabled()) return;
    // cyclic Jacobi; returns A unchanged when it is already PSD
    std::vector<double> a(A, A + n * n), V(n * n, 0.0);
    for (int i = 0; i < n; i++) V[i * n + i] = 1.0;
    double prev_off = INFINITY;
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0, tot = 0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                tot += a[i * n + j] * a[i * n + j];
                if (i != j) off += a[i * n + j] * a[i * n + j];
            }
        if (off <= 1e-30 * tot || off == 0) break;
        // rank-deficient matrices settle at a rounding floor just above that
        // bound: stop once a sweep no longer cuts the off-diagonal mass by 4
        // and it is already below 1e-24 of the total
        if (off > 0.25 * prev_off && off <= 1e-24 * tot) break;
        prev_off = off;
        for (int p = 0; p < n - 1; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = a[p * n + q];
                if (apq == 0.0) continue;
                const double theta = (a[q * n + q] - a[p * n + p]) / (2 * apq);
                const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1));
                const double c = 1 / std::sqrt(t * t + 1), sn = t * c;
                for (int k = 0; k < n; k++) {
                    const double akp = a[k * n + p], akq = a[k * n + q];
                    a[k * n + p] = c * akp - sn * akq;
                    a[k * n + q] = sn * akp + c * akq;
                }
                for (int k = 0; k < n; k++) {
                    const double apk = a[p * n + k], aqk = a[q * n + k];
                    a[p * n + k] = c * apk - sn * aqk;
                    a[q * n + k] = sn * apk + c * aqk;
                }
                for (int k = 0; k < n; k++) {
                    const double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - sn * vkq;
                    V[k * n + q] = sn * vkp + c * vkq;
                }
            }
    }
    bool psd = true;
    for (int i = 0; i < n; i++) psd = psd && a[i * n + i] >= 0.0;
    if (psd) return;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            double sum = 0;
            for (int k = 0; k < n; k++) sum += V[i * n + k] * std::max(a[k * n + k], 0.0) * V[j * n + k];
            A[i * n + j] = sum;
        }
}

// ------------------------------------------------------------------------
namespace {

void constraint_vertices(const Scene& s, Constraint& c) {
    switch (c.type) {
    case VV: c.nv = 2; c.v[0] = c.a; c.v[1] = c.b; break;
    case EV: c.nv = 3; c.v[0] = c.b; c.v[1] = s.E[2 * c.a]; c.v[2] = s.E[2 * c.a + 1]; break;
    case EE: c.nv = 4; c.v[0] = s.E[2 * c.a]; c.v[1] = s.E[2 * c.a + 1]; c.v[2] = s.E[2 * c.b]; c.v[3] = s.E[2 * c.b + 1]; break;
    default: c.nv = 4; c.v[0] = c.b; c.v[1] = s.F[3 * c.a]; c.v[2] = s.F[3 * c.a + 1]; c.v[3] = s.F[3 * c.a + 2]; break;
    }
    // RigidBody*Constraint::body_ids / vertex_local_body_ids
    // (src/problems/rigid_body_collision_constraint.hpp:12-240)
    c.body[0] = s.v_body[c.v[0]];
    c.body[1] = s.v_body[c.v[c.nv - 1]];
    for (int i = 0; i < c.nv; i++) c.local_body[i] = (s.v_body[c.v[i]] == c.body[0] && i == 0) ? 0 : 1;
    if (c.type == EE) { c.local_body[0] = c.local_body[1] = 0; c.local_body[2] = c.local_body[3] = 1; }
}

void gather(const std::vector<double>& W, const Constraint& c, double* X) {
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 3; k++) X[3 * i + k] = i < c.nv ? W[3 * c.v[i] + k] : 0.0;
}

std::vector<double> all_world_vertices(const Scene& s, const double* poses) {
    std::vector<double> W(3 * (size_t)s.nv), w;
    for (int b = 0; b < s.nb; b++) {
        world_vertices(s, b, poses + 6 * b, w);
        std::copy(w.begin(), w.end(), W.begin() + 3 * (size_t)s.bv[b]);
    }
    return W;
}

// apply_chain_rule (distance_barrier_rb_problem.cpp:1182-1248): vertex-space
// gradient g (3 nv) and Hessian H (3nv x 3nv) -> 12-vector and projected
// 12x12 in the two bodies' dofs.
void chain_rule(const Scene& s, const double* poses, const Constraint& c, const double* g, const double* H,
                double* lg, double* lh) {
    double J[4][18], HV[4][27], w[3];
    for (int i = 0; i < c.nv; i++) vertex_derivatives(s, poses + 6 * s.v_body[c.v[i]], c.v[i], w, J[i], HV[i]);
    for (int k = 0; k < 12; k++) lg[k] = 0;
    for (int i = 0; i < c.nv; i++)
        for (int d = 0; d < 6; d++) {
            double sum = 0;
            for (int j = 0; j < 3; j++) sum += J[i][6 * j + d] * g[3 * i + j];
            lg[6 * c.local_body[i] + d] += sum;
        }
    if (!lh) return;
    const int n = 3 * c.nv;
    std::vector<double> Jf(n * 12, 0.0);
    for (int i = 0; i < c.nv; i++)
        for (int j = 0; j < 3; j++)
            for (int d = 0; d < 6; d++) Jf[(3 * i + j) * 12 + 6 * c.local_body[i] + d] = J[i][6 * j + d];
    std::vector<double> HJ(n * 12, 0.0);
    for (int r = 0; r < n; r++)
        for (int col = 0; col < 12; col++) {
            double sum = 0;
            for (int k = 0; k < n; k++) sum += H[r * n + k] * Jf[k * 12 + col];
            HJ[r * 12 + col] = sum;
        }
    for (int r = 0; r < 12; r++)
        for (int col = 0; col < 12; col++) {
            double sum = 0;
            for (int k = 0; k < n; k++) sum += Jf[k * 12 + r] * HJ[k * 12 + col];
            lh[r * 12 + col] = sum;
        }
    for (int i = 0; i < c.nv; i++)
        for (int j = 0; j < 3; j++) {
            const int o = 6 * c.local_body[i] + 3;
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++) lh[(o + k) * 12 + o + l] += HV[i][9 * j + 3 * k + l] * g[3 * i + j];
        }
    project_to_psd(lh, 12);
}

}  // namespace

void construct_constraint_set(const Scene& s, const double* poses, double dhat, double dmin,
                              std::vector<Constraint>& out) {
    out.clear();
    if (s.nb <= 1) return;
    Candidates cand;
    broadphase_static(s, poses, (dhat + dmin) / 2.0, cand);
    const std::vector<double> W = all_world_vertices(s, poses);
    const double offset_sqr = (dmin + dhat) * (dmin + dhat);
    std::vector<Constraint> vv, ev, ee, fv;
    std::map<std::pair<int, int>, size_t> vv_idx, ev_idx;
    auto add_vv = [&](int a, int b) {
        const std::pair<int, int> k(std::min(a, b), std::max(a, b));
        auto it = vv_idx.find(k);
        if (it != vv_idx.end()) { vv[it->second].multiplicity++; return; }
        Constraint c; c.type = VV; c.a = k.first; c.b = k.second;
        vv_idx[k] = vv.size();
        vv.push_back(c);
    };
    auto add_ev = [&](int e, int v) {
        const std::pair<int, int> k(e, v);
        auto it = ev_idx.find(k);
        if (it != ev_idx.end()) { ev[it->second].multiplicity++; return; }
        Constraint c; c.type = EV; c.a = e; c.b = v;
        ev_idx[k] = ev.size();
        ev.push_back(c);
    };
    for (const auto& p : cand.ee) {
        const int ei = p[0], ej = p[1];
        const int i0 = s.E[2 * ei], i1 = s.E[2 * ei + 1], j0 = s.E[2 * ej], j1 = s.E[2 * ej + 1];
        const double *a0 = &W[3 * i0], *a1 = &W[3 * i1], *b0 = &W[3 * j0], *b1 = &W[3 * j1];
        EEType t = edge_edge_type(a0, a1, b0, b1);
        Constraint probe; probe.type = EE; probe.a = ei; probe.b = ej;
        constraint_vertices(s, probe);
        double X[12];
        gather(W, probe, X);
        const double d = constraint_distance_value(probe, X);
        if (!(d < offset_sqr)) continue;
        // eps_x from V_rest = V (distance_barrier_constraint.cpp:297-299)
        double ua[3], ub[3], n[3];
        sub3(a1, a0, ua); sub3(b1, b0, ub); cross3(ua, ub, n);
        const double eps_x = 1e-3 * dot3(ua, ua) * dot3(ub, ub);
        if (dot3(n, n) < eps_x) t = EA_EB;
        switch (t) {
        case EA0_EB0: add_vv(i0, j0); break;
        case EA0_EB1: add_vv(i0, j1); break;
        case EA1_EB0: add_vv(i1, j0); break;
        case EA1_EB1: add_vv(i1, j1); break;
        case EA_EB0: add_ev(ei, j0); break;
        case EA_EB1: add_ev(ei, j1); break;
        case EA0_EB: add_ev(ej, i0); break;
        case EA1_EB: add_ev(ej, i1); break;
        default: probe.eps_x = eps_x; ee.push_back(probe); break;
        }
    }
    for (const auto& p : cand.fv) {
        const int f = p[0], v = p[1];
        const int f0 = s.F[3 * f], f1 = s.F[3 * f + 1], f2 = s.F[3 * f + 2];
        Constraint probe; probe.type = FV; probe.a = f; probe.b = v;
        constraint_vertices(s, probe);
        double X[12];
        gather(W, probe, X);
        const double d = constraint_distance_value(probe, X);
        if (!(d < offset_sqr)) continue;
        switch (point_triangle_type(&W[3 * v], &W[3 * f0], &W[3 * f1], &W[3 * f2])) {
        case P_T0: add_vv(v, f0); break;
        case P_T1: add_vv(v, f1); break;
        case P_T2: add_vv(v, f2); break;
        case P_TE0: add_ev(s.FE[3 * f], v); break;
        case P_TE1: add_ev(s.FE[3 * f + 1], v); break;
        case P_TE2: add_ev(s.FE[3 * f + 2], v); break;
        default: fv.push_back(probe); break;
        }
    }
    for (auto& c : vv) constraint_vertices(s, c);
    for (auto& c : ev) constraint_vertices(s, c);
    out.insert(out.end(), vv.begin(), vv.end());
    out.insert(out.end(), ev.begin(), ev.end());
    out.insert(out.end(), ee.begin(), ee.end());
    out.insert(out.end(), fv.begin(), fv.end());
}

double barrier_term(const Scene& s, const double* poses, const std::vector<Constraint>& C, double dhat, double dmin,
                    double* grad, std::vector<std::array<double, 144>>* local_hess) {
    const std::vector<double> W = all_world_vertices(s, poses);
    const double dh = 2 * dmin * dhat + dhat * dhat, dm2 = dmin * dmin;
    const size_t n = C.size();
    std::vector<double> energy(n);
    if (!grad && !local_hess) {
        // value only (compute_potential)
#pragma omp parallel for schedule(dynamic, 64)
        for (long ci = 0; ci < (long)n; ci++) {
            const Constraint& c = C[ci];
            double X[12];
            gather(W, c, X);
            const double d = constraint_distance_value(c, X);
            double b = barrier(d - dm2, dh);
            if (c.type == EE && c.eps_x >= 0) b = mollifier_value(X, c.eps_x) * b;
            energy[ci] = (double)c.multiplicity * b;
        }
        double e = 0;
        for (size_t ci = 0; ci < n; ci++) e += energy[ci];
        return e;
    }
    std::vector<std::array<double, 12>> lgrad(n);
    if (local_hess) local_hess->assign(n, std::array<double, 144>{});
#pragma omp parallel for schedule(dynamic, 8)
    for (long ci = 0; ci < (long)n; ci++) {
        const Constraint& c = C[ci];
        double X[12];
        gather(W, c, X);
        const D12 d = constraint_distance(c, X);
        const double mult = (double)c.multiplicity;
        const double b = barrier(d.v - dm2, dh), db = barrier_gradient(d.v - dm2, dh), ddb = barrier_hessian(d.v - dm2, dh);
        D12 B = chain(d, b, db, ddb);
        if (c.type == EE && c.eps_x >= 0) B = mollifier(X, c.eps_x) * B;
        B = mult * B;
        energy[ci] = B.v;
        const int nn = 3 * c.nv;
        double g[12], H[144];
        for (int i = 0; i < nn; i++) {
            g[i] = B.g[i];
            for (int j = 0; j < nn; j++) H[i * nn + j] = B.h[i][j];
        }
        chain_rule(s, poses, c, g, H, lgrad[ci].data(), local_hess ? (*local_hess)[ci].data() : nullptr);
    }
    double e = 0;
    for (size_t ci = 0; ci < n; ci++) {
        e += energy[ci];
        if (grad)
            for (int k = 0; k < 12; k++) grad[6 * C[ci].body[k / 6] + k % 6] += lgrad[ci][k];
    }
    return e;
}

// One constraint's potential in its vertices' coordinates (X: 4 slots x 3),
// unprojected: used to check closed-form derivatives.
double constraint_local(int type, const double* X, double eps_x, double dhat, double dmin, long multiplicity,
                        double* g, double* H) {
    Constraint c;
    c.type = type;
    c.eps_x = eps_x;
    c.multiplicity = multiplicity;
    const D12 d = constraint_distance(c, X);
    const double dh = 2 * dmin * dhat + dhat * dhat, dm2 = dmin * dmin;
    D12 B = chain(d, barrier(d.v - dm2, dh), barrier_gradient(d.v - dm2, dh), barrier_hessian(d.v - dm2, dh));
    if (c.type == EE && c.eps_x >= 0) B = mollifier(X, c.eps_x) * B;
    B = (double)c.multiplicity * B;
    for (int i = 0; i < 12; i++) {
        g[i] = B.g[i];
        for (int j = 0; j < 12; j++) H[12 * i + j] = B.h[i][j];
    }
    return B.v;
}

double minimum_distance(const Scene& s, const double* poses, const std::vector<Constraint>& C) {
    const std::vector<double> W = all_world_vertices(s, poses);
    double m = INFINITY;
    for (const auto& c : C) {
        double X[12];
        gather(W, c, X);
        m = std::min(m, constraint_distance_value(c, X));
    }
    return m;
}

// ---------------------------------------------------------------- friction
namespace {

void normalize3(double* v) {
    const double n = std::sqrt(dot3(v, v));
    for (int k = 0; k < 3; k++) v[k] /= n;
}

// ipc tangent bases (columns) and closest-point parameters; weights w[i]
// give the relative displacement sum_i w_i U_i of the constraint's vertices.
void friction_frame(const Constraint& c, const double* X, double coords[2], double basis[3][2], double w[4]) {
    const double *x0 = X, *x1 = X + 3, *x2 = X + 6, *x3 = X + 9;
    double c0[3], c1[3], t[3], u[3], n[3];
    switch (c.type) {
    case VV: {
        sub3(x1, x0, t);
        const double ex[3] = { 1, 0, 0 }, ey[3] = { 0, 1, 0 };
        double cx[3], cy[3];
        cross3(ex, t, cx); cross3(ey, t, cy);
        const double* cc = dot3(cx, cx) > dot3(cy, cy) ? cx : cy;
        for (int k = 0; k < 3; k++) c0[k] = cc[k];
        cross3(t, cc, c1);
        coords[0] = coords[1] = 0;
        w[0] = 1; w[1] = -1; w[2] = w[3] = 0;
        break;
    }
    case EV: {
        sub3(x2, x1, t); sub3(x0, x1, u);
        const double a = dot3(u, t) / dot3(t, t);
        coords[0] = a; coords[1] = 0;
        for (int k = 0; k < 3; k++) c0[k] = t[k];
        cross3(t, u, c1);
        w[0] = 1; w[1] = -(1 - a); w[2] = -a; w[3] = 0;
        break;
    }
    case EE: {
        double ea[3], eb[3], ba[3];
        sub3(x1, x0, ea); sub3(x3, x2, eb); sub3(x0, x2, ba);
        double al, be;
        solve2_ldlt(dot3(ea, ea), -dot3(ea, eb), dot3(eb, eb), -dot3(ba, ea), dot3(ba, eb), al, be);
        coords[0] = al; coords[1] = be;
        cross3(ea, eb, n);
        if (dot3(n, n) == 0) { const double ez[3] = { 0, 0, 1 }; cross3(ea, ez, n); }
        for (int k = 0; k < 3; k++) c0[k] = ea[k];
        cross3(n, ea, c1);
        w[0] = 1 - al; w[1] = al; w[2] = -(1 - be); w[3] = -be;
        break;
    }
    default: {
        double e0[3], e1[3], d[3];
        sub3(x2, x1, e0); sub3(x3, x1, e1); sub3(x0, x1, d);
        double b1, b2;
        solve2_ldlt(dot3(e0, e0), dot3(e0, e1), dot3(e1, e1), dot3(e0, d), dot3(e1, d), b1, b2);
        coords[0] = b1; coords[1] = b2;
        cross3(e0, e1, n);
        for (int k = 0; k < 3; k++) c0[k] = e0[k];
        cross3(n, e0, c1);
        w[0] = 1; w[1] = -(1 - b1 - b2); w[2] = -b1; w[3] = -b2;
        break;
    }
    }
    normalize3(c0);
    normalize3(c1);
    for (int k = 0; k < 3; k++) { basis[k][0] = c0[k]; basis[k][1] = c1[k]; }
}

double f0_SF(double x, double e) { return x >= e ? x : x * x * (-x / 3.0 / e + 1.0) / e + e / 3.0; }
double f1_SF_over_x(double x, double e) { return x >= e ? 1 / x : (-x / e + 2.0) / e; }
double df1_x_minus_f1_over_x3(double x, double e) { return x >= e ? -1 / (x * x * x) : -1 / (x * e * e); }

}  // namespace

void construct_friction_constraint_set(const Scene& s, const double* poses, const std::vector<Constraint>& C,
                                       double dhat, double dmin, double kappa, double mu,
                                       std::vector<FrictionConstraint>& out) {
    out.clear();
    const std::vector<double> W = all_world_vertices(s, poses);
    const double dh = 2 * dmin * dhat + dhat * dhat, dm2 = dmin * dmin;
    for (const auto& c : C) {
        FrictionConstraint f;
        f.c = c;
        double X[12];
        gather(W, c, X);
        const double d = constraint_distance_value(c, X);
        double w[4];
        friction_frame(c, X, f.coords, f.basis, w);
        // compute_normal_force_magnitude: -kappa * b'(d) * 2 sqrt(d)
        f.normal_force = -kappa * (barrier_gradient(d - dm2, dh) * (2 * std::sqrt(d)));
        f.mu = mu;
        out.push_back(f);
    }
}

// Adjoint of the friction gradient with respect to the start-of-step poses
// (the device's rigc_friction_adjoint), by second-order dual numbers over
// the 12 coordinates of a record's vertices at x (variables 0-11) and at x0
// (12-23): out += J(x0)^T (d2 D / dW0 dW) J(x) lambda, record by record. D
// depends on x0 through the displacement and through the normal force
// -kappa b'(d(x0)) 2 sqrt(d(x0)); the tangent basis and weights are held
// fixed. Only the mixed block of the dual Hessian is used, so the normal
// force's own second derivative is not seeded.
void friction_adjoint(const Scene& s, const double* poses, const double* poses_t0,
                      const std::vector<FrictionConstraint>& FC, double eh, double dhat, double dmin, double kappa,
                      const double* lambda, double* out) {
    using D24 = D2<24>;
    using V24 = V3<24>;
    const std::vector<double> W1 = all_world_vertices(s, poses), W0 = all_world_vertices(s, poses_t0);
    for (int i = 0; i < 6 * s.nb; i++) out[i] = 0;
    const double dh = 2 * dmin * dhat + dhat * dhat, dm2 = dmin * dmin;
    for (const FrictionConstraint& f : FC) {
        const Constraint& c = f.c;
        double X1[12], X0[12];
        gather(W1, c, X1);
        gather(W0, c, X0);
        V24 P1[4], P0[4];
        for (int i = 0; i < 4; i++)
            for (int k = 0; k < 3; k++) {
                P1[i].c[k] = i < c.nv ? D24::var(X1[3 * i + k], 3 * i + k) : D24(0.0);
                P0[i].c[k] = i < c.nv ? D24::var(X0[3 * i + k], 12 + 3 * i + k) : D24(0.0);
            }
        double w[4];
        switch (c.type) {
        case VV: w[0] = 1; w[1] = -1; w[2] = w[3] = 0; break;
        case EV: w[0] = 1; w[1] = -(1 - f.coords[0]); w[2] = -f.coords[0]; w[3] = 0; break;
        case EE: w[0] = 1 - f.coords[0]; w[1] = f.coords[0]; w[2] = -(1 - f.coords[1]); w[3] = -f.coords[1]; break;
        default: w[0] = 1; w[1] = -(1 - f.coords[0] - f.coords[1]); w[2] = -f.coords[0]; w[3] = -f.coords[1]; break;
        }
        D24 ua[2] = { D24(0.0), D24(0.0) };
        for (int a = 0; a < 2; a++)
            for (int i = 0; i < c.nv; i++)
                for (int k = 0; k < 3; k++) ua[a] = ua[a] + (w[i] * f.basis[k][a]) * (P1[i].c[k] - P0[i].c[k]);
        const D24 s2 = ua[0] * ua[0] + ua[1] * ua[1];
        D24 F0;
        if (s2.v > 0) {
            const double sq = std::sqrt(s2.v);
            const D24 nu = chain(s2, sq, 0.5 / sq, -0.25 / (s2.v * sq));
            const double x = nu.v;
            const double f0 = x >= eh ? x : x * x * (-x / 3.0 / eh + 1.0) / eh + eh / 3.0;
            const double f1 = x >= eh ? 1.0 : -x * x / (eh * eh) + 2.0 * x / eh;
            const double f1d = x >= eh ? 0.0 : -2.0 * x / (eh * eh) + 2.0 / eh;
            F0 = chain(nu, f0, f1, f1d);
        } else {
            // no lagged displacement: f0 = s / eh + eh / 3 - s^(3/2) / (3 eh^2)
            // in s = |T^T u|^2, whose last term has no first or second
            // derivative at s = 0 (the gradient vanishes, the Hessian
            // 2 T T^T / eh does not)
            F0 = chain(s2, eh / 3.0, 1.0 / eh, 0.0);
        }
        const D24 d0 = constraint_distance_p<24>(c, X0, P0);
        const double sd = std::sqrt(d0.v);
        const double b1 = barrier_gradient(d0.v - dm2, dh), b2 = barrier_hessian(d0.v - dm2, dh);
        const D24 lamN = chain(d0, -kappa * (b1 * (2 * sd)), -kappa * (b2 * (2 * sd) + b1 / sd), 0.0);
        const D24 Dk = ((double)c.multiplicity * f.mu) * (lamN * F0);
        // y = J(x) lambda on the record's vertex slots
        double y[12] = { 0 }, J1[4][18], J0[4][18], Hv[27], wv[3];
        for (int i = 0; i < c.nv; i++) {
            const int b = s.v_body[c.v[i]];
            vertex_derivatives(s, poses + 6 * b, c.v[i], wv, J1[i], Hv);
            vertex_derivatives(s, poses_t0 + 6 * b, c.v[i], wv, J0[i], Hv);
            for (int k = 0; k < 3; k++)
                for (int d = 0; d < 6; d++) y[3 * i + k] += J1[i][6 * k + d] * lambda[6 * b + d];
        }
        for (int i = 0; i < c.nv; i++) {
            const int b = s.v_body[c.v[i]];
            double z[3];
            for (int k = 0; k < 3; k++) {
                double sum = 0;
                for (int r = 0; r < 12; r++) sum += Dk.h[12 + 3 * i + k][r] * y[r];
                z[k] = sum;
            }
            for (int d = 0; d < 6; d++) {
                double sum = 0;
                for (int k = 0; k < 3; k++) sum += J0[i][6 * k + d] * z[k];
                out[6 * b + d] += sum;
            }
        }
    }
}

double friction_term(const Scene& s, const double* poses, const double* poses_t0,
                     const std::vector<FrictionConstraint>& FC, double eh, double* grad,
                     std::vector<std::array<double, 144>>* local_hess) {
    const std::vector<double> W1 = all_world_vertices(s, poses), W0 = all_world_vertices(s, poses_t0);
    const size_t n = FC.size();
    std::vector<double> energy(n);
    const bool value_only = !grad && !local_hess;
    std::vector<std::array<double, 12>> lgrad(n);
    if (local_hess) local_hess->assign(n, std::array<double, 144>{});
#pragma omp parallel for schedule(dynamic, 8)
    for (long ci = 0; ci < (long)n; ci++) {
        const FrictionConstraint& f = FC[ci];
        const Constraint& c = f.c;
        double X0[12], coords[2], basis[3][2], w[4];
        gather(W0, c, X0);
        // weights from the lagged closest point
        switch (c.type) {
        case VV: w[0] = 1; w[1] = -1; w[2] = w[3] = 0; break;
        case EV: w[0] = 1; w[1] = -(1 - f.coords[0]); w[2] = -f.coords[0]; w[3] = 0; break;
        case EE: w[0] = 1 - f.coords[0]; w[1] = f.coords[0]; w[2] = -(1 - f.coords[1]); w[3] = -f.coords[1]; break;
        default: w[0] = 1; w[1] = -(1 - f.coords[0] - f.coords[1]); w[2] = -f.coords[0]; w[3] = -f.coords[1]; break;
        }
        (void)coords; (void)basis;
        const int nn = 3 * c.nv;
        double T[12][2];
        for (int i = 0; i < c.nv; i++)
            for (int k = 0; k < 3; k++)
                for (int j = 0; j < 2; j++) T[3 * i + k][j] = w[i] * f.basis[k][j];
        double u[2] = { 0, 0 };
        for (int i = 0; i < c.nv; i++)
            for (int k = 0; k < 3; k++) {
                const double U = W1[3 * c.v[i] + k] - W0[3 * c.v[i] + k];
                u[0] += T[3 * i + k][0] * U;
                u[1] += T[3 * i + k][1] * U;
            }
        const double nu = std::sqrt(u[0] * u[0] + u[1] * u[1]);
        const double scale = (double)c.multiplicity * f.mu * f.normal_force;
        energy[ci] = scale * f0_SF(nu, eh);
        if (value_only) continue;
        const double f1x = f1_SF_over_x(nu, eh);
        double g[12], H[144];
        for (int r = 0; r < nn; r++) g[r] = scale * f1x * (T[r][0] * u[0] + T[r][1] * u[1]);
        double M[2][2] = { { f1x, 0 }, { 0, f1x } };
        if (nu > 0) {
            const double f2 = df1_x_minus_f1_over_x3(nu, eh);
            for (int a = 0; a < 2; a++)
                for (int b = 0; b < 2; b++) M[a][b] += f2 * u[a] * u[b];
        }
        for (int r = 0; r < nn; r++)
            for (int q = 0; q < nn; q++) {
                double sum = 0;
                for (int a = 0; a < 2; a++)
                    for (int b = 0; b < 2; b++) sum += T[r][a] * M[a][b] * T[q][b];
                H[r * nn + q] = scale * sum;
            }
        chain_rule(s, poses, c, g, H, lgrad[ci].data(), local_hess ? (*local_hess)[ci].data() : nullptr);
    }
    double e = 0;
    for (size_t ci = 0; ci < n; ci++) {
        e += energy[ci];
        if (grad && !value_only)
            for (int k = 0; k < 12; k++) grad[6 * FC[ci].c.body[k / 6] + k % 6] += lgrad[ci][k];
    }
    return e;
}

}  // namespace oracle
