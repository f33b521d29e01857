// gdt_exact.h -- exact signs of the one-ring predicates (host/device), for the
// exact precision mode.
//
// A decision whose FP64 value lies within its error bound of the threshold is
// re-evaluated exactly, on the exact displacements of the caller's points:
//     d = (x_c - x_v) + m L        (m the image shift the FP64 displacement
//                                   chose, recovered by rounding)
// written as floating-point expansions -- sums of doubles, nonoverlapping,
// in increasing magnitude, with zero components removed -- after Shewchuk,
// "Adaptive Precision Floating-Point Arithmetic and Fast Robust Geometric
// Predicates" (Discrete Comput. Geom. 18, 1997): Two-Sum, Two-Product (here
// with the fused multiply-add), Grow-Expansion, Scale-Expansion by repeated
// growth, and Compress.  The sign of an expansion is the sign of its largest
// component.  Every routine is built with FP contraction off, which the
// error-free transformations need.
//
// An expansion holds at most GDT_XN components; an operation whose result
// would not fit sets `overflow` (the caller reports the point as
// non-generic) instead of writing past the end.  Inputs inside a box of a few
// box lengths keep every intermediate within a handful of components.
#pragma once

#include "gdt_point.h"

#define GDT_XN 32

struct GdtX {
    double t[GDT_XN];
    int n;
};

struct GdtXCtx {
    bool overflow;
};

GDT_HD inline void gdt_two_sum(double a, double b, double& x, double& y)
{
    x = a + b;
    const double bv = x - a;
    const double av = x - bv;
    y = (a - av) + (b - bv);
}

GDT_HD inline void gdt_fast_two_sum(double a, double b, double& x, double& y)   // |a| >= |b|
{
    x = a + b;
    y = b - (x - a);
}

GDT_HD inline void gdt_two_prod(double a, double b, double& x, double& y)
{
    x = a * b;
    y = fma(a, b, -x);
}

GDT_HD inline void gdt_x_set(GdtX& e, double v)
{
    e.n = 0;
    if (v != 0.) e.t[e.n++] = v;
}

// e += b (Grow-Expansion, zero components dropped), in place.
GDT_HD inline void gdt_x_grow(GdtX& e, double b, GdtXCtx& cx)
{
    if (b == 0.) return;
    if (e.n + 1 > GDT_XN) {
        cx.overflow = true;
        return;
    }
    double Q = b;
    int k = 0;
    for (int i = 0; i < e.n; ++i) {
        double Qn, h;
        gdt_two_sum(Q, e.t[i], Qn, h);
        Q = Qn;
        if (h != 0.) e.t[k++] = h;      // k <= i: writes behind the read position
    }
    if (Q != 0.) e.t[k++] = Q;
    e.n = k;
}

// Compress (fewer, larger components; the value is unchanged), in place.
GDT_HD inline void gdt_x_compress(GdtX& e)
{
    if (e.n < 2) return;
    double h[GDT_XN];
    int bottom = e.n - 1;
    double Q = e.t[bottom];
    for (int i = e.n - 2; i >= 0; --i) {
        double Qn, q;
        gdt_fast_two_sum(Q, e.t[i], Qn, q);
        if (q != 0.) {
            h[bottom--] = Qn;
            Q = q;
        } else {
            Q = Qn;
        }
    }
    int top = 0;
    for (int i = bottom + 1; i < e.n; ++i) {
        double Qn, q;
        gdt_fast_two_sum(h[i], Q, Qn, q);
        if (q != 0.) e.t[top++] = q;
        Q = Qn;
    }
    if (Q != 0. || top == 0) e.t[top++] = Q;
    e.n = (top == 1 && e.t[0] == 0.) ? 0 : top;
}

// h += s f  (s = +1 or -1).
GDT_HD inline void gdt_x_acc(GdtX& h, const GdtX& f, double s, GdtXCtx& cx)
{
    for (int i = 0; i < f.n && !cx.overflow; ++i) {
        if (h.n + 1 > GDT_XN) gdt_x_compress(h);
        gdt_x_grow(h, s * f.t[i], cx);
    }
    gdt_x_compress(h);
}

// h = e f (the partial products grown in one at a time, compressed whenever
// the expansion nears its capacity).
GDT_HD inline void gdt_x_mul(GdtX& h, const GdtX& e, const GdtX& f, GdtXCtx& cx)
{
    h.n = 0;
    for (int j = 0; j < f.n && !cx.overflow; ++j)
        for (int i = 0; i < e.n && !cx.overflow; ++i) {
            double p, q;
            gdt_two_prod(e.t[i], f.t[j], p, q);
            if (h.n + 2 > GDT_XN) gdt_x_compress(h);
            gdt_x_grow(h, q, cx);
            gdt_x_grow(h, p, cx);
        }
    gdt_x_compress(h);
}

GDT_HD inline int gdt_x_sign(const GdtX& e)
{
    return e.n == 0 ? 0 : (e.t[e.n - 1] > 0. ? 1 : -1);
}

// Exact displacement of caller point c seen from caller point v (rectangle
// lx x ly), the image the FP64 displacement (fx, fy) chose:
// (x_c - x_v) + m lx with m = round((fx - (x_c - x_v)) / lx), likewise y.
GDT_HD inline void gdt_x_disp(const gdt_d2& pv, const gdt_d2& pc, double lx, double ly, double fx,
                              double fy, GdtX& X, GdtX& Y, GdtXCtx& cx)
{
    double s, e;
    gdt_two_sum(pc.x, -pv.x, s, e);
    X.n = 0;
    gdt_x_grow(X, e, cx);
    gdt_x_grow(X, s, cx);
    const double mx = rint((fx - s) / lx);
    if (mx != 0.) {
        double p, q;
        gdt_two_prod(mx, lx, p, q);
        gdt_x_grow(X, q, cx);
        gdt_x_grow(X, p, cx);
    }
    gdt_x_compress(X);
    gdt_two_sum(pc.y, -pv.y, s, e);
    Y.n = 0;
    gdt_x_grow(Y, e, cx);
    gdt_x_grow(Y, s, cx);
    const double my = rint((fy - s) / ly);
    if (my != 0.) {
        double p, q;
        gdt_two_prod(my, ly, p, q);
        gdt_x_grow(Y, q, cx);
        gdt_x_grow(Y, p, cx);
    }
    gdt_x_compress(Y);
}

struct GdtXVec {
    GdtX x, y;
};

// h = p q - r s.
GDT_HD inline void gdt_x_det2(GdtX& h, const GdtX& p, const GdtX& q, const GdtX& r, const GdtX& s,
                              GdtX& tmp, GdtXCtx& cx)
{
    gdt_x_mul(h, p, q, cx);
    gdt_x_mul(tmp, r, s, cx);
    gdt_x_acc(h, tmp, -1., cx);
}

// sign(a x c), the orientation of gdt_orient64.
GDT_HD inline int gdt_x_orient(const GdtXVec& a, const GdtXVec& c, GdtXCtx& cx)
{
    GdtX h, tmp;
    gdt_x_det2(h, a.x, c.y, a.y, c.x, tmp, cx);
    return gdt_x_sign(h);
}

GDT_HD inline void gdt_x_norm2(GdtX& h, const GdtXVec& a, GdtX& tmp, GdtXCtx& cx)
{
    gdt_x_mul(h, a.x, a.x, cx);
    gdt_x_mul(tmp, a.y, a.y, cx);
    gdt_x_acc(h, tmp, 1., cx);
}

// sign(|c|^2 - |b|^2).
GDT_HD inline int gdt_x_dist_cmp(const GdtXVec& c, const GdtXVec& b, GdtXCtx& cx)
{
    GdtX p, q, tmp;
    gdt_x_norm2(p, c, tmp, cx);
    gdt_x_norm2(q, b, tmp, cx);
    gdt_x_acc(p, q, -1., cx);
    return gdt_x_sign(p);
}

// sign of P(c) = K0 |c|^2 + Kx cx + Ky cy for the circle through the origin,
// a and b (the polynomial of gdt_circ64 / gdt_incircle64: < 0 iff c strictly
// inside when a, b are counter-clockwise).
GDT_HD inline int gdt_x_incircle(const GdtXVec& a, const GdtXVec& b, const GdtXVec& c,
                                 GdtXCtx& cx)
{
    GdtX aa, bb, k, t, tmp, acc;
    gdt_x_det2(k, a.x, b.y, a.y, b.x, tmp, cx);      // K0 = ax by - ay bx
    gdt_x_norm2(t, c, tmp, cx);
    gdt_x_mul(acc, k, t, cx);                         // K0 |c|^2
    gdt_x_norm2(aa, a, tmp, cx);
    gdt_x_norm2(bb, b, tmp, cx);
    gdt_x_det2(k, bb, a.y, aa, b.y, tmp, cx);        // Kx = |b|^2 ay - |a|^2 by
    gdt_x_mul(t, k, c.x, cx);
    gdt_x_acc(acc, t, 1., cx);
    gdt_x_det2(k, aa, b.x, bb, a.x, tmp, cx);        // Ky = |a|^2 bx - |b|^2 ax
    gdt_x_mul(t, k, c.y, cx);
    gdt_x_acc(acc, t, 1., cx);
    return gdt_x_sign(acc);
}

// Error bounds of the FP64 path's values against the exact ones, for a box of
// side L with caller coordinates within four box lengths of it (the exact
// mode checks that).  Per displacement coordinate: the wrap of x into the box
// (4uL), the cell offset x - cx bs (about 2uL), the cell base dc bs - o_v and
// the sum (2uL + u|d|), for both points, with a factor 4 of slack:
//     e = 64 u (L + |d|).
// With M the largest displacement coordinate and e the largest error:
//     orientation  |dO| <= 4(e M + e^2) + 4u M^2
//     |c|^2        |dD| <= 4(e M + e^2) + 4u M^2
//     in-circle    |dP| <= 48 e (M + e)^3 + 32 u M^4
// (first-order growth of the degree-2 and degree-4 polynomials in the inputs,
// plus the rounding of their FP64 evaluation), each doubled below.
#define GDT_XU 1.1102230246251565e-16

#ifndef GDT_XTOL_SCALE     // (a host test shrinks it to show that a too-tight bound is caught)
#define GDT_XTOL_SCALE 1.0
#endif
GDT_HD inline double gdt_x_err(double L, double d)
{
    return GDT_XTOL_SCALE * 64. * GDT_XU * (L + fabs(d));
}

GDT_HD inline double gdt_x_tol2(double e, double M)
{
    return 2. * (4. * (e * M + e * e) + GDT_XTOL_SCALE * 4. * GDT_XU * M * M);
}

GDT_HD inline double gdt_x_tol4(double e, double M)
{
    const double Me = M + e;
    return 2. * (48. * e * Me * Me * Me + GDT_XTOL_SCALE * 32. * GDT_XU * M * M * M * M);
}

// a - b, componentwise.
GDT_HD inline void gdt_x_sub(GdtXVec& h, const GdtXVec& a, const GdtXVec& b, GdtXCtx& cx)
{
    h.x = a.x;
    gdt_x_acc(h.x, b.x, -1., cx);
    h.y = a.y;
    gdt_x_acc(h.y, b.y, -1., cx);
}

GDT_HD inline void gdt_x_neg(GdtXVec& h, const GdtXVec& a)
{
    h = a;
    for (int i = 0; i < h.x.n; ++i) h.x.t[i] = -h.x.t[i];
    for (int i = 0; i < h.y.n; ++i) h.y.t[i] = -h.y.t[i];
}

// The symbolic perturbation of gdt_incircle_tie (caller indices pv, pa, pb,
// pc; v at the origin), with the orientations it reads taken exactly: whether
// c, on the circle through v, a, b, counts as inside.
GDT_HD inline bool gdt_x_incircle_tie(int pv, int pa, int pb, int pc, const GdtXVec& a,
                                      const GdtXVec& b, const GdtXVec& c, GdtXCtx& cx)
{
    const int m = gdt_imin(gdt_imin(pv, pa), gdt_imin(pb, pc));
    if (m == pc) return false;
    if (m == pb) return true;
    if (m == pa) return (gdt_x_orient(b, a, cx) > 0) == (gdt_x_orient(b, c, cx) > 0);
    GdtXVec u, na, ca;
    gdt_x_sub(u, b, a, cx);
    gdt_x_neg(na, a);
    gdt_x_sub(ca, c, a, cx);
    return (gdt_x_orient(u, na, cx) > 0) == (gdt_x_orient(u, ca, cx) > 0);
}
