// gdt_circle.h -- empty-circle test of one triangle against a cell list
// (host/device): the kernels of the circumcircle test and the CPU mode share it.
// Replaces test_circumcircle_kernel_function (DelaunayGPU.cu:192-284).
#pragma once

#include "gdt_point.h"

// Empty-circle test.  The circle through the triangle (circumcentre q, radius
// R, both FP64 relative to the first vertex) is scanned by Chebyshev rings of
// cells around q's cell until the covered square contains the whole circle.
// The test is conservative: a point within relative 1e-9 of the circle counts
// as inside, and a circle that cannot be covered is reported, so no invalid
// triangle escapes (a flagged valid one is merely re-triangulated to itself).

struct GdtTriCircle {
    bool bad;            // not counter-clockwise: invalid whatever else holds
    double qx, qy;       // circumcentre relative to the first vertex
    double R2, lim;      // squared radius and the conservative test bound
};

// Minimum images use the inverse box sides (ilx, ily; general boxes: the
// rounded unit-cell coordinates): neighbours are far closer than half a box,
// so the rounded quotient picks the same image.  The
// centre uses the hardware reciprocal with one Newton step (relative error
// ~1e-15, far inside the test's 1e-9 margin).
GDT_HD inline GdtTriCircle tri_circle(const GdtSet& S, double ilx, double ily,
                                          const gdt_d2* __restrict__ xy, int i, int a, int b)
{
    GdtTriCircle c;
    const gdt_d2 v = xy[i], pa = xy[a], pb = xy[b];
    double ax = pa.x - v.x, ay = pa.y - v.y, bx = pb.x - v.x, by = pb.y - v.y;
    if (S.general) {
        gdt_min_image(S, ax, ay);
        gdt_min_image(S, bx, by);
    } else {
        ax -= S.lx * rint(ax * ilx);
        ay -= S.ly * rint(ay * ily);
        bx -= S.lx * rint(bx * ilx);
        by -= S.ly * rint(by * ily);
    }
    const double D = ax * by - ay * bx;
    c.bad = !(D > 0.);
    c.qx = c.qy = c.R2 = c.lim = 0.;
    if (!c.bad) {
        const double aa = ax * ax + ay * ay, bb = bx * bx + by * by;
        const double i2d = gdt_rcp64(2. * D);
        c.qx = (aa * by - bb * ay) * i2d;
        c.qy = (bb * ax - aa * bx) * i2d;
        c.R2 = c.qx * c.qx + c.qy * c.qy;
        c.lim = c.R2 * (1. + 1e-9);
    }
    return c;
}

// Global-memory scan of the circle of triangle (i, a, b); returns the
// intruder's caller index (-1 none; the vertices, which lie on the circle, are
// told apart by caller index) and whether the circle was covered.
GDT_HD inline int tri_scan_global(const GdtSet& S, const gdt_d2* __restrict__ xy, int i,
                                      const GdtTriCircle& c, int a, int b,
                                      const uint32_t* __restrict__ cstart,
                                      const int* __restrict__ perm,
                                      const gdt_d2* __restrict__ off64, bool& covered)
{
    const gdt_d2 v = xy[i];
    int qcx, qcy;
    double oqx, oqy;
    gdt_bin_point(S, v.x + c.qx, v.y + c.qy, qcx, qcy, oqx, oqy);
    const double covb = gdt_cover0(S, oqx, oqy);
    const double bsmin = S.hmin;
    int intruder = -1;
    covered = false;
    for (int r = 0; r <= S.rmax && intruder < 0; ++r) {
        const int ncell = r ? 8 * r : 1;
        for (int q = 0; q < ncell && intruder < 0; ++q) {
            int dx, dy;
            gdt_ring_cell(r, q, dx, dy);
            const int x = gdt_wrap(qcx + dx, S.xsize), y = gdt_wrap(qcy + dy, S.ysize);
            const int key = S.c0 + gdt_cell_key(S, x, y);
            const uint32_t cb = cstart[key], ce = cstart[key + 1];
            const double ebx = gdt_cbase_x(S, dx, dy, oqx), eby = gdt_cbase_y(S, dx, dy, oqy);
            for (uint32_t k = cb; k < ce; ++k) {
                const gdt_d2 o = off64[k];
                const double ex = gdt_disp64(ebx, o.x), ey = gdt_disp64(eby, o.y);
                if (ex * ex + ey * ey < c.lim) {
                    const int p = perm[k];
                    if (p != i && p != a && p != b) { intruder = p; break; }
                }
            }
        }
        const double cov = (double)r * bsmin + covb;
        if (intruder < 0 && cov > 0. && cov * cov >= c.R2 * (1. + 4e-9)) { covered = true; break; }
    }
    return intruder;
}

GDT_HD inline void tri_flag(int* __restrict__ repair, int i, int a, int b, int intruder_caller)
{
    repair[i] = i;
    repair[a] = a;
    repair[b] = b;
    if (intruder_caller >= 0) repair[intruder_caller] = intruder_caller;
}
