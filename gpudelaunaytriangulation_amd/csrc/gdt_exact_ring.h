// gdt_exact_ring.h -- one point's ring with exact decisions (host/device): the
// exact precision mode.
//
// The search of gdt_one_ring_f64 (nearest neighbour, then gift wrapping over
// Chebyshev rings of cells with certified stops), with every decision taken
// on the exact displacements of the caller's coordinates: a comparison whose
// FP64 value lies outside its error bound (gdt_exact.h) keeps the FP64
// answer, which is then the exact one; inside it the sign is computed with
// expansions, and an exact zero falls to the symbolic perturbation of
// gdt_nn_tie / gdt_incircle_tie (with exact orientations), so the rings are
// the Delaunay triangulation of the caller's points perturbed as the FP64
// paths perturb exact ties.  Rectangular boxes; caller coordinates within four
// box lengths of the box (ERR_RANGE otherwise); an expansion that outgrows its
// capacity reports ERR_NONEIGHBOUR (non-generic beyond this mode's reach).
#pragma once

#include "gdt_exact.h"
#include "gdt_point.h"

// Audit hook of host test builds: every decision taken from its FP64 value
// alone (kind 0 orientation, 1 in-circle, 2 distance), the sign it claims and
// the exact sign of the same predicate.
#ifdef GDT_AUDIT_EXACT
void gdt_exact_audit(int kind, int claimed, int exact);
#define GDT_XAUDIT(kind, claimed, exact_sign) gdt_exact_audit(kind, claimed, exact_sign)
#else
#define GDT_XAUDIT(kind, claimed, exact_sign) ((void)0)
#endif

// A candidate: slot, FP64 displacement, and the pool entry that holds its
// exact displacement once needed.  Taking another candidate's role swaps the
// pool entries instead of copying expansions.
struct GdtXCand {
    int k;
    double x, y;
    int xs;         // pool entry owned by this role
    bool have;      // entry xs holds the exact displacement of slot k
};

GDT_HD inline void gdt_xc_set(GdtXCand& c, int k, double x, double y)
{
    c.k = k;
    c.x = x;
    c.y = y;
    c.have = false;
}

// dst takes src's candidate (src's role is reset by its next set)
GDT_HD inline void gdt_xc_take(GdtXCand& dst, GdtXCand& src)
{
    const int t = dst.xs;
    dst.k = src.k;
    dst.x = src.x;
    dst.y = src.y;
    dst.have = src.have;
    dst.xs = src.xs;
    src.xs = t;
    src.have = false;
}

GDT_HD inline GdtRingResult gdt_one_ring_exact(const GdtSet& S, int j, const GdtRingIO& io)
{
    GdtRingResult res = {0, 0, 0u, 0u};
    const uint32_t pc = io.scell[j];
    const int cx = (int)(pc & 0xffffu), cy = (int)(pc >> 16);
    const gdt_d2 ov = io.off64[j];
    const double covb = gdt_cover0(S, ov.x, ov.y);
    const double bsmin = S.hmin;
    const size_t row = (size_t)io.perm[j] * (size_t)io.stride;
    const double L = S.lx > S.ly ? S.lx : S.ly;
    const gdt_d2 pv = io.xy[io.perm[j]];
    if (S.general || !(fabs(pv.x) <= 4. * S.lx && fabs(pv.y) <= 4. * S.ly)) {
        res.err = ERR_RANGE;
        return res;
    }
    unsigned npred = 0;
    GdtXCtx xc;
    xc.overflow = false;
    GdtXVec pool[4];
    auto exact = [&](GdtXCand& c) -> const GdtXVec& {
        GdtXVec& X = pool[c.xs];
        if (!c.have) {
            gdt_x_disp(pv, io.xy[io.perm[c.k]], S.lx, S.ly, c.x, c.y, X.x, X.y, xc);
            c.have = true;
        }
        return X;
    };
    auto mag = [](const GdtXCand& c) { return fmax(fabs(c.x), fabs(c.y)); };

    // 1. nearest neighbour: c beats b when |c|^2 < |b|^2, exact ties by caller index
    GdtXCand nn, cand, a, best;
    nn.xs = 0;
    cand.xs = 1;
    a.xs = 2;
    best.xs = 3;
    nn.k = -1;
    nn.have = cand.have = a.have = best.have = false;
    double nd2 = 0.;
    auto nn_cell = [&](int dx, int dy, int, uint32_t b, uint32_t e) {
        const double bx = gdt_cbase_x(S, dx, dy, ov.x), by = gdt_cbase_y(S, dx, dy, ov.y);
        for (uint32_t k = b; k < e; ++k) {
            if ((int)k == j) continue;
            const gdt_d2 o = io.off64[k];
            const double px = gdt_disp64(bx, o.x), py = gdt_disp64(by, o.y);
            const double d2 = gdt_dist2_64(px, py);
            ++npred;
            gdt_xc_set(cand, (int)k, px, py);
            bool take = nn.k < 0;
            if (!take) {
                const double mc = mag(cand), mb = mag(nn);
                const double tol = gdt_x_tol2(gdt_x_err(L, mc), mc) + gdt_x_tol2(gdt_x_err(L, mb), mb);
                if (d2 < nd2 - tol) {
                    take = true;
                    GDT_XAUDIT(2, -1, gdt_x_dist_cmp(exact(cand), exact(nn), xc));
                } else if (d2 > nd2 + tol) {
                    GDT_XAUDIT(2, 1, gdt_x_dist_cmp(exact(cand), exact(nn), xc));
                } else {
                    const int s = gdt_x_dist_cmp(exact(cand), exact(nn), xc);
                    take = s < 0 || (s == 0 && gdt_nn_tie(io.perm[k], io.perm[nn.k]));
                }
            }
            if (take) {
                gdt_xc_take(nn, cand);
                nd2 = d2;
            }
        }
    };
    // certified when the nearest distance, widened by the error of every
    // position, stays inside the covered square
    auto nn_done = [&](int r) {
        if (nn.k < 0) return false;
        const double cov = (double)r * bsmin + covb - 4. * gdt_x_err(L, (double)(r + 1) * bsmin);
        return cov > 0. && nd2 <= cov * cov * (1.0 - 1e-9);
    };
    if (gdt_scan_rings(S, cx, cy, io.cstart, nn_cell, nn_done) < 0) {
        res.err = nn.k < 0 ? ERR_NONEIGHBOUR : ERR_SEARCH;
        res.npred = npred;
        return res;
    }

    // 2. walk the ring counter-clockwise
    gdt_xc_take(a, nn);
    nn.k = a.k;         // (only its slot is read from here on)
    int count = 0;
    for (;;) {
        if (count < io.stride) io.nbrs[row + count] = io.perm[a.k];
        ++count;
        if (count > GDT_HARD_CAP) { res.err |= ERR_NOCLOSE; break; }
        best.k = -1;
        Circ64 K = {0., 0., 0.};
        auto sc_cell = [&](int dx, int dy, int, uint32_t b, uint32_t e) {
            const double bx = gdt_cbase_x(S, dx, dy, ov.x), by = gdt_cbase_y(S, dx, dy, ov.y);
            for (uint32_t k = b; k < e; ++k) {
                if ((int)k == j || (int)k == a.k) continue;
                const gdt_d2 o = io.off64[k];
                const double px = gdt_disp64(bx, o.x), py = gdt_disp64(by, o.y);
                ++npred;
                gdt_xc_set(cand, (int)k, px, py);
                // strictly left of v->a
                {
                    const double O = gdt_orient64(a.x, a.y, px, py);
                    const double M = fmax(mag(a), mag(cand));
                    const double tol = gdt_x_tol2(gdt_x_err(L, M), M);
                    bool left = O > tol;
                    if (left) GDT_XAUDIT(0, 1, gdt_x_orient(exact(a), exact(cand), xc));
                    else if (O < -tol) GDT_XAUDIT(0, -1, gdt_x_orient(exact(a), exact(cand), xc));
                    else left = gdt_x_orient(exact(a), exact(cand), xc) > 0;
                    if (!left) continue;
                }
                if (best.k >= 0) {
                    ++npred;
                    const double P = gdt_incircle64(K, px, py);
                    const double M = fmax(fmax(mag(a), mag(best)), mag(cand));
                    const double tol = gdt_x_tol4(gdt_x_err(L, M), M);
                    bool inside = P < -tol;
                    if (inside)
                        GDT_XAUDIT(1, -1, gdt_x_incircle(exact(a), exact(best), exact(cand), xc));
                    else if (P > tol)
                        GDT_XAUDIT(1, 1, gdt_x_incircle(exact(a), exact(best), exact(cand), xc));
                    else {
                        const int s = gdt_x_incircle(exact(a), exact(best), exact(cand), xc);
                        inside = s < 0 ||
                                 (s == 0 && gdt_x_incircle_tie(io.perm[j], io.perm[a.k],
                                                               io.perm[best.k], io.perm[k],
                                                               exact(a), exact(best), exact(cand), xc));
                    }
                    if (!inside) continue;
                }
                gdt_xc_take(best, cand);
                K = gdt_circ64(a.x, a.y, px, py);
            }
        };
        auto sc_done = [&](int r) {
            if (best.k < 0) return false;
            const double cov = (double)r * bsmin + covb;
            const double k0min = 1e-6 * (fabs(a.x * best.y) + fabs(a.y * best.x));
            if (!(K.k0 > k0min) || !(cov > 0.)) return false;
            return K.kx * K.kx + K.ky * K.ky <= cov * cov * (K.k0 * K.k0) * (1.0 - 1e-6);
        };
        if (gdt_scan_rings(S, cx, cy, io.cstart, sc_cell, sc_done) < 0) {
            res.err |= best.k < 0 ? ERR_NONEIGHBOUR : ERR_SEARCH;
            break;
        }
        if (best.k == nn.k) break;
        gdt_xc_take(a, best);
    }
    if (xc.overflow) res.err |= ERR_NONEIGHBOUR;
    res.count = count;
    res.npred = npred;
    return res;
}
