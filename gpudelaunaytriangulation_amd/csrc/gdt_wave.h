// gdt_wave.h -- one-ring construction of a single point by a whole wavefront
// (host/device routines), for the points the tile kernel hands back (circles
// reaching past the staged window, windows too full for LDS).
//
// The search is the one of the tile routine (gdt_tile.h): nearest neighbour,
// then gift wrapping, each search scanning the cells [-2,2]^2 around v's cell
// and then the rest of its current circle's bounding box (or, while that is not
// possible, further Chebyshev rings) until its answer is certified.  Here the
// cells of a pass are dealt to the lanes of a wavefront; each lane keeps a best
// candidate, and the pass ends with a reduction of the lane bests to the
// wave's best, which every lane then starts the next pass from.  Arithmetic is FP64
// with the displacement formula of the direct routine, so both precision paths
// give these points the same row.
//
// Reduction order does not change the answer: "c lies inside circle (v,a,b)"
// orders the candidates strictly left of v->a by the centre of their circle
// along the bisector of (v,a), a total order for generic input, so any
// reduction tree keeps the candidate a sequential scan keeps.  Nearest
// neighbour: minimum of (squared distance, slot).
//
// The routine is written against a wave policy W with
//     GdtWaveBest W::ring(const GdtWaveBest& start, bool wrap, double ax,
//                         double ay, GdtTieIds& t, Fn& lane_fn)
// (template argument TIES: whether exact ties are decided or only reported)
// which runs lane_fn(lane, best) on every lane (each starting from `start`)
// and returns the reduction of the lane bests, identical on every lane: a
// device policy with cross-lane shuffles, and a host policy that runs the
// lanes one after another and reduces with the same tree.
#pragma once

#include "gdt_point.h"

#define GDT_WAVE 64

struct GdtWaveBest {
    int k;             // slot, -1 none
    double x, y;       // displacement from v
    double d2;         // squared distance (nearest-neighbour search)
};

GDT_HD inline GdtWaveBest gdt_wave_none()
{
    GdtWaveBest b;
    b.k = -1;
    b.x = b.y = b.d2 = 0.;
    return b;
}

// What exact ties need: slot -> caller index, and the slots of v and a (a -1
// in the nearest-neighbour search).  Caller indices are read only for a tie.
// `tie` records that one was met by a routine built without the tie rule.
struct GdtTieIds {
    const int* perm;
    int v, a;
    bool tie;
};

// Whether candidate c beats the current best b.  TIES: exact ties by
// gdt_nn_tie / gdt_incircle_tie of the caller indices (a total order);
// otherwise a tie keeps b and is recorded in t.tie (the kernel that inlines
// this variant keeps its registers for its own loop).
template <bool TIES>
GDT_HD inline bool gdt_wave_better(bool wrap, double ax, double ay, GdtTieIds& t,
                                   const GdtWaveBest& b, const GdtWaveBest& c)
{
    if (c.k < 0) return false;
    if (b.k < 0) return true;
    if (!wrap) {
        if (!(c.d2 == b.d2)) return c.d2 < b.d2;
        if (!TIES) {
            t.tie = true;
            return false;
        }
        return gdt_nn_tie(t.perm[c.k], t.perm[b.k]);
    }
    const Circ64 K = gdt_circ64(ax, ay, b.x, b.y);
    const double P = gdt_incircle64(K, c.x, c.y);
    if (!(P == 0.)) return P < 0.;
    if (!TIES) {
        t.tie = true;
        return false;
    }
    return gdt_incircle_tie(t.perm[t.v], t.perm[t.a], t.perm[b.k], t.perm[c.k], ax, ay, b.x, b.y,
                            c.x, c.y);
}

// Reduction tree of the lane bests (the device policy's shuffle-down order):
// after the pass with offset `off`, lane i < off holds the better of lanes i
// and i + off.
struct GdtWaveHost {
    template <bool TIES, class Fn>
    GdtWaveBest ring(const GdtWaveBest& start, bool wrap, double ax, double ay, GdtTieIds& t,
                     Fn& lane_fn)
    {
        GdtWaveBest lb[GDT_WAVE];
        for (int l = 0; l < GDT_WAVE; ++l) {
            lb[l] = start;
            lane_fn(l, lb[l]);
        }
        for (int off = GDT_WAVE / 2; off > 0; off >>= 1)
            for (int i = 0; i < off; ++i)
                if (gdt_wave_better<TIES>(wrap, ax, ay, t, lb[i], lb[i + off])) lb[i] = lb[i + off];
        return lb[0];
    }
};

// One point's ring.  `lead` is true on the one lane that writes the row.
// npred counts the predicates this lane evaluated.  Built without the tie rule
// (TIES false), a point that meets an exact tie returns ERR_TIE (its row
// partly written) for the caller to hand to the variant with it.  Built for
// rectangles only (GEN false: the LDS tile kernel, whose sets are rectangles),
// the geometry takes the rectangular expressions, which the general ones
// reduce to bit for bit, and the kernel keeps the registers of the others.
template <bool TIES = true, bool GEN = true, class W>
GDT_HD inline GdtRingResult gdt_wave_ring(const GdtSet& S, int j, const GdtRingIO& io, W& w,
                                          bool lead)
{
    GdtRingResult res = {0, 0, 0u, 0u};
    const uint32_t pc = io.scell[j];
    const int cx = (int)(pc & 0xffffu), cy = (int)(pc >> 16);
    const gdt_d2 ov = io.off64[j];
    const double covb = GEN ? gdt_cover0(S, ov.x, ov.y)
                            : GDT_FMIN(GDT_FMIN(ov.x, S.bsx - ov.x), GDT_FMIN(ov.y, S.bsy - ov.y));
    const double bsmin = S.hmin;
    const size_t row = (size_t)io.perm[j] * (size_t)io.stride;
    unsigned npred = 0;

    GdtTieIds t;
    t.perm = io.perm;
    t.v = j;
    t.tie = false;
    // One search: nearest neighbour (wrap false) or successor of edge (v,a).
    auto search = [&](bool wrap, int a, double ax, double ay, GdtWaveBest& g) -> bool {
        t.a = a;
        g = gdt_wave_none();
        // One pass over the cells of [x0,x1] x [y0,y1] outside [-hole,hole]^2
        // (hole < 0: none), dealt to the lanes, then the wave reduction.
        auto pass = [&](int x0, int x1, int y0, int y1, int hole) {
            const int nx = x1 - x0 + 1, ncell = nx * (y1 - y0 + 1);
            auto lane_fn = [&](int lane, GdtWaveBest& m) {
                Circ64 K = {0., 0., 0.};
                if (wrap && m.k >= 0) K = gdt_circ64(ax, ay, m.x, m.y);
                for (int q = lane; q < ncell; q += GDT_WAVE) {
                    const int dx = x0 + q % nx, dy = y0 + q / nx;
                    if (dx >= -hole && dx <= hole && dy >= -hole && dy <= hole) continue;
                    const int x = gdt_wrap(cx + dx, S.xsize), y = gdt_wrap(cy + dy, S.ysize);
                    const int key = S.c0 + gdt_cell_key(S, x, y);
                    const uint32_t b = io.cstart[key], e = io.cstart[key + 1];
                    const double bx = GEN ? gdt_cbase_x(S, dx, dy, ov.x) : gdt_base64(dx, S.bsx, ov.x);
                    const double by = GEN ? gdt_cbase_y(S, dx, dy, ov.y) : gdt_base64(dy, S.bsy, ov.y);
                    for (uint32_t k = b; k < e; ++k) {
                        if ((int)k == j || (int)k == a) continue;
                        const gdt_d2 o = io.off64[k];
                        GdtWaveBest c;
                        c.k = (int)k;
                        c.x = gdt_disp64(bx, o.x);
                        c.y = gdt_disp64(by, o.y);
                        c.d2 = gdt_dist2_64(c.x, c.y);
                        ++npred;
                        if (!wrap) {
                            if (gdt_wave_better<TIES>(false, 0., 0., t, m, c)) m = c;
                            continue;
                        }
                        if (!(gdt_orient64(ax, ay, c.x, c.y) > 0.)) continue;
                        if (m.k >= 0) {
                            ++npred;
                            const double P = gdt_incircle64(K, c.x, c.y);
                            if (!(P <= 0.)) continue;
                            if (P == 0.) {
                                if constexpr (!TIES) {
                                    t.tie = true;
                                    continue;
                                } else if (!gdt_incircle_tie(io.perm[j], io.perm[a], io.perm[m.k],
                                                             io.perm[k], ax, ay, m.x, m.y, c.x,
                                                             c.y)) {
                                    continue;
                                }
                            }
                        }
                        m = c;
                        K = gdt_circ64(ax, ay, c.x, c.y);
                    }
                }
            };
            g = w.template ring<TIES>(g, wrap, ax, ay, t, lane_fn);
        };
        // The square [-hs,hs]^2 first (hs = 2), then either the rest of the
        // bounding box of the current circle in one pass -- every later best's
        // circle lies inside it (circles through v and a with centres left of
        // v->a are nested along the bisector), so the search is then certified
        // -- or, while there is no candidate or the box reaches past rmax, the
        // next Chebyshev ring.
        int hs = S.rmax < 2 ? S.rmax : 2;
        pass(-hs, hs, -hs, hs, -1);
        for (;;) {
            if (g.k >= 0) {
                const double cov = (double)hs * bsmin + covb;
                int dx0 = 0, dx1 = 0, dy0 = 0, dy1 = 0;
                bool boxed = false;
                if (!wrap) {
                    if (cov > 0. && g.d2 <= cov * cov * (1.0 - 1e-12)) return true;
                    const double rad = gdt_sqrt64(g.d2) * (1.0 + 1e-9);
                    if constexpr (GEN)
                        gdt_disc_cells(S, 0., 0., 0., 0., rad, ov.x, ov.y, 1e-3, dx0, dx1, dy0, dy1);
                    else
                        gdt_circle_cells<double>(0., 0., 0., 0., rad, ov.x, ov.y, S.ibsx, S.ibsy,
                                                 1e-3, dx0, dx1, dy0, dy1);
                    boxed = true;
                } else {
                    const Circ64 K = gdt_circ64(ax, ay, g.x, g.y);
                    const double k0min = 1e-6 * (fabs(ax * g.y) + fabs(ay * g.x));
                    if (K.k0 > k0min) {
                        if (cov > 0. &&
                            K.kx * K.kx + K.ky * K.ky <= cov * cov * (K.k0 * K.k0) * (1.0 - 1e-6))
                            return true;
                        const double i2k = gdt_rcp64(2. * K.k0);
                        const double qx = -K.kx * i2k, qy = -K.ky * i2k;
                        const double rad = gdt_sqrt64(qx * qx + qy * qy);
                        const double m = 1e-6 * (fabs(qx) + fabs(qy) + rad) + 1e-9 * S.bsx;
                        if constexpr (GEN)
                            gdt_disc_cells(S, qx - m, qx + m, qy - m, qy + m, rad, ov.x, ov.y, 1e-3,
                                           dx0, dx1, dy0, dy1);
                        else
                            gdt_circle_cells<double>(qx - m, qx + m, qy - m, qy + m, rad, ov.x,
                                                     ov.y, S.ibsx, S.ibsy, 1e-3, dx0, dx1, dy0, dy1);
                        boxed = true;
                    }
                }
                if (boxed) {
                    if (dx0 >= -hs && dx1 <= hs && dy0 >= -hs && dy1 <= hs) return true;
                    if (dx0 >= -S.rmax && dx1 <= S.rmax && dy0 >= -S.rmax && dy1 <= S.rmax) {
                        pass(dx0, dx1, dy0, dy1, hs);
                        return true;
                    }
                }
            }
            if (hs >= S.rmax) return false;
            ++hs;
            pass(-hs, hs, -hs, hs, hs - 1);
        }
    };

    GdtWaveBest g;
    if (!search(false, -1, 0., 0., g)) {
        res.err = g.k < 0 ? ERR_NONEIGHBOUR : ERR_SEARCH;
        res.npred = npred;
        return res;
    }
    const int nn = g.k;
    int a = nn, count = 0;
    double ax = g.x, ay = g.y;
    for (;;) {
        if (lead && count < io.stride) io.nbrs[row + count] = io.perm[a];
        ++count;
        if (count > GDT_HARD_CAP) { res.err |= ERR_NOCLOSE; break; }
        if (!search(true, a, ax, ay, g)) {
            res.err |= g.k < 0 ? ERR_NONEIGHBOUR : ERR_SEARCH;
            break;
        }
        if (!TIES && t.tie) break;      // handed on (the flag is the same on every lane)
        if (g.k == nn) break;
        a = g.k;
        ax = g.x;
        ay = g.y;
    }
    res.count = count;
    if (!TIES && t.tie) res.err |= ERR_TIE;
    res.npred = npred;
    return res;
}
