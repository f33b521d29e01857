// gdt_point.h -- per-point Delaunay one-ring construction (host/device).
//
// For one point v the ring is built by gift wrapping over the cell grid:
//   1. n0 = nearest neighbour of v (a Delaunay edge);
//   2. given the Delaunay edge (v, a), the next neighbour counter-clockwise is
//      the candidate c strictly left of v->a for which no other left candidate
//      lies strictly inside the circle (v, a, c): keep a running best b and
//      replace it by c whenever c is inside circle (v, a, b), i.e. whenever
//      P(c) = K0|c|^2 + Kx cx + Ky cy < 0 for the coefficients of (v, a, b);
//   3. stop when the walk returns to n0.
// Candidates are visited by Chebyshev rings of cells around v's cell.  After
// ring r every point closer than cov(r) = r*bs_min + (distance of v to its own
// cell walls) has been seen, so a search may stop as soon as its answer is
// certified: the nearest distance is below cov(r), or the current circle
// (v, a, b) has diameter 2R <= cov(r) (any point strictly inside it is closer
// than 2R to v).
//
// This replaces the reference's two-phase virtual-polygon clipping
// (virtual_voronoi_calc_function, DelaunayGPU.cu:318-608, and
// get_oneRing_function, DelaunayGPU.cu:680-943) with the same output: the
// counter-clockwise Delaunay one-ring of every point, which is unique for
// generic input.  Each reference test "point inside circle (v, P_j, P_j+1)"
// (DelaunayGPU.cu:766-791) is the in-circle predicate evaluated here.
#pragma once

#include "gdt_internal.h"

struct alignas(16) gdt_d2 { double x, y; };
struct alignas(8) gdt_f2 { float x, y; };

GDT_HD inline int gdt_pack_dxy(int dx, int dy)
{
    return (int)(((unsigned)dx & 0xffffu) | ((unsigned)dy << 16));
}
GDT_HD inline int gdt_dx_of(int p) { return (int)(short)(p & 0xffff); }
GDT_HD inline int gdt_dy_of(int p) { return p >> 16; }

#if defined(__HIPCC__)
#define GDT_FMIN fmin
#define GDT_FMAX fmax
#define GDT_FMINF fminf
#define GDT_FMAXF fmaxf
#else
#define GDT_FMIN std::fmin
#define GDT_FMAX std::fmax
#define GDT_FMINF std::fmin
#define GDT_FMAXF std::fmax
#endif

// Arrays the one-ring construction reads (cell order) and writes (caller order).
struct GdtRingIO {
    const uint32_t* cstart;   // [keys+1] first sorted slot of each cell key
    const int* perm;          // [N] sorted slot -> caller index
    const gdt_d2* off64;      // [N] FP64 offset of each slot inside its cell
    const gdt_f2* off32;      // [N] the same offsets rounded to FP32
    const uint32_t* scell;    // [N] packed (cx | cy << 16) of each slot
    int* nbrs;                // [N][stride] rings, caller order
    int stride;
    const gdt_d2* xy;         // [N] caller coordinates (exact mode; else may be null)
};

struct GdtRingResult {
    int count;          // ring length (may exceed stride; then the row is truncated)
    int err;            // ERR_* bits
    unsigned npred;     // predicate evaluations
    unsigned nfb;       // FP64 fallbacks (filtered path)
};

// Visit the Chebyshev rings r = 0..rmax around cell (cx,cy); cell_fn(dx,dy,r,b,e)
// is called for each non-empty cell (slots [b,e)), done(r) after each ring.
// Returns the ring at which done() returned true, or -1.
template <class CellFn, class DoneFn>
GDT_HD inline int gdt_scan_rings(const GdtSet& S, int cx, int cy, const uint32_t* cstart,
                                 CellFn& cell_fn, DoneFn& done)
{
    for (int r = 0; r <= S.rmax; ++r) {
        const int ncell = r ? 8 * r : 1;
        for (int q = 0; q < ncell; ++q) {
            int dx, dy;
            gdt_ring_cell(r, q, dx, dy);
            const int x = gdt_wrap(cx + dx, S.xsize), y = gdt_wrap(cy + dy, S.ysize);
            const int key = S.c0 + gdt_cell_key(S, x, y);
            const uint32_t b = cstart[key], e = cstart[key + 1];
            if (b < e) cell_fn(dx, dy, r, b, e);
        }
        if (done(r)) return r;
    }
    return -1;
}

// FP64 displacement of slot k seen from slot j (with j's offset ov) when k
// lies in the cell at offset (dx,dy) from j's cell: the one expression both
// paths use.
GDT_HD inline void gdt_disp64_of(const GdtSet& S, const gdt_d2& ov, const gdt_d2* off64, int k,
                                 int dxy, double& x, double& y)
{
    const gdt_d2 o = off64[k];
    const int dx = gdt_dx_of(dxy), dy = gdt_dy_of(dxy);
    x = gdt_disp64(gdt_cbase_x(S, dx, dy, ov.x), o.x);
    y = gdt_disp64(gdt_cbase_y(S, dx, dy, ov.y), o.y);
}

// ---------------------------------------------------------------------------
// FP64 path
// ---------------------------------------------------------------------------
GDT_HD inline GdtRingResult gdt_one_ring_f64(const GdtSet& S, int j, const GdtRingIO& io)
{
    GdtRingResult res = {0, 0, 0u, 0u};
    const uint32_t pc = io.scell[j];
    const int cx = (int)(pc & 0xffffu), cy = (int)(pc >> 16);
    const gdt_d2 ov = io.off64[j];
    const double covb = gdt_cover0(S, ov.x, ov.y);
    const double bsmin = S.hmin;
    const size_t row = (size_t)io.perm[j] * (size_t)io.stride;
    unsigned npred = 0;

    // 1. nearest neighbour
    int nn = -1;
    double nd2 = 0., nx = 0., ny = 0.;
    auto nn_cell = [&](int dx, int dy, int, uint32_t b, uint32_t e) {
        const double bx = gdt_cbase_x(S, dx, dy, ov.x), by = gdt_cbase_y(S, dx, dy, ov.y);
        for (uint32_t k = b; k < e; ++k) {
            if ((int)k == j) continue;
            const gdt_d2 o = io.off64[k];
            const double px = gdt_disp64(bx, o.x), py = gdt_disp64(by, o.y);
            const double d2 = gdt_dist2_64(px, py);
            ++npred;
            if (nn < 0 || d2 < nd2 || (d2 == nd2 && gdt_nn_tie(io.perm[k], io.perm[nn]))) {
                nn = (int)k; nd2 = d2; nx = px; ny = py;
            }
        }
    };
    auto nn_done = [&](int r) {
        if (nn < 0) return false;
        const double cov = (double)r * bsmin + covb;
        return cov > 0. && nd2 <= cov * cov * (1.0 - 1e-12);
    };
    if (gdt_scan_rings(S, cx, cy, io.cstart, nn_cell, nn_done) < 0) {
        res.err = nn < 0 ? ERR_NONEIGHBOUR : ERR_SEARCH;
        res.npred = npred;
        return res;
    }

    // 2. walk the ring counter-clockwise
    int a = nn, count = 0;
    double ax = nx, ay = ny;
    for (;;) {
        if (count < io.stride) io.nbrs[row + count] = io.perm[a];
        ++count;
        if (count > GDT_HARD_CAP) { res.err |= ERR_NOCLOSE; break; }
        int best = -1;
        double bx_ = 0., by_ = 0.;
        Circ64 K = {0., 0., 0.};
        auto sc_cell = [&](int dx, int dy, int, uint32_t b, uint32_t e) {
            const double bx = gdt_cbase_x(S, dx, dy, ov.x), by = gdt_cbase_y(S, dx, dy, ov.y);
            for (uint32_t k = b; k < e; ++k) {
                if ((int)k == j || (int)k == a) continue;
                const gdt_d2 o = io.off64[k];
                const double px = gdt_disp64(bx, o.x), py = gdt_disp64(by, o.y);
                ++npred;
                if (!(gdt_orient64(ax, ay, px, py) > 0.)) continue;
                if (best >= 0) {
                    ++npred;
                    const double P = gdt_incircle64(K, px, py);
                    if (!(P < 0.) && !(P == 0. && gdt_incircle_tie(io.perm[j], io.perm[a], io.perm[best],
                                                                    io.perm[k], ax, ay, bx_, by_, px, py)))
                        continue;
                }
                best = (int)k; bx_ = px; by_ = py;
                K = gdt_circ64(ax, ay, px, py);
            }
        };
        auto sc_done = [&](int r) {
            if (best < 0) return false;
            const double cov = (double)r * bsmin + covb;
            const double k0min = 1e-6 * (fabs(ax * by_) + fabs(ay * bx_));
            if (!(K.k0 > k0min) || !(cov > 0.)) return false;
            return K.kx * K.kx + K.ky * K.ky <= cov * cov * (K.k0 * K.k0) * (1.0 - 1e-6);
        };
        if (gdt_scan_rings(S, cx, cy, io.cstart, sc_cell, sc_done) < 0) {
            res.err |= best < 0 ? ERR_NONEIGHBOUR : ERR_SEARCH;
            break;
        }
        if (best == nn) break;
        a = best; ax = bx_; ay = by_;
    }
    res.count = count;
    res.npred = npred;
    return res;
}

// ---------------------------------------------------------------------------
// FP32-filtered path with FP64 fallback
//
// Input error: |d32 - d64| <= 5u|d| + 7u*bs per coordinate (bs = larger cell
// side).  With max-norms ma, mb, mc of the three displacements:
//   orientation a x c : |err| <= 24u*ma*mc + 14u*bs*(ma+mc)
//   in-circle P(a,b,c): |err| <= 112u*ma*mb*mc*S + 28u*bs*S^3,  S = ma+mb+mc
//   squared distance  : |err| <= 13u*|c|^2 + 28u*bs*mc
// (first-order terms; all three are doubled below).  A decision is taken in
// FP32 only when the FP32 value clears its bound; otherwise the FP64
// expression of the FP64 path decides.
// ---------------------------------------------------------------------------
GDT_HD inline GdtRingResult gdt_one_ring_f32(const GdtSet& S, int j, const GdtRingIO& io)
{
    const float u = GDT_U32;
    GdtRingResult res = {0, 0, 0u, 0u};
    const uint32_t pc = io.scell[j];
    const int cx = (int)(pc & 0xffffu), cy = (int)(pc >> 16);
    const gdt_d2 ov = io.off64[j];
    const gdt_f2 ov32 = io.off32[j];
    const float bsx = S.bsx32, bsy = S.bsy32, bsM = S.bsmax32, bsm = S.bsmin32;
    const float covb = GDT_FMINF(GDT_FMINF(ov32.x, bsx - ov32.x), GDT_FMINF(ov32.y, bsy - ov32.y));
    const size_t row = (size_t)io.perm[j] * (size_t)io.stride;
    unsigned npred = 0, nfb = 0;
    auto covs = [&](int r) {
        return (float)r * bsm + covb - 9.5367431640625e-07f * (float)(r + 2) * bsM;
    };
    auto mcell = [&](int dx, int dy) {
        const int r = dx < 0 ? -dx : dx, s = dy < 0 ? -dy : dy;
        return (float)((r > s ? r : s) + 1) * bsM * 1.0001f;
    };

    // 1. nearest neighbour
    int nn = -1, ndxy = 0;
    float nd2 = 0.f, nE = 0.f, nx = 0.f, ny = 0.f;
    auto nn_cell = [&](int dx, int dy, int, uint32_t b, uint32_t e) {
        const float bx = (float)dx * bsx - ov32.x, by = (float)dy * bsy - ov32.y;
        const float eabs = 56.f * u * bsM * mcell(dx, dy) + 1e-30f;
        for (uint32_t k = b; k < e; ++k) {
            if ((int)k == j) continue;
            const gdt_f2 o = io.off32[k];
            const float px = bx + o.x, py = by + o.y;
            const float d2 = fmaf(px, px, py * py);
            const float E = fmaf(26.f * u, d2, eabs);
            ++npred;
            bool take;
#ifdef GDT_AUDIT_FILTER
            auto audit_nn = [&](int sign) {
                double cx64, cy64, nx64, ny64;
                gdt_disp64_of(S, ov, io.off64, (int)k, gdt_pack_dxy(dx, dy), cx64, cy64);
                gdt_disp64_of(S, ov, io.off64, nn, ndxy, nx64, ny64);
                GDT_FILTER_AUDIT(GDT_AUD_DIST, sign, 0., 0., nx64, ny64, cx64, cy64);
            };
#endif
            if (nn < 0) take = true;
            else if (GDT_FILTER_DECIDES(d2 + E < nd2 - nE)) {
                take = true;
#ifdef GDT_AUDIT_FILTER
                audit_nn(-1);
#endif
            } else if (GDT_FILTER_DECIDES(d2 - E > nd2 + nE)) {
                take = false;
#ifdef GDT_AUDIT_FILTER
                audit_nn(1);
#endif
            } else {
                ++nfb;
                double cx64, cy64, nx64, ny64;
                gdt_disp64_of(S, ov, io.off64, (int)k, gdt_pack_dxy(dx, dy), cx64, cy64);
                gdt_disp64_of(S, ov, io.off64, nn, ndxy, nx64, ny64);
                const double dc = gdt_dist2_64(cx64, cy64), dn = gdt_dist2_64(nx64, ny64);
                take = dc < dn || (dc == dn && gdt_nn_tie(io.perm[k], io.perm[nn]));
            }
            if (take) { nn = (int)k; ndxy = gdt_pack_dxy(dx, dy); nd2 = d2; nE = E; nx = px; ny = py; }
        }
    };
    auto nn_done = [&](int r) {
        if (nn < 0) return false;
        const float c = covs(r);
        return c > 0.f && nd2 + nE <= c * c * 0.9999f;
    };
    if (gdt_scan_rings(S, cx, cy, io.cstart, nn_cell, nn_done) < 0) {
        res.err = nn < 0 ? ERR_NONEIGHBOUR : ERR_SEARCH;
        res.npred = npred;
        res.nfb = nfb;
        return res;
    }

    // 2. walk the ring counter-clockwise
    int a = nn, adxy = ndxy, count = 0;
    float ax = nx, ay = ny;
    for (;;) {
        if (count < io.stride) io.nbrs[row + count] = io.perm[a];
        ++count;
        if (count > GDT_HARD_CAP) { res.err |= ERR_NOCLOSE; break; }
        const float ma = GDT_FMAXF(fabsf(ax), fabsf(ay));
        const float alphaO = 48.f * u * ma + 28.f * u * bsM;
        const float betaO = 28.f * u * bsM * ma + 1e-30f;
        int best = -1, bdxy = 0;
        float bx_ = 0.f, by_ = 0.f;
        float K0 = 0.f, Kx = 0.f, Ky = 0.f, A0 = 0.f, A1 = 0.f, A2 = 0.f, A3 = 0.f, EK0 = 0.f, EK = 0.f;
        auto set_best = [&](int k, int dxy, float px, float py) {
            best = k; bdxy = dxy; bx_ = px; by_ = py;
            const float aa = fmaf(ax, ax, ay * ay), bb = fmaf(px, px, py * py);
            K0 = fmaf(ax, py, -(ay * px));
            Kx = fmaf(bb, ay, -(aa * py));
            Ky = fmaf(aa, px, -(bb * ax));
            const float mb = GDT_FMAXF(fabsf(px), fabsf(py));
            const float sab = ma + mb, pab = ma * mb;
            A3 = 56.f * u * bsM;
            A2 = 224.f * u * pab + 168.f * u * bsM * sab;
            A1 = 224.f * u * pab * sab + 168.f * u * bsM * sab * sab;
            A0 = 56.f * u * bsM * sab * sab * sab + 1e-30f;
            EK0 = 48.f * u * pab + 28.f * u * bsM * sab;
            EK = 84.f * u * pab * sab + 56.f * u * bsM * sab * sab;
        };
        auto sc_cell = [&](int dx, int dy, int, uint32_t b, uint32_t e) {
            const float bx = (float)dx * bsx - ov32.x, by = (float)dy * bsy - ov32.y;
            const float mc = mcell(dx, dy);
            const float BO = fmaf(alphaO, mc, betaO);
            float BP = fmaf(mc, fmaf(mc, fmaf(mc, A3, A2), A1), A0);
            const int dxy = gdt_pack_dxy(dx, dy);
            for (uint32_t k = b; k < e; ++k) {
                if ((int)k == j || (int)k == a) continue;
                const gdt_f2 o = io.off32[k];
                const float px = bx + o.x, py = by + o.y;
                const float O = fmaf(ax, py, -(ay * px));
                ++npred;
#ifdef GDT_AUDIT_FILTER
                auto audit_sc = [&](int kind, int sign) {
                    double a64x, a64y, b64x = 0., b64y = 0., c64x, c64y;
                    gdt_disp64_of(S, ov, io.off64, a, adxy, a64x, a64y);
                    if (best >= 0) gdt_disp64_of(S, ov, io.off64, best, bdxy, b64x, b64y);
                    gdt_disp64_of(S, ov, io.off64, (int)k, dxy, c64x, c64y);
                    GDT_FILTER_AUDIT(kind, sign, a64x, a64y, b64x, b64y, c64x, c64y);
                };
                if (O > BO) audit_sc(GDT_AUD_ORIENT, 1);
                if (O < -BO) audit_sc(GDT_AUD_ORIENT, -1);
#endif
                bool left;
                if (GDT_FILTER_DECIDES(O > BO)) left = true;
                else if (GDT_FILTER_DECIDES(O < -BO)) left = false;
                else {
                    ++nfb;
                    double a64x, a64y, c64x, c64y;
                    gdt_disp64_of(S, ov, io.off64, a, adxy, a64x, a64y);
                    gdt_disp64_of(S, ov, io.off64, (int)k, dxy, c64x, c64y);
                    left = gdt_orient64(a64x, a64y, c64x, c64y) > 0.;
                }
                if (!left) continue;
                if (best >= 0) {
                    const float cc = fmaf(px, px, py * py);
                    const float P = fmaf(K0, cc, fmaf(Kx, px, Ky * py));
                    ++npred;
                    bool inside;
#ifdef GDT_AUDIT_FILTER
                    if (P < -BP) audit_sc(GDT_AUD_INCIRCLE, -1);
                    if (P > BP) audit_sc(GDT_AUD_INCIRCLE, 1);
#endif
                    if (GDT_FILTER_DECIDES(P < -BP)) inside = true;
                    else if (GDT_FILTER_DECIDES(P > BP)) inside = false;
                    else {
                        ++nfb;
                        double a64x, a64y, b64x, b64y, c64x, c64y;
                        gdt_disp64_of(S, ov, io.off64, a, adxy, a64x, a64y);
                        gdt_disp64_of(S, ov, io.off64, best, bdxy, b64x, b64y);
                        gdt_disp64_of(S, ov, io.off64, (int)k, dxy, c64x, c64y);
                        const Circ64 K64 = gdt_circ64(a64x, a64y, b64x, b64y);
                        const double P64 = gdt_incircle64(K64, c64x, c64y);
                        inside = P64 < 0. ||
                                 (P64 == 0. && gdt_incircle_tie(io.perm[j], io.perm[a], io.perm[best],
                                                                io.perm[k], a64x, a64y, b64x, b64y,
                                                                c64x, c64y));
                    }
                    if (!inside) continue;
                }
                set_best((int)k, dxy, px, py);
                BP = fmaf(mc, fmaf(mc, fmaf(mc, A3, A2), A1), A0);
            }
        };
        auto sc_done = [&](int r) {
            if (best < 0) return false;
            const float k0l = K0 - EK0;
            const float c = covs(r);
            if (!(k0l > 0.f) || !(c > 0.f)) return false;
            const float gx = fabsf(Kx) + EK, gy = fabsf(Ky) + EK;
            return (gx * gx + gy * gy) * 1.001f <= (c * c) * (k0l * k0l);
        };
        if (gdt_scan_rings(S, cx, cy, io.cstart, sc_cell, sc_done) < 0) {
            res.err |= best < 0 ? ERR_NONEIGHBOUR : ERR_SEARCH;
            break;
        }
        if (best == nn) break;
        a = best; adxy = bdxy; ax = bx_; ay = by_;
    }
    res.count = count;
    res.npred = npred;
    res.nfb = nfb;
    return res;
}

// Reciprocal and square root for the certified boxes and the scan clipping:
// a few ulp of error on the device (hardware approximation; one Newton step
// for the FP64 reciprocal), far below the margins those boxes carry.  The
// host build (CPU emulation) uses the correctly rounded operations.
GDT_HD inline double gdt_rcp64(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
#else
    return 1.0 / x;
#endif
}
GDT_HD inline double gdt_sqrt64(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrt(x);
#else
    return std::sqrt(x);
#endif
}
GDT_HD inline float gdt_rcp32(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.f / x;
#endif
}
GDT_HD inline float gdt_sqrt32(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return std::sqrt(x);
#endif
}

// Cell-offset bounding box of the circle through the origin, a and b, given
// centre bounds [qxlo,qxhi] x [qylo,qyhi] and radius bound R, for a point whose
// offset inside its cell is (ox,oy); ibsx, ibsy are the reciprocal cell sides
// and `margin` is in cells.
GDT_HD inline int gdt_cell_floor(double x)
{
    return (int)floor(GDT_FMIN(GDT_FMAX(x, -1e9), 1e9));
}
template <class T>
GDT_HD inline void gdt_circle_cells(T qxlo, T qxhi, T qylo, T qyhi, T R, T ox, T oy, T ibsx,
                                    T ibsy, T margin, int& dx0, int& dx1, int& dy0, int& dy1)
{
    dx0 = gdt_cell_floor((double)((ox + (qxlo - R)) * ibsx - margin));
    dx1 = gdt_cell_floor((double)((ox + (qxhi + R)) * ibsx + margin));
    dy0 = gdt_cell_floor((double)((oy + (qylo - R)) * ibsy - margin));
    dy1 = gdt_cell_floor((double)((oy + (qyhi + R)) * ibsy + margin));
}

// The same cell box for every box shape: cells (relative to the home cell of
// a point at offset (ox, oy)) that a disc of radius R, centred in
// [qxlo,qxhi] x [qylo,qyhi] relative to the point, can touch, with `margin`
// cells of slack.  General boxes: the cell coordinates of o + q + w are
// R(o + q) + Rw, and |(Rw)_x| <= |w| * (norm of R's first row).
GDT_HD inline void gdt_disc_cells(const GdtSet& S, double qxlo, double qxhi, double qylo,
                                  double qyhi, double R, double ox, double oy, double margin,
                                  int& dx0, int& dx1, int& dy0, int& dy1)
{
    if (!S.general) {
        gdt_circle_cells<double>(qxlo, qxhi, qylo, qyhi, R, ox, oy, S.ibsx, S.ibsy, margin, dx0,
                                 dx1, dy0, dy1);
        return;
    }
    const double fx = S.rxx * ox + S.rxy * oy, fy = S.ryx * ox + S.ryy * oy;
    const double ax0 = S.rxx * qxlo, ax1 = S.rxx * qxhi, ay0 = S.rxy * qylo, ay1 = S.rxy * qyhi;
    const double bx0 = S.ryx * qxlo, bx1 = S.ryx * qxhi, by0 = S.ryy * qylo, by1 = S.ryy * qyhi;
    const double xlo = fx + GDT_FMIN(ax0, ax1) + GDT_FMIN(ay0, ay1) - R * S.nrx;
    const double xhi = fx + GDT_FMAX(ax0, ax1) + GDT_FMAX(ay0, ay1) + R * S.nrx;
    const double ylo = fy + GDT_FMIN(bx0, bx1) + GDT_FMIN(by0, by1) - R * S.nry;
    const double yhi = fy + GDT_FMAX(bx0, bx1) + GDT_FMAX(by0, by1) + R * S.nry;
    dx0 = gdt_cell_floor(xlo - margin);
    dx1 = gdt_cell_floor(xhi + margin);
    dy0 = gdt_cell_floor(ylo - margin);
    dy1 = gdt_cell_floor(yhi + margin);
}
