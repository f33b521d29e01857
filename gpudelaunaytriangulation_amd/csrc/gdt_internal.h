// gdt_internal.h -- shared device-side definitions of the gfx950 triangulation
// library: per-set geometry, cell keys, cell-local displacements and the
// FP64 / certified-FP32 predicates.
//
// Coordinates are never subtracted globally.  Every point is stored as (cell,
// offset inside the cell); the displacement of a candidate p from a point v is
//     d = (dcx * bs - off_v) + off_p          (per axis)
// with dcx the (unwrapped) cell distance.  The FP64 path evaluates exactly this
// expression in FP64; the filtered path evaluates it in FP32 from offsets
// rounded to FP32 and certifies every orientation / in-circle / distance
// decision against a forward error bound that includes the difference between
// the two inputs; undecided cases re-evaluate the FP64 expression with the same
// operation order, so both paths take identical decisions.
#pragma once

#include <stdint.h>
#include <string.h>
#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GDT_HD __host__ __device__
#else
// Host-only build of the per-point routines (CPU emulation tests).
#include <cmath>
#include <algorithm>
#define GDT_HD
#endif

#define GDT_BLOCK 256

// Audit of the FP32 filter (host test builds only): every decision the filter
// takes from its FP32 value alone is reported as a claimed sign (+1 / -1) of
// an exact predicate on the FP64 displacements the FP64 path would use (a, b,
// c relative to v):  ORIENT  a x c;  INCIRCLE  K0|c|^2 + Kx cx + Ky cy of the
// circle (v, a, b);  DIST  |c|^2 - |b|^2.  Compiled out everywhere else.
enum { GDT_AUD_ORIENT = 0, GDT_AUD_INCIRCLE = 1, GDT_AUD_DIST = 2 };
#ifndef GDT_FILTER_AUDIT
#define GDT_FILTER_AUDIT(kind, sign, ax, ay, bx, by, cx, cy) ((void)0)
#endif
#define GDT_HARD_CAP 256      // ring walk cap: a longer ring means non-generic input

// Device status word layout (gdt_status[...]).
enum {
    ST_REQUIRED = 0,     // largest ring size seen (atomicMax)
    ST_ERRFLAGS = 1,     // bit 0: search radius exhausted, bit 1: ring did not close, bit 2: no neighbour
    ST_NFLAGGED = 2,     // number of points flagged by the circumcircle test
    ST_NTRI = 3,         // number of triangles enumerated
    ST_NDEFER = 4,       // points deferred to the wavefront kernel (tile or per-point kernels)
    ST_NDEFER_TILE = 5,  // deferred points a tile block finished itself with a wavefront
    ST_WORDS = 8
};
// ERR_DEFER is internal: the LDS routine hands the point to the direct routine.
// ERR_TIE is internal: the wavefront routine built without the tie rule met
// an exact tie and hands the point to the one built with it.
// The precision code of the exact mode (GDT_PREC_EXACT of the public header).
enum { GDT_EXACT = 128 };
// Lanes per point of the global triangulation's star kernel (one wavefront).
enum { GDT_STAR_LANES = 64 };
enum { ERR_SEARCH = 1, ERR_NOCLOSE = 2, ERR_NONEIGHBOUR = 4, ERR_DEFER = 8, ERR_TIE = 16,
       ERR_RANGE = 32 /* exact mode: coordinates too far outside the box, or a general box */ };
// 64-bit counters (gdt_counters[...]).
// CT_NBRSUM: sum of ring lengths written; CT_INSUM: sum of input counts seen by
// the triangle enumeration (Euler check).
enum { CT_PRED = 0, CT_FALLBACK = 1, CT_NBRSUM = 2, CT_INSUM = 3, CT_WORDS = 4 };

// One independent periodic point set.  Points of set s occupy the range
// [p0, p0+n) both in the caller's (original) order and in cell (sorted) order;
// its cells occupy keys [c0, c0+nkeys).
struct GdtSet {
    int p0, n;
    int c0, nkeys;
    int xsize, ysize;
    int tiles_x, tiled;
    int rmax;            // largest Chebyshev search radius in cells
    int tshift;          // tiles of 2^tshift x 2^tshift cells (tiled orders)
    double lx, ly;       // box
    double bsx, bsy;     // cell sides
    double ibsx, ibsy;   // their reciprocals
    float bsx32, bsy32;
    float bsmax32, bsmin32;
    float ibsx32, ibsy32;
    float pad1, pad2;
    // Cell geometry of every box shape.  Rectangles: ex = (bsx, 0),
    // ey = (0, bsy).  General boxes (general = 1; periodicBoundaries::setGeneral,
    // periodicBoundaries.h:86-101): the cells are parallelograms spanned by
    // ex = a/xsize and ey = b/ysize, (a, b) the Lagrange-Gauss reduced basis
    // of the box's lattice (the same torus, its shortest and most orthogonal
    // basis), and bsx, bsy hold the cell heights.
    int general, pad3;
    double ex_x, ex_y, ey_x, ey_y;   // cell edge vectors
    double rxx, rxy, ryx, ryy;       // inverse of [ex ey]: cell coordinates of a displacement
    double nrx, nry;                 // norms of its rows: 1 / cell height
    double hmin;                     // smaller cell height
    double lat[4], ilat[4];          // reduced basis (a.x, b.x, a.y, b.y) and its inverse
    double box[4], ibox[4];          // the caller's box matrix and its inverse (wrapping)
};

struct GdtChunk { int set; int start; };   // a block's work: set id and first local item

// Cell order.  tiled = 1: tiles of 2^tshift x 2^tshift cells (32x32 or 16x16)
// in row-major tile order, cells row-major inside a tile, so one tile's points
// are one contiguous slot range (the unit of work of the LDS-staged kernels);
// tiled = 2: the same tiles with the cells of a tile in Morton (Z) order, so
// that consecutive slots (the points of one wavefront) form square patches;
// tiled = 0: row-major cells.  The tile shift is a field of the set; the
// LDS-staged kernels are compiled once per supported shift, GDT_TILE_SHIFT
// being the one a translation unit is compiled for.
#define GDT_TSHIFT_MIN 4
#define GDT_TSHIFT_MAX 5
#ifndef GDT_TILE_SHIFT
#define GDT_TILE_SHIFT 5
#endif
#define GDT_TILE (1 << GDT_TILE_SHIFT)
#define GDT_TILE_CELLS (GDT_TILE * GDT_TILE)

// Bits of v spread to the even positions (v < 2^8).
GDT_HD inline int gdt_spread_bits(int v)
{
    v = (v | (v << 4)) & 0x0f0f;
    v = (v | (v << 2)) & 0x3333;
    v = (v | (v << 1)) & 0x5555;
    return v;
}

GDT_HD inline int gdt_cell_key(const GdtSet& s, int cx, int cy)
{
    if (s.tiled) {
        const int ts = s.tshift, m = (1 << ts) - 1;
        const int tile = ((cy >> ts) * s.tiles_x) + (cx >> ts);
        const int ix = cx & m, iy = cy & m;
        const int inner = s.tiled == 2 ? (gdt_spread_bits(ix) | (gdt_spread_bits(iy) << 1))
                                       : ((iy << ts) | ix);
        return (tile << (2 * ts)) | inner;
    }
    return cy * s.xsize + cx;
}

GDT_HD inline int gdt_nkeys(int xsize, int ysize, int tiled, int tshift)
{
    if (!tiled) return xsize * ysize;
    const int t = 1 << tshift;
    return ((xsize + t - 1) / t) * ((ysize + t - 1) / t) * t * t;
}

// Offsets (dx,dy) of the i-th cell of the Chebyshev ring of radius r (8r cells).
GDT_HD inline void gdt_ring_cell(int r, int i, int& dx, int& dy)
{
    if (r == 0) { dx = 0; dy = 0; return; }
    const int side = i / (2 * r), t = i - side * 2 * r;
    switch (side) {
        case 0: dx = -r + t; dy = -r; break;
        case 1: dx = r; dy = -r + t; break;
        case 2: dx = r - t; dy = r; break;
        default: dx = -r; dy = r - t; break;
    }
}

GDT_HD inline int gdt_wrap(int x, int n)
{
    if (x < 0) x += n;
    if (x >= n) x -= n;
    return x;
}

// Set of a global point index (binary search over set starts).
GDT_HD inline int gdt_set_of(const GdtSet* sets, int nsets, int p)
{
    int lo = 0, hi = nsets - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (sets[mid].p0 <= p) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Wrap a point into the box (rectangles: [0,lx) x [0,ly); general boxes: the
// unit cell of the reduced basis), find its cell and its FP64 offset inside
// the cell (general boxes: relative to the cell's corner cx*ex + cy*ey).
// Every kernel that needs a point's cell calls this one routine, so all of
// them agree bit for bit.
GDT_HD inline void gdt_bin_point(const GdtSet& S, double x, double y, int& cx, int& cy,
                                 double& ox, double& oy)
{
    if (S.general) {
        double u = S.ilat[0] * x + S.ilat[1] * y, v = S.ilat[2] * x + S.ilat[3] * y;
        if (!(u >= 0. && u < 1. && v >= 0. && v < 1.)) {
            if (!(fabs(u) < 1e15 && fabs(v) < 1e15)) {
                x = y = u = v = 0.;
            } else {
                const double ku = floor(u), kv = floor(v);
                x -= ku * S.lat[0] + kv * S.lat[1];
                y -= ku * S.lat[2] + kv * S.lat[3];
                u -= ku;
                v -= kv;
            }
        }
        cx = (int)floor(u * (double)S.xsize);
        cy = (int)floor(v * (double)S.ysize);
        if (cx > S.xsize - 1) cx = S.xsize - 1;
        if (cy > S.ysize - 1) cy = S.ysize - 1;
        if (cx < 0) cx = 0;
        if (cy < 0) cy = 0;
        ox = x - ((double)cx * S.ex_x + (double)cy * S.ey_x);
        oy = y - ((double)cx * S.ex_y + (double)cy * S.ey_y);
        return;
    }
    if (!(x >= 0. && x < S.lx)) {
        x -= floor(x / S.lx) * S.lx;
        if (x >= S.lx) x -= S.lx;
        if (!(x >= 0.)) x = 0.;
    }
    if (!(y >= 0. && y < S.ly)) {
        y -= floor(y / S.ly) * S.ly;
        if (y >= S.ly) y -= S.ly;
        if (!(y >= 0.)) y = 0.;
    }
    cx = (int)floor(x / S.bsx);
    cy = (int)floor(y / S.bsy);
    if (cx > S.xsize - 1) cx = S.xsize - 1;
    if (cy > S.ysize - 1) cy = S.ysize - 1;
    if (cx < 0) cx = 0;
    if (cy < 0) cy = 0;
    ox = x - (double)cx * S.bsx;
    oy = y - (double)cy * S.bsy;
}

// Displacement of the corner of the cell (dx, dy) cells away minus the offset
// o of a point in the home cell: the candidate at offset p in that cell lies
// at base + p.  One fixed operation order for every box shape: for
// rectangles (ey_x = ex_y = 0) the fused term adds an exact zero, so this is
// dx*bsx - o.x, dy*bsy - o.y bit for bit.
GDT_HD inline double gdt_cbase_x(const GdtSet& S, int dx, int dy, double ox)
{
    return fma((double)dy, S.ey_x, (double)dx * S.ex_x) - ox;
}
GDT_HD inline double gdt_cbase_y(const GdtSet& S, int dx, int dy, double oy)
{
    return fma((double)dx, S.ex_y, (double)dy * S.ey_y) - oy;
}

// Distance from a point at offset (ox, oy) in its cell to the cell's walls:
// after the Chebyshev rings 0..r around the cell every point closer than
// r * hmin + this distance has been seen.
GDT_HD inline double gdt_cover0(const GdtSet& S, double ox, double oy)
{
    if (!S.general) {
        const double a = ox < S.bsx - ox ? ox : S.bsx - ox;
        const double b = oy < S.bsy - oy ? oy : S.bsy - oy;
        return a < b ? a : b;
    }
    const double fx = S.rxx * ox + S.rxy * oy, fy = S.ryx * ox + S.ryy * oy;
    const double a = (fx < 1. - fx ? fx : 1. - fx) * S.bsx;
    const double b = (fy < 1. - fy ? fy : 1. - fy) * S.bsy;
    return a < b ? a : b;
}

// Minimum image of a displacement between two nearby points (much closer than
// half a box thickness): rectangles by the rounded quotient of each side,
// general boxes by the rounded unit-cell coordinates in the reduced basis.
GDT_HD inline void gdt_min_image(const GdtSet& S, double& x, double& y)
{
    if (!S.general) {
        x -= S.lx * rint(x * (1.0 / S.lx));
        y -= S.ly * rint(y * (1.0 / S.ly));
        return;
    }
    const double ku = rint(S.ilat[0] * x + S.ilat[1] * y);
    const double kv = rint(S.ilat[2] * x + S.ilat[3] * y);
    x -= ku * S.lat[0] + kv * S.lat[1];
    y -= ku * S.lat[2] + kv * S.lat[3];
}

// Cell grid of a box (cellListGPU::setGridSize, cellListGPU.cpp:97-117): the
// count per axis is floor(L/a) rounded up to even; sides are L/count per axis.
GDT_HD inline void gdt_grid(double lx, double ly, double a, int& xsize, int& ysize)
{
    xsize = (int)floor(lx / a);
    if (xsize % 2 == 1) xsize += 1;
    ysize = (int)floor(ly / a);
    if (ysize % 2 == 1) ysize += 1;
    if (xsize < 2) xsize = 2;
    if (ysize < 2) ysize = 2;
}

// Geometry of one set; c0 is its first cell key.  rmax keeps every searched
// cell distinct on the torus (2*rmax+1 <= cells per axis).
// tshift: tiles of 2^tshift x 2^tshift cells (GDT_TSHIFT_MIN..GDT_TSHIFT_MAX).
inline GdtSet gdt_make_set(int p0, int n, int c0, double lx, double ly, double cell_size, int tiled,
                           int tshift = GDT_TILE_SHIFT)
{
    GdtSet s;
    memset(&s, 0, sizeof(s));
    s.p0 = p0; s.n = n; s.c0 = c0;
    gdt_grid(lx, ly, cell_size, s.xsize, s.ysize);
    s.tiled = tiled == 2 ? 2 : (tiled ? 1 : 0);
    s.tshift = tshift;
    s.tiles_x = (s.xsize + (1 << tshift) - 1) >> tshift;
    s.nkeys = gdt_nkeys(s.xsize, s.ysize, s.tiled, tshift);
    const int rx = (s.xsize - 1) / 2, ry = (s.ysize - 1) / 2;
    s.rmax = rx < ry ? rx : ry;
    s.lx = lx; s.ly = ly;
    s.bsx = lx / s.xsize;
    s.bsy = ly / s.ysize;
    s.ibsx = 1.0 / s.bsx;
    s.ibsy = 1.0 / s.bsy;
    s.bsx32 = (float)s.bsx;
    s.bsy32 = (float)s.bsy;
    s.ibsx32 = (float)s.ibsx;
    s.ibsy32 = (float)s.ibsy;
    s.pad1 = s.pad2 = 0.f;
    s.bsmax32 = (float)(s.bsx > s.bsy ? s.bsx : s.bsy);
    s.bsmin32 = (float)(s.bsx < s.bsy ? s.bsx : s.bsy);
    s.general = 0;
    s.ex_x = s.bsx; s.ey_y = s.bsy;
    s.rxx = s.ibsx; s.ryy = s.ibsy;
    s.nrx = s.ibsx; s.nry = s.ibsy;
    s.hmin = s.bsx < s.bsy ? s.bsx : s.bsy;
    s.lat[0] = s.box[0] = lx; s.lat[3] = s.box[3] = ly;
    s.ilat[0] = s.ibox[0] = 1.0 / lx; s.ilat[3] = s.ibox[3] = 1.0 / ly;
    return s;
}

// Whether a box matrix {x11, x12, x21, x22} is a rectangle with positive sides.
inline bool gdt_box_is_rect(const double* box)
{
    return box[1] == 0. && box[2] == 0. && box[0] > 0. && box[3] > 0.;
}

// Lagrange-Gauss reduction of the lattice basis (a, b): the shortest basis
// of the same lattice, |a| <= |b| and |a.b| <= |a|^2 / 2, oriented so that
// a x b > 0.  A sheared box and its reduced form are the same torus.
inline void gdt_reduce_basis(double& ax, double& ay, double& bx, double& by)
{
    for (int it = 0; it < 200; ++it) {
        if (ax * ax + ay * ay > bx * bx + by * by) {
            const double tx = ax, ty = ay;
            ax = bx; ay = by;
            bx = tx; by = ty;
        }
        const double mu = rint((ax * bx + ay * by) / (ax * ax + ay * ay));
        if (mu == 0.) break;
        bx -= mu * ax;
        by -= mu * ay;
    }
    if (ax * by - ay * bx < 0.) {
        bx = -bx;
        by = -by;
    }
}

// A set in a general box {x11, x12, x21, x22} (lattice vectors as columns),
// cells binned in the unit-cell coordinates of the reduced basis: cell
// counts from the thicknesses T of the unit cell across each pair of walls
// (the rectangle's side lengths for a rectangle), by the rule of gdt_grid.
// Row-major cell order (the LDS tile kernels handle rectangles only).
// A rectangular box gives gdt_make_set's set.
inline GdtSet gdt_make_set_box(int p0, int n, int c0, const double* box, double cell_size,
                               int tiled, int tshift = GDT_TILE_SHIFT)
{
    if (gdt_box_is_rect(box)) return gdt_make_set(p0, n, c0, box[0], box[3], cell_size, tiled, tshift);
    double ax = box[0], ay = box[2], bx = box[1], by = box[3];
    gdt_reduce_basis(ax, ay, bx, by);
    const double det = ax * by - ay * bx;
    const double tx = det / sqrt(bx * bx + by * by), ty = det / sqrt(ax * ax + ay * ay);
    GdtSet s = gdt_make_set(p0, n, c0, tx, ty, cell_size, 0, tshift);
    s.general = 1;
    s.lat[0] = ax; s.lat[1] = bx; s.lat[2] = ay; s.lat[3] = by;
    s.ilat[0] = by / det; s.ilat[1] = -bx / det; s.ilat[2] = -ay / det; s.ilat[3] = ax / det;
    const double fx = (double)s.xsize, fy = (double)s.ysize;
    s.ex_x = ax / fx; s.ex_y = ay / fx;
    s.ey_x = bx / fy; s.ey_y = by / fy;
    s.rxx = fx * s.ilat[0]; s.rxy = fx * s.ilat[1];
    s.ryx = fy * s.ilat[2]; s.ryy = fy * s.ilat[3];
    s.nrx = sqrt(s.rxx * s.rxx + s.rxy * s.rxy);
    s.nry = sqrt(s.ryx * s.ryx + s.ryy * s.ryy);
    s.bsx = tx / fx;            // cell heights
    s.bsy = ty / fy;
    s.ibsx = s.nrx;
    s.ibsy = s.nry;
    s.hmin = s.bsx < s.bsy ? s.bsx : s.bsy;
    for (int k = 0; k < 4; ++k) s.box[k] = box[k];
    const double bd = box[0] * box[3] - box[1] * box[2];
    s.ibox[0] = box[3] / bd; s.ibox[1] = -box[1] / bd;
    s.ibox[2] = -box[2] / bd; s.ibox[3] = box[0] / bd;
    return s;
}

// Wrap a moved point into the caller's box (periodicBoundaries::putInBoxReal,
// periodicBoundaries.h:131-152): rectangles per axis into [0, L), general
// boxes by whole lattice vectors of the caller's basis into its unit cell.
GDT_HD inline void gdt_wrap_into_box(const GdtSet& S, double& x, double& y)
{
    if (!S.general) {
        x = x - floor(x / S.lx) * S.lx;
        if (x >= S.lx) x -= S.lx;
        if (x < 0.0) x = 0.0;
        y = y - floor(y / S.ly) * S.ly;
        if (y >= S.ly) y -= S.ly;
        if (y < 0.0) y = 0.0;
        return;
    }
    const double ku = floor(S.ibox[0] * x + S.ibox[1] * y);
    const double kv = floor(S.ibox[2] * x + S.ibox[3] * y);
    if (fabs(ku) < 1e15 && fabs(kv) < 1e15) {
        x -= ku * S.box[0] + kv * S.box[1];
        y -= ku * S.box[2] + kv * S.box[3];
    }
}

// ---------------------------------------------------------------------------
// FP64 displacement and predicates.  One fixed operation order with explicit
// fused multiply-adds (the library is compiled with -ffp-contract=off), used
// by every FP64 decision of both paths, so the filtered path's fallback
// reproduces the FP64 path bit for bit.
// ---------------------------------------------------------------------------
GDT_HD inline double gdt_base64(int dc, double bs, double ov)
{
    return (double)dc * bs - ov;
}
GDT_HD inline double gdt_disp64(double base, double op) { return base + op; }

// a x c  (> 0: c strictly left of the ray from the origin through a)
GDT_HD inline double gdt_orient64(double ax, double ay, double cx, double cy)
{
    return fma(ax, cy, -(ay * cx));
}

GDT_HD inline double gdt_dist2_64(double x, double y) { return fma(x, x, y * y); }

// In-circle coefficients of the circle through the origin, a and b:
//   P(c) = K0 |c|^2 + Kx cx + Ky cy = K0 (|c - q|^2 - R^2),
// q = -(Kx,Ky)/(2 K0) the centre: P < 0 iff c strictly inside (K0 > 0).
struct Circ64 { double k0, kx, ky; };
GDT_HD inline Circ64 gdt_circ64(double ax, double ay, double bx, double by)
{
    const double aa = gdt_dist2_64(ax, ay);
    const double bb = gdt_dist2_64(bx, by);
    Circ64 c;
    c.k0 = fma(ax, by, -(ay * bx));
    c.kx = fma(bb, ay, -(aa * by));
    c.ky = fma(aa, bx, -(bb * ax));
    return c;
}
GDT_HD inline double gdt_incircle_cc64(const Circ64& k, double cx, double cy, double cc)
{
    return fma(k.k0, cc, fma(k.kx, cx, k.ky * cy));
}
GDT_HD inline double gdt_incircle64(const Circ64& k, double cx, double cy)
{
    return gdt_incircle_cc64(k, cx, cy, gdt_dist2_64(cx, cy));
}

// Exact ties (degenerate input: a candidate exactly on the circle, two points
// at exactly the same distance) are decided by a symbolic perturbation, the
// same one in every search of every point, so that the rings of a lattice form
// one triangulation: each point's lifted coordinate |p|^2 is raised by
// eps^(1 + caller index), the lowest caller index most.  Nearest neighbour:
// of two points at the same distance the lower caller index wins.  In-circle:
// c, exactly on the circle through v, a and b (displacements from v; b and c
// left of v->a), counts as inside when the most raised of the four points is
// b (b and c lie on the same side of line va), or v, a and c lies on that
// point's side of the line through the two others (raising one point's lift
// tilts the plane through the three about the line through the other two),
// and as outside when it is c.  p* are caller indices.
GDT_HD inline int gdt_imin(int x, int y) { return x < y ? x : y; }
GDT_HD inline bool gdt_nn_tie(int pc, int pb) { return pc < pb; }
GDT_HD inline bool gdt_incircle_tie(int pv, int pa, int pb, int pc, double ax, double ay,
                                    double bx, double by, double cx, double cy)
{
    const int m = gdt_imin(gdt_imin(pv, pa), gdt_imin(pb, pc));
    if (m == pc) return false;
    if (m == pb) return true;
    if (m == pa)   // c on a's side of line v b
        return (gdt_orient64(bx, by, ax, ay) > 0.) == (gdt_orient64(bx, by, cx, cy) > 0.);
    // v dominant: c on v's side of line a b (v is the origin)
    const double ux = bx - ax, uy = by - ay;
    return (gdt_orient64(ux, uy, -ax, -ay) > 0.) == (gdt_orient64(ux, uy, cx - ax, cy - ay) > 0.);
}

// ---------------------------------------------------------------------------
// FP32 filter constants.  u = 2^-24.  The FP32 displacement differs from the
// FP64 one by at most EPS1*|d| + EPS0*bs per coordinate (rounding of the two
// offsets, of dcx*bs and of the two sums); all bounds below include a factor 2
// of slack for the rounding of the bound itself.
// ---------------------------------------------------------------------------
// A decision the FP32 filter takes from its FP32 value alone.  A host test
// build (GDT_FILTER_FORCE_FP64) takes none: every predicate then goes to the
// FP64 fallback, whose rows must be those of the FP64 path.
// GDT_FILTER_FORCED(x): a condition that only such a build needs (where the
// FP32 test would otherwise have decided it).
#ifdef GDT_FILTER_FORCE_FP64
#define GDT_FILTER_DECIDES(x) false
#define GDT_FILTER_FORCED(x) (x)
#else
#define GDT_FILTER_DECIDES(x) (x)
#define GDT_FILTER_FORCED(x) false
#endif

#ifndef GDT_U32     // (a host test shrinks it to show that a too-tight bound is caught)
#define GDT_U32 5.9604644775390625e-08f
#endif
