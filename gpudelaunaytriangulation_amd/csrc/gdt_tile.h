// gdt_tile.h -- LDS-staged one-ring construction (host/device routines).
//
// A block owns one GDT_TILE x GDT_TILE-cell tile (32x32; a contiguous range of cell-ordered
// slots).  It stages the tile's window -- the tile plus a halo of GDT_HALO
// cells on every side, wrapped on the torus -- into LDS, with every point
// written as a coordinate in the window frame,
//     X = f * bs + off,   f = (window cell index) - GDT_FRAME,
// in FP64 and (filtered path) FP32.  A lane then walks its point's one-ring
// as in gdt_point.h (nearest neighbour, then gift wrapping with certified
// stops), each search step scanning the block of cells [-GDT_SCAN,GDT_SCAN]^2
// around the point and, when that does not certify the step, the rest of the
// step's circle bounding box while it stays inside the window.  A
// displacement is d = X(c) - X(v).
//
// The filtered path certifies against the FP64 displacement of the same
// frame formula.  Its FP32 frame coordinate is the FP64 one rounded to nearest,
// so per coordinate (u = 2^-24)
//   |X32 - X64| <= u|X64|   =>   |d32 - d64| <= u(|Xk| + |Xv|) + u|d| <= 2u|d| + 2u|Xv|,
// i.e. eps1 = 2.5u (relative) and eps0 = 2.5u|Xv|_inf (absolute), with margin.
// The orientation, in-circle and distance bounds are the ones of gdt_point.h
// written for general (eps1, eps0), doubled for slack.
#pragma once

#include "gdt_point.h"

#ifndef GDT_SCAN0
#define GDT_SCAN0 2         // half-width (cells) of the block every search step scans first
#endif
#define GDT_SCAN 3          // half-width of the block scanned when the first finds nothing
#ifndef GDT_SCAN_NN
#define GDT_SCAN_NN 1       // half-width of the nearest-neighbour search's first block
#endif
#ifndef GDT_HALO
#define GDT_HALO 4          // window halo (cells): reach of the circle-box extension
#endif
#define GDT_WIN (GDT_TILE + 2 * GDT_HALO)
#define GDT_WIN_CELLS (GDT_WIN * GDT_WIN)
#ifndef GDT_WIN_MAXPTS
#define GDT_WIN_MAXPTS 2048 // LDS entries per window (a fuller window defers its tile)
#endif
#define GDT_FRAME (GDT_WIN / 2)
static_assert(GDT_SCAN_NN <= GDT_SCAN0 && GDT_SCAN0 <= 2 && GDT_SCAN0 <= GDT_SCAN &&
                  GDT_SCAN < GDT_HALO,
              "first blocks fit a 5-row table; the scan blocks lie in the window");

struct GdtTileWin {
    int tx, ty;          // tile coordinates
    int gx0, gy0;        // unwrapped global cell of window cell (0,0)
};

struct GdtTileLds {
    uint32_t* cs;        // [GDT_WIN_CELLS+1] first LDS entry of each window cell
    int* gstart;         // [GDT_WIN_CELLS] first global slot of each window cell (or null)
    gdt_d2* X64;         // [GDT_WIN_MAXPTS + 1] FP64 frame coordinates
    gdt_f2* X32;         // [GDT_WIN_MAXPTS + 1] FP32 frame coordinates (filtered path)
    int* orig;           // [GDT_WIN_MAXPTS] caller index
};

GDT_HD inline int gdt_wrapm(int v, int n)
{
    v %= n;
    return v < 0 ? v + n : v;
}

GDT_HD inline GdtTileWin gdt_tile_window(const GdtSet& S, int t)
{
    GdtTileWin w;
    w.tx = t % S.tiles_x;
    w.ty = t / S.tiles_x;
    w.gx0 = w.tx * GDT_TILE - GDT_HALO;
    w.gy0 = w.ty * GDT_TILE - GDT_HALO;
    return w;
}

// Cell key of window cell c (wrapped on the torus).
GDT_HD inline int gdt_tile_key(const GdtSet& S, const GdtTileWin& w, int c)
{
    const int wx = c % GDT_WIN, wy = c / GDT_WIN;
    const int gx = gdt_wrapm(w.gx0 + wx, S.xsize), gy = gdt_wrapm(w.gy0 + wy, S.ysize);
    return S.c0 + gdt_cell_key(S, gx, gy);
}

// Staging, phase 1: global slot range of window cell c.
GDT_HD inline void gdt_tile_count(const GdtSet& S, const GdtTileWin& w, int c,
                                  const uint32_t* cstart, int& gstart, uint32_t& cnt)
{
    const int key = gdt_tile_key(S, w, c);
    const uint32_t b = cstart[key], e = cstart[key + 1];
    gstart = (int)b;
    cnt = e - b;
}

GDT_HD inline double gdt_frame64(int f, double bs, double off) { return (double)f * bs + off; }
GDT_HD inline float gdt_frame32(int f, float bs, float off) { return (float)f * bs + off; }

// Staging, phase 2: copy window cell c's points to LDS entries [at, at+cnt).
template <bool F32>
GDT_HD inline void gdt_tile_copy(const GdtSet& S, int c, int gstart, uint32_t at, uint32_t cnt,
                                 const GdtRingIO& io, const GdtTileLds& L)
{
    const int fx = c % GDT_WIN - GDT_FRAME, fy = c / GDT_WIN - GDT_FRAME;
    for (uint32_t k = 0; k < cnt; ++k) {
        const int s = gstart + (int)k;
        const gdt_d2 o = io.off64[s];
        gdt_d2 X;
        X.x = gdt_frame64(fx, S.bsx, o.x);
        X.y = gdt_frame64(fy, S.bsy, o.y);
        L.X64[at + k] = X;
        L.orig[at + k] = io.perm[s];
        if (F32) {
            gdt_f2 Y;
            Y.x = (float)X.x;
            Y.y = (float)X.y;
            L.X32[at + k] = Y;
        }
    }
}

// LDS entries of the block of cells [x0,x1] x [y0,y1] (offsets from window
// cell (wx,wy), inside the window) minus the already scanned square
// [-hole,hole]^2 (hole < 0: none), visited as one flat loop over row segments
// (each row of cells is one contiguous LDS range; a row crossing the hole is
// two segments).  Segments are set up lazily, when the previous one ends, so
// the loop costs the points it visits plus one short set-up per segment.
template <class Fn>
GDT_HD inline unsigned gdt_lds_block(const GdtTileLds& L, int wx, int wy, int x0, int x1, int y0,
                                     int y1, int hole, Fn& fn)
{
    const int xl = x1 < -hole - 1 ? x1 : -hole - 1;     // left segment ends here
    const int xr = x0 > hole + 1 ? x0 : hole + 1;       // right segment starts here
    int y = y0, side = 0, base = (wy + y0) * GDT_WIN + wx;
    uint32_t k = 0, e = 0;
    unsigned n = 0;
    for (;;) {
        while (k == e) {
            // next segment: (row y, side 0 = left or whole row, side 1 = right)
            if (side == 0) {
                const bool split = y >= -hole && y <= hole;
                const int a = x0, z = split ? xl : x1;
                side = split ? 1 : 2;
                if (z >= a) { k = L.cs[base + a]; e = L.cs[base + z + 1]; }
                continue;
            }
            if (side == 1) {
                side = 2;
                if (x1 >= xr) { k = L.cs[base + xr]; e = L.cs[base + x1 + 1]; }
                continue;
            }
            if (++y > y1) return n;
            base += GDT_WIN;
            side = 0;
        }
        fn((int)k);
        ++k;
        ++n;
    }
}

// Columns of the scan block that may hold points strictly left of v->a.  A
// point of cell (dx, dy) has displacement x in [dx bsx - ox, (dx+1) bsx - ox),
// y likewise; it is left when ax y - ay x > 0, i.e. x < (ax/ay) y for ay > 0
// and x > (ax/ay) y for ay < 0.  Over a row of cells the bound is linear in
// dy, so row dy keeps the columns dx <= floor(s dy + c) (mode 1) or
// dx >= floor(s dy + c) (mode 2), with a margin of m cells; mode 0 keeps whole
// rows (nearest-neighbour search, or a nearly horizontal v->a, where the float
// slope would be too coarse).
struct GdtClip {
    float s, c;
    int mode;
};

GDT_HD inline GdtClip gdt_left_clip(bool wrap, float ax, float ay, float ox, float oy, float bsx,
                                    float bsy, float m)
{
    GdtClip cl;
    cl.s = cl.c = 0.f;
    cl.mode = 0;
    if (!wrap || !(fabsf(ay) > 0.01f * fabsf(ax))) return cl;
    const float r = ax * gdt_rcp32(ay);           // |r| < 100
    const float span = r * bsy, ib = gdt_rcp32(bsx);
    cl.s = span * ib;
    if (ay > 0.f) {
        cl.mode = 1;
        cl.c = (fmaf(-r, oy, span > 0.f ? span : 0.f) + ox) * ib + m;
    } else {
        cl.mode = 2;
        cl.c = (fmaf(-r, oy, span < 0.f ? span : 0.f) + ox) * ib - m;
    }
    return cl;
}

// Part sub of W equal parts of the LDS range [k, e) (W lanes sharing a point).
GDT_HD inline void gdt_range_part(uint32_t& k, uint32_t& e, int sub, int W)
{
    if (W > 1) {
        const uint32_t len = e - k;
        e = k + len * (uint32_t)(sub + 1) / (uint32_t)W;
        k = k + len * (uint32_t)sub / (uint32_t)W;
    }
}

// Row table of the block [-h,h]^2 around window cell c0 (h <= GDT_SCAN), each row
// clipped to the columns cl keeps (and cut to part sub of W): one word per
// non-empty row, first LDS entry | end << 16.  Returns the number of entries.
GDT_HD inline unsigned gdt_block_rows(const GdtTileLds& L, int c0, int h, const GdtClip& cl,
                                      uint32_t* tbl, int sub = 0, int W = 1)
{
    unsigned n = 0;
    int ns = 0;
    for (int r = 0; r <= 2 * h; ++r) {
        const int dy = r - h;
        int lo = -h, hi = h;
        if (cl.mode) {
            float X = fmaf(cl.s, (float)dy, cl.c);
            X = GDT_FMINF(GDT_FMAXF(X, -(float)h - 1.f), (float)h + 1.f);
            const int bnd = (int)floorf(X);
            if (cl.mode == 1) hi = bnd < hi ? bnd : hi;
            else lo = bnd > lo ? bnd : lo;
        }
        const int rb = c0 + dy * GDT_WIN;
        if (lo <= hi) {
            uint32_t k = L.cs[rb + lo], e = L.cs[rb + hi + 1];
            gdt_range_part(k, e, sub, W);
            if (e > k) tbl[ns++] = k | (e << 16);
            n += e - k;
        }
    }
    return n;
}

// Extension rows: from row dy (advanced) up to dy1, the columns [dx0, dx1] of
// the circle's cell box, clipped to the columns cl keeps, without the scanned
// square [-hs,hs]^2, as row segments in tbl (at most 5: rows are added while
// two more segments fit).  Rows through the square contribute the parts left
// and right of it.  Segments are cut to part sub of W.  Returns the entries.
// (The box holds the circle's part left of v->a with its margins; scanning
// whole box rows costs a few more entries than the rows' chords would, and
// no square root per row.)
GDT_HD inline unsigned gdt_box_rows(const GdtTileLds& L, int c0, int& dy, int dy1, int dx0,
                                    int dx1, int hs, const GdtClip& cl, uint32_t* tbl,
                                    int sub = 0, int W = 1)
{
    unsigned n = 0;
    int ns = 0;
    for (; dy <= dy1 && ns <= 3; ++dy) {
        int lo = dx0, hi = dx1;
        if (cl.mode) {
            float X = fmaf(cl.s, (float)dy, cl.c);
            X = GDT_FMINF(GDT_FMAXF(X, -1e6f), 1e6f);
            const int bnd = (int)floorf(X);
            if (cl.mode == 1) hi = bnd < hi ? bnd : hi;
            else lo = bnd > lo ? bnd : lo;
        }
        const int rb = c0 + dy * GDT_WIN;
        auto seg = [&](int sa, int sb) {
            if (sa > sb) return;
            uint32_t k = L.cs[rb + sa], e = L.cs[rb + sb + 1];
            gdt_range_part(k, e, sub, W);
            if (e > k) tbl[ns++] = k | (e << 16);
            n += e - k;
        };
        if (dy >= -hs && dy <= hs) {
            seg(lo, hi < -hs - 1 ? hi : -hs - 1);
            seg(lo > hs + 1 ? lo : hs + 1, hi);
        } else {
            seg(lo, hi);
        }
    }
    return n;
}

// The n entries of a row table (non-empty rows only) as one flat loop.  The
// row switch is branch-free: each half-trip reads the word after the current
// row and takes it when the row ends, so every LDS access of the loop is
// unconditional and the coordinates of the next entry, loaded, load(k),
// before the current one is evaluated, fn(k, coords), are waited for one
// half-trip later.  The word after the last row is read but never used (the
// tables are laid out so that it exists), and so is the entry after the last
// row's end, read as the last prefetch: its index is a cell end, at most the
// window's point count, so the coordinate arrays hold GDT_WIN_MAXPTS + 1
// entries.
template <class Load, class Fn>
GDT_HD inline void gdt_scan_tbl(const uint32_t* tbl, unsigned n, Load& load, Fn& fn)
{
    if (n == 0) return;
    const uint32_t* p = tbl;
    // the current entry k and its segment's end e stay packed as in the table
    // (k in the low half, e in the high half; k < e <= 2^16, so ++ke steps k)
    uint32_t ke = p[0];
    uint32_t wn = p[1];
    auto advance = [&](bool more) {
        const bool sw = more & ((ke & 0xffffu) == (ke >> 16));
        ke = sw ? wn : ke;
        p += sw ? 1 : 0;
        wn = p[1];
    };
    auto XA = load((int)(ke & 0xffffu));
    decltype(XA) XB;
    for (unsigned i = 1;; ++i) {
        uint32_t kc = ke & 0xffffu;
        ++ke;
        bool more = i < n;
        advance(more);
        XB = load((int)(ke & 0xffffu));
        fn((int)kc, XA);
        if (!more) break;
        ++i;
        kc = ke & 0xffffu;
        ++ke;
        more = i < n;
        advance(more);
        XA = load((int)(ke & 0xffffu));
        fn((int)kc, XB);
        if (!more) break;
    }
}

// A value the caller knows to be the same in every active lane, made visibly
// uniform (scalar registers on the device).
GDT_HD inline uint32_t gdt_uniform(uint32_t v)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
#else
    return v;
#endif
}

// gdt_scan_tbl for a table that every active lane of the wave holds the same
// (the star kernel's): the same visits in the same order, with the row switch
// and the entry index uniform, so that they cost scalar instructions and
// every lane reads the same coordinates (one LDS broadcast per entry).
template <class Fn>
GDT_HD inline void gdt_scan_tbl_uniform(const uint32_t* tbl, unsigned n, const gdt_d2* X, Fn& fn)
{
    n = gdt_uniform(n);
    if (n == 0) return;
    int r = 0;
    uint32_t t = gdt_uniform(tbl[0]);
    uint32_t k = t & 0xffffu, e = t >> 16;
    unsigned i = 1;     // k is entry i of n
    auto next = [&]() {
        uint32_t kn = k + 1;
        if (kn == e && i < n) {
            t = gdt_uniform(tbl[++r]);
            kn = t & 0xffffu;
            e = t >> 16;
        }
        return kn;      // (after the last entry: its row's end, which exists)
    };
    // two entries a trip, so that the prefetched coordinates need no copy
    gdt_d2 XA = X[k], XB;
    for (;;) {
        uint32_t kn = next();
        XB = X[kn];
        fn((int)k, XA);
        if (i == n) break;
        ++i;
        k = kn;
        kn = next();
        XA = X[kn];
        fn((int)k, XB);
        if (i == n) break;
        ++i;
        k = kn;
    }
}

// Whether the cell block [dx0,dx1] x [dy0,dy1] around window cell c0 lies in
// the staged window and within the grid's search radius (no cell twice).
GDT_HD inline bool gdt_in_window(const GdtSet& S, int c0, int dx0, int dx1, int dy0, int dy1)
{
    const int wx = c0 % GDT_WIN, wy = c0 / GDT_WIN;
    return wx + dx0 >= 0 && wx + dx1 < GDT_WIN && wy + dy0 >= 0 && wy + dy1 < GDT_WIN &&
           dx0 >= -S.rmax && dx1 <= S.rmax && dy0 >= -S.rmax && dy1 <= S.rmax;
}

// ---------------------------------------------------------------------------
// Per point, one search step at a time (the kernel interleaves the points of
// a tile across its lanes step by step).  A step scans the block of cells
// [-h,h]^2 around v's cell, h = GDT_SCAN0, widened to h = GDT_SCAN when that
// block holds no candidate; the loop body evaluates the predicates of the
// lane's current search and takes a single (rare) branch when the candidate
// wins or a filtered predicate is undecided.
//   step 0 (nearest neighbour): certified when its distance is below
//     cov = h*bs_min + (distance of v to its cell walls);
//   step k (successor of edge (v,a)): the final circle (v, a, best) holds no
//     scanned point, so it is certified when its conservative bounding box
//     lies inside the block, or 2R <= cov.
// Otherwise the step scans the rest of the bounding box of its current
// circle (nearest neighbour: of the disk through the current best), provided
// the box lies in the window.  The circles through v and a with centres on
// the left of v->a are nested along the bisector, so every later best's
// circle lies inside the scanned box and the step is then certified.  A box
// leaving the window returns STEP_DEFER: the kernel hands the point to the
// wavefront routine (gdt_wave.h), which rewrites its whole row.  Candidates
// are LDS entries k.
// ---------------------------------------------------------------------------
template <class Real>
struct GdtTileLane {
    int j;            // slot of v
    int iv;           // LDS entry of v
    int c0;           // window cell of v
    Real vx, vy;      // v's window-frame coordinates
    Real covb;        // distance of v to its cell walls
    int ak, n0;       // LDS entries of the current neighbour a and of the first one
    Real ax, ay;      // displacement of a
    int count;        // ring entries emitted
    int* row;
    uint32_t* tbl;    // the lane's row table: 5 words, and a sixth readable one after them
};

// LDS entry of caller index i in window cell c (a cell holds distinct points;
// the search stays inside the cell's range whatever it finds).
GDT_HD inline int gdt_tile_entry(const GdtTileLds& L, int c, int i)
{
    int k = (int)L.cs[c];
    const int last = (int)L.cs[c + 1] - 1;
    while (k < last && L.orig[k] != i) ++k;
    return k;
}

template <class Real>
GDT_HD inline void gdt_tile_lane_init(const GdtTileWin& w, const GdtTileLds& L, int j,
                                      const GdtRingIO& io, const GdtSet& S, GdtTileLane<Real>& st);

template <>
GDT_HD inline void gdt_tile_lane_init<double>(const GdtTileWin& w, const GdtTileLds& L, int j,
                                              const GdtRingIO& io, const GdtSet& S,
                                              GdtTileLane<double>& st)
{
    const uint32_t pc = io.scell[j];
    const int i = io.perm[j];
    const int wxv = (int)(pc & 0xffffu) - w.gx0, wyv = (int)(pc >> 16) - w.gy0;
    st.j = j;
    st.c0 = wyv * GDT_WIN + wxv;
    st.iv = gdt_tile_entry(L, st.c0, i);
    const gdt_d2 X = L.X64[st.iv];
    st.vx = X.x;
    st.vy = X.y;
    const gdt_d2 ov = io.off64[j];
    st.covb = GDT_FMIN(GDT_FMIN(ov.x, S.bsx - ov.x), GDT_FMIN(ov.y, S.bsy - ov.y));
    st.ak = st.n0 = -1;
    st.ax = st.ay = 0.;
    st.count = 0;
    st.row = io.nbrs + (size_t)i * (size_t)io.stride;
}

template <>
GDT_HD inline void gdt_tile_lane_init<float>(const GdtTileWin& w, const GdtTileLds& L, int j,
                                             const GdtRingIO& io, const GdtSet& S,
                                             GdtTileLane<float>& st)
{
    const uint32_t pc = io.scell[j];
    const int i = io.perm[j];
    const int wxv = (int)(pc & 0xffffu) - w.gx0, wyv = (int)(pc >> 16) - w.gy0;
    st.j = j;
    st.c0 = wyv * GDT_WIN + wxv;
    st.iv = gdt_tile_entry(L, st.c0, i);
    const gdt_f2 X = L.X32[st.iv];
    st.vx = X.x;
    st.vy = X.y;
    const gdt_f2 ov = io.off32[j];
    st.covb = GDT_FMINF(GDT_FMINF(ov.x, S.bsx32 - ov.x), GDT_FMINF(ov.y, S.bsy32 - ov.y));
    st.ak = st.n0 = -1;
    st.ax = st.ay = 0.f;
    st.count = 0;
    st.row = io.nbrs + (size_t)i * (size_t)io.stride;
}

// Offset of v inside its cell, recovered from the frame coordinate.
GDT_HD inline double gdt_cell_offset(double X, int wx, double bs)
{
    return X - (double)(wx - GDT_FRAME) * bs;
}

// Result of one step: 0 running, 1 ring closed, 2 defer, 3 non-generic (no close)
enum { STEP_RUN = 0, STEP_DONE = 1, STEP_DEFER = 2, STEP_NOCLOSE = 3 };
// Best entry of an FP64 tile search that ties exactly with the best circle.
#define GDT_TIE (-2)

// b of any of the W lanes that share a point (all W lanes must call it).
template <int W>
GDT_HD inline bool gdt_group_any(bool b)
{
#if defined(__HIPCC__)
    if constexpr (W > 1) {
#pragma unroll
        for (int o = 1; o < W; o <<= 1) b |= __shfl_xor((int)b, o, 64) != 0;
    }
#endif
    return b;
}

// Emit the step's neighbour and advance the lane (shared tail of both paths).
template <class Real>
GDT_HD inline int gdt_tile_advance(const GdtTileLds& L, const GdtRingIO& io, GdtTileLane<Real>& st,
                                   bool wrap, int bk, Real bx, Real by)
{
    if (wrap && bk == st.n0) return STEP_DONE;
    if (!wrap) st.n0 = bk;
    if (st.count < io.stride) st.row[st.count] = L.orig[bk];
    ++st.count;
    if (st.count > GDT_HARD_CAP) return STEP_NOCLOSE;
    st.ak = bk; st.ax = bx; st.ay = by;
    return STEP_RUN;
}

// The search of one FP64 step: the next neighbour bk after st.ak (st.ak < 0:
// the nearest neighbour) and its displacement (bx, by), certified
// (STEP_RUN), or STEP_DEFER.  ext = false gives up (STEP_DEFER) where the
// search would scan beyond its first block or the rest of a circle's box.
// The lane state is only read (its row table is written).  With `pre` (a row
// table of the whole block [-ph, ph]^2, ph = GDT_SCAN0 or GDT_SCAN, npre
// entries, the same in every active lane of the wave) both searches scan that
// block first, unclipped, and certify
// against it: the nearest neighbour is the same, and the successor is the
// same best of the same left candidates (the clip only drops points that are
// not strictly left of v->a).
// W lanes may share the point (W = 2 or 4, device only): each scans its part
// of every row segment, and the group then agrees on the best of their bests,
// so all end every step in the same state.
template <int W = 1>
GDT_HD inline int gdt_tile_search_f64(const GdtSet& S, const GdtTileLds& L,
                                      const GdtTileLane<double>& st, unsigned& npred, int sub,
                                      bool ext, int& bk, double& bx, double& by,
                                      const uint32_t* pre = nullptr, unsigned npre = 0,
                                      int ph = GDT_SCAN0)
{
    static_assert(W == 1 || W == 2 || W == 4, "1, 2 or 4 lanes per point");
    const bool wrap = st.ak >= 0;
    bk = -1;
    const double ax = st.ax, ay = st.ay, vx = st.vx, vy = st.vy;
    const int wx = st.c0 % GDT_WIN, wy = st.c0 / GDT_WIN;
    // Both searches as one test, c taken when O(c) > Othr, P(c) < Pthr and
    // c is not the excluded entry:
    //   successor (a != 0): Othr = 0 (strictly left; v itself has O = 0),
    //     P = in-circle value of (v, a, best), Pthr = 0, a excluded; before a
    //     best exists K = (-1,0,0), so that P = -|c|^2 < 0 and the first left
    //     candidate is taken;
    //   nearest neighbour (a = 0, so O = 0): Othr = -1, v excluded,
    //     K = (1,0,0), so that P = |c|^2, and Pthr = the best squared distance.
    // A new best c sets K = circle (v, a, c) -- K0 = O(c) -- and Pthr = 0 in
    // the successor search; with a = 0 the same expressions give K = (0,0,0),
    // so K0 = O + nnf and Pthr = nnf |c|^2 (nnf = 1 for the nearest neighbour)
    // serve both without a select.
    const int ex = wrap ? st.ak : st.iv;
    const double nnf = wrap ? 0. : 1.;
    const double Othr = wrap ? 0. : -1.;
    const double aa = gdt_dist2_64(ax, ay);
    double Pthr = wrap ? 0. : 1e300;
    Circ64 K;
    K.k0 = wrap ? -1. : 1.;
    K.kx = K.ky = 0.;
    auto load = [&](int k) { return L.X64[k]; };
    // An exact tie with the current best (P == Pthr: degenerate input) sets
    // bk = GDT_TIE and Pthr = -1e300, so nothing more is taken and the step
    // hands the point to the wavefront routine, which decides ties by the
    // symbolic perturbation of gdt_incircle_tie / gdt_nn_tie.  No other state:
    // the test is the branch's own comparison.
    auto visit = [&](int k, const gdt_d2 X) {
        const double px = X.x - vx, py = X.y - vy;
        const double cc = gdt_dist2_64(px, py);
        const double O = gdt_orient64(ax, ay, px, py);
        const double P = gdt_incircle_cc64(K, px, py, cc);
        // one branch, taken only for a new best (or an exact tie)
        if ((O > Othr) & (P <= Pthr) & (k != ex)) {
            const bool t = P == Pthr;
            bk = t ? GDT_TIE : k;
            K.k0 = O + nnf;                        // gdt_circ64(a, c).k0, or 1
            K.kx = fma(cc, ay, -(aa * py));        // gdt_circ64(a, c).kx, or 0
            K.ky = fma(aa, px, -(cc * ax));        // gdt_circ64(a, c).ky, or 0
            Pthr = t ? -1e300 : nnf * cc;
        }
    };
    auto visit_k = [&](int k) { visit(k, load(k)); };
    // Lane groups, pairwise by lane distance o = 1, 2, ...: the lower lane of
    // each pair keeps its partner's best when it lies strictly inside its own
    // circle (nearest neighbour: strictly closer), both take that decision and
    // recompute K and Pthr with the expressions of the visit, so the group's
    // states end identical.  A tie in either lane, or between them, makes the
    // group's best GDT_TIE for the rest of the step (Pthr = -1e300 takes
    // nothing more), so the point is deferred.
    auto pair_reduce = [&]() {
#if defined(__HIPCC__)
        if constexpr (W > 1) {
#pragma unroll
            for (int o = 1; o < W; o <<= 1) {
                const int pk = __shfl_xor(bk, o, 64);
                int win = bk;
                if (bk == GDT_TIE || pk == GDT_TIE) {
                    win = GDT_TIE;
                } else if (pk >= 0 && pk != bk) {
                    if (bk < 0) {
                        win = pk;
                    } else {
                        const gdt_d2 X = L.X64[pk];
                        const double px = X.x - vx, py = X.y - vy;
                        const double P = gdt_incircle_cc64(K, px, py, gdt_dist2_64(px, py));
                        if (P < Pthr) win = pk;
                        else if (P == Pthr) win = GDT_TIE;
                    }
                }
                win = __shfl(win, (int)(__lane_id() & ~(unsigned)(2 * o - 1)), 64);
                if (win == GDT_TIE) {
                    bk = GDT_TIE;
                    Pthr = -1e300;
                } else if (win != bk) {
                    bk = win;
                    const gdt_d2 X = L.X64[bk];
                    const double px = X.x - vx, py = X.y - vy;
                    const double cc = gdt_dist2_64(px, py);
                    K.k0 = gdt_orient64(ax, ay, px, py) + nnf;
                    K.kx = fma(cc, ay, -(aa * py));
                    K.ky = fma(aa, px, -(cc * ax));
                    Pthr = nnf * cc;
                }
            }
        }
#endif
    };
    // the best's displacement and squared distance (the loop keeps only its entry)
    double bd2 = 0.;
    bx = by = 0.;
    auto best = [&]() {
        if (bk >= 0) {
            const gdt_d2 X = L.X64[bk];
            bx = X.x - vx;
            by = X.y - vy;
            bd2 = gdt_dist2_64(bx, by);
        }
    };
    // nearest neighbour: the 3x3 block; successor: the 5x5 block, clipped to
    // the columns that may hold points left of v->a
    int hs = pre ? ph : (wrap ? GDT_SCAN0 : GDT_SCAN_NN);
    const double ox = gdt_cell_offset(vx, wx, S.bsx), oy = gdt_cell_offset(vy, wy, S.bsy);
    const GdtClip cl = gdt_left_clip(wrap, (float)ax, (float)ay, (float)ox, (float)oy,
                                     (float)S.bsx, (float)S.bsy, 1e-3f);
    unsigned nb;
    if (pre) {
        nb = npre;
        gdt_scan_tbl_uniform(pre, nb, L.X64, visit);
    } else {
        nb = gdt_block_rows(L, st.c0, hs, cl, st.tbl, sub, W);
        gdt_scan_tbl(st.tbl, nb, load, visit);
    }
    npred += nb;
    pair_reduce();
    if (bk < 0 && hs < GDT_SCAN) {
        if (!ext) return STEP_DEFER;
        npred += gdt_lds_block(L, wx, wy, -GDT_SCAN, GDT_SCAN, -GDT_SCAN, GDT_SCAN, hs, visit_k);
        hs = GDT_SCAN;
    }

    bool ok = false, boxed = false;
    int dx0 = 0, dx1 = 0, dy0 = 0, dy1 = 0;
    // Certification against the scanned block [-hs,hs]^2 (sets the box too),
    // one code path for both searches: the circle is centred at
    // q = -(Kx,Ky)/(2 K0) -- the origin for the nearest neighbour, whose K is
    // (1,0,0) -- with R^2 = |q|^2 (successor) or the best distance (nearest
    // neighbour); certified when its box lies in the block, or when all of it
    // is within cov of v: 2R <= cov (successor), R <= cov (nearest neighbour).
    auto certify = [&]() {
        best();
        const double cov = (double)hs * GDT_FMIN(S.bsx, S.bsy) + st.covb;
        ok = boxed = false;
        if (bk >= 0 && K.k0 > 1e-6 * (fabs(ax * by) + fabs(ay * bx))) {
            const double i2k = gdt_rcp64(2. * K.k0);
            const double qx = -K.kx * i2k, qy = -K.ky * i2k;
            const double R2 = wrap ? qx * qx + qy * qy : bd2;
            const double rad = gdt_sqrt64(R2) * (1.0 + 1e-9);
            const double m = 1e-6 * (fabs(qx) + fabs(qy) + rad) + 1e-9 * S.bsx;
            gdt_circle_cells<double>(qx - m, qx + m, qy - m, qy + m, rad, ox, oy, S.ibsx, S.ibsy,
                                     1e-3, dx0, dx1, dy0, dy1);
            ok = (dx0 >= -hs && dx1 <= hs && dy0 >= -hs && dy1 <= hs) ||
                 (cov > 0. && (wrap ? 4. : 1.) * R2 <= cov * cov * (1.0 - 1e-6));
            boxed = true;
        }
    };
    certify();
    if (!ok && !ext) return STEP_DEFER;
    // a box leaving the window: the full block first (it lies in the window)
    if (!ok && hs < GDT_SCAN && !(boxed && gdt_in_window(S, st.c0, dx0, dx1, dy0, dy1))) {
        npred += gdt_lds_block(L, wx, wy, -GDT_SCAN, GDT_SCAN, -GDT_SCAN, GDT_SCAN, hs, visit_k);
        hs = GDT_SCAN;
        certify();
    }
    if (!ok && boxed && gdt_in_window(S, st.c0, dx0, dx1, dy0, dy1)) {
        // the rest of the circle's cell box, left of v->a, outside the scanned
        // square (every later best's circle lies inside)
        for (int dy = dy0; dy <= dy1;) {
            const unsigned ne = gdt_box_rows(L, st.c0, dy, dy1, dx0, dx1, hs, cl, st.tbl, sub, W);
            gdt_scan_tbl(st.tbl, ne, load, visit);
            npred += ne;
        }
        pair_reduce();
        ok = true;
    }
    if (!ok) return STEP_DEFER;     // (a best of GDT_TIE is never certified)
    best();
    return STEP_RUN;
}

template <int W = 1>
GDT_HD inline int gdt_tile_step_f64(const GdtSet& S, const GdtTileLds& L, const GdtRingIO& io,
                                    GdtTileLane<double>& st, unsigned& npred, int sub = 0)
{
    int bk;
    double bx, by;
    const int rc = gdt_tile_search_f64<W>(S, L, st, npred, sub, true, bk, bx, by);
    if (rc != STEP_RUN) return rc;
    return gdt_tile_advance<double>(L, io, st, st.ak >= 0, bk, bx, by);
}

// W = 2: two neighbouring lanes share the point, as in the FP64 step.
template <int W = 1>
GDT_HD inline int gdt_tile_step_f32(const GdtSet& S, const GdtTileLds& L, const GdtRingIO& io,
                                    GdtTileLane<float>& st, unsigned& npred, unsigned& nfb,
                                    int sub = 0)
{
    static_assert(W == 1 || W == 2, "one or two lanes per point");
    const float u = GDT_U32;
    const float bsM = S.bsmax32;
    const float eps1 = 2.5f * u;
    const float eps0 = 2.5f * u * (GDT_FMAXF(fabsf(st.vx), fabsf(st.vy)) + 1e-6f * bsM);
    const float cP = 8.f * (4.f * eps1 + 8.f * u), eP = 8.f * eps0;
    const bool wrap = st.ak >= 0;
    const float ax = st.ax, ay = st.ay, vx = st.vx, vy = st.vy;
    const int iv = st.iv, ak = st.ak;
    const float ma = GDT_FMAXF(fabsf(ax), fabsf(ay));
    const float cE = 2.f * (2.f * eps1 + 3.f * u);
    int bk = -1;
    float bd2 = 0.f, mbb = 0.f, BP = 0.f;
    // Both searches as one test on (O, P), as in the FP64 path (nearest
    // neighbour: a = 0 so O = 0, K = (1,0,0) so P = |c|^2; successor before a
    // best exists: K = (-1,0,0) so P = -|c|^2; a new best sets K0 = O + nnf).
    // A candidate is certainly a new best when O > Oc and P < Pc, possibly one
    // (the FP64 expression of the FP64 path decides) when O >= Op and P <= Pp,
    // and certainly not otherwise.  v and a are excluded by entry: v's O = 0
    // lies inside the orientation bound, so it would reach the FP64 decision.
    const float nnf = wrap ? 0.f : 1.f;
    const float aa = fmaf(ax, ax, ay * ay);
    float K0 = wrap ? -1.f : 1.f, Kx = 0.f, Ky = 0.f;
    float Oc = -1.f, Op = -1.f, Pc = wrap ? 0.f : 3.0e38f, Pp = Pc;
    // Bounds that depend on mc, a bound on the max-norm of the displacements
    // of the region being scanned: (h+1) cells for offsets |dx|,|dy| <= h.
    float mc = 0.f, tA = 0.f, qa = 0.f, BO = 0.f, eabs = 0.f, e2 = 0.f;
    // nearest neighbour: |c|^2 + E(c) < bd2 - E(b) certain, |c|^2 - E(c) >
    // bd2 + E(b) certainly not, with E(x) = cE |x|^2 + eabs (eabs of the
    // current region, which bounds that of the region b was found in).  As
    // thresholds on |c|^2, each one rounding away from the exact one:
    //   Pc = bd2 (1 - 3cE) - e2 <= (bd2 (1 - cE) - 2 eabs) / (1 + cE),
    //   Pp = bd2 (1 + 3cE) + e2 >= (bd2 (1 + cE) + 2 eabs) / (1 - cE),
    // with e2 = 2.0001 eabs; 1 +- 3cE are exact in FP32 (cE = 16u).
    const float nn_c = 1.f - 3.f * cE, nn_p = 1.f + 3.f * cE;
    auto set_nn = [&]() {
        Pc = fmaf(bd2, nn_c, -e2);
        Pp = fmaf(bd2, nn_p, e2);
    };
    // in-circle bound of the current best b, mb = |b|_inf (with sab = ma + mb,
    // pab = ma mb):  BP = eP (sab + mc)^3 + cP pab mc (sab + mc)
    //                   = t (eP t^2 + (cP ma mc) mb),  t = (ma + mc) + mb
    auto set_bp = [&](float mb) {
        const float t = tA + mb;
        BP = fmaf(t, fmaf(qa, mb, (eP * t) * t), 1e-30f);
        Pc = -BP;
        Pp = BP;
    };
    auto set_region = [&](int h) {
        mc = (float)(h + 1) * bsM * 1.0001f;
        tA = ma + mc;
        qa = cP * mc * ma;
        BO = fmaf(4.f * (2.f * eps1 + 2.f * u) * ma + 4.f * eps0, mc, 4.f * eps0 * ma + 1e-30f);
        eabs = 8.f * eps0 * mc + 4.f * eps0 * eps0 + 1e-30f;
        e2 = 2.0001f * eabs;
        if (wrap) {
            Oc = BO;
            Op = -BO;
            if (bk >= 0) set_bp(mbb);
        } else if (bk >= 0) {
            set_nn();
        }
    };
    int hs = wrap ? GDT_SCAN0 : GDT_SCAN_NN;
    set_region(hs);
    const int wx = st.c0 % GDT_WIN, wy = st.c0 / GDT_WIN;
    // FP64 displacement of LDS entry k: the FP64 path's value (fallbacks only)
    auto x64 = [&](int k, double& x, double& y) {
        const gdt_d2 V = L.X64[iv], X = L.X64[k];
        x = X.x - V.x;
        y = X.y - V.y;
    };
    auto load = [&](int k) { return L.X32[k]; };
    // an exact tie met in the FP64 decision (the FP64 path's GDT_TIE): the
    // point is deferred at the end of the step
    bool tie = false;
#ifdef GDT_AUDIT_FILTER
    auto audit = [&](int kind, int sign, int k) {
        double a64x = 0., a64y = 0., b64x = 0., b64y = 0., c64x, c64y;
        if (ak >= 0) x64(ak, a64x, a64y);
        if (bk >= 0) x64(bk, b64x, b64y);
        x64(k, c64x, c64y);
        GDT_FILTER_AUDIT(kind, sign, a64x, a64y, b64x, b64y, c64x, c64y);
    };
    // the claims of an FP32-only decision: the orientation (successor), and
    // against a best the in-circle value (successor) or the distance order
    auto audit_fp32 = [&](int k, float O, float P, bool taken) {
        if (taken) {
            if (wrap) audit(GDT_AUD_ORIENT, 1, k);
            if (bk >= 0) audit(wrap ? GDT_AUD_INCIRCLE : GDT_AUD_DIST, -1, k);
        } else if (wrap && O < Op) {
            audit(GDT_AUD_ORIENT, -1, k);
        } else if (bk >= 0 && P > Pp) {
            audit(wrap ? GDT_AUD_INCIRCLE : GDT_AUD_DIST, 1, k);
        }
    };
#endif
    auto visit = [&](int k, const gdt_f2 Y) {
        const float px = Y.x - vx, py = Y.y - vy;
        const float cc = fmaf(px, px, py * py);
        const float O = fmaf(ax, py, -(ay * px));
        const float P = fmaf(K0, cc, fmaf(Kx, px, Ky * py));
        // one branch, taken only when the candidate may be a new best
        if (!GDT_FILTER_DECIDES(!((O >= Op) & (P <= Pp))) & (k != iv) & (k != ak)) {
            // (the nearest-neighbour search's first candidate is always
            // taken: there is no best to compare it with; the FP32 test
            // takes it, a build without FP32 verdicts says so)
            bool take = GDT_FILTER_FORCED(!wrap & (bk < 0)) |
                        GDT_FILTER_DECIDES((O > Oc) & (P < Pc));
#ifdef GDT_AUDIT_FILTER
            if (take) audit_fp32(k, O, P, true);
#endif
            if (!take) {                      // undecided: the FP64 path's decision
                ++nfb;
                double c64x, c64y;
                x64(k, c64x, c64y);
                if (wrap) {
                    double a64x, a64y;
                    x64(ak, a64x, a64y);
#ifdef GDT_AUDIT_FILTER
                    if (fabsf(O) > BO) audit(GDT_AUD_ORIENT, O > BO ? 1 : -1, k);
                    if ((fabsf(O) > BO ? O > BO : gdt_orient64(a64x, a64y, c64x, c64y) > 0.) &&
                        bk >= 0 && fabsf(P) > BP)
                        audit(GDT_AUD_INCIRCLE, P < -BP ? -1 : 1, k);
#endif
                    const bool left = GDT_FILTER_DECIDES(fabsf(O) > BO)
                                          ? O > BO
                                          : gdt_orient64(a64x, a64y, c64x, c64y) > 0.;
                    if (!left) take = false;
                    else if (bk < 0) take = true;
                    else if (GDT_FILTER_DECIDES(fabsf(P) > BP)) take = P < -BP;
                    else {
                        double b64x, b64y;
                        x64(bk, b64x, b64y);
                        const Circ64 K64 = gdt_circ64(a64x, a64y, b64x, b64y);
                        const double P64 = gdt_incircle64(K64, c64x, c64y);
                        take = P64 < 0.;
                        tie |= P64 == 0.;
                    }
                } else {
                    double b64x, b64y;
                    x64(bk, b64x, b64y);
                    const double dc = gdt_dist2_64(c64x, c64y), db = gdt_dist2_64(b64x, b64y);
                    take = dc < db;
                    tie |= dc == db;
                }
            }
            if (take) {
                bk = k;
                bd2 = cc;
                K0 = O + nnf;                        // fmaf(ax, py, -(ay px)), or 1
                Kx = fmaf(cc, ay, -(aa * py));       // 0 when a = 0
                Ky = fmaf(aa, px, -(cc * ax));       // 0 when a = 0
                if (wrap) {
                    mbb = GDT_FMAXF(fabsf(px), fabsf(py));
                    set_bp(mbb);
                } else {
                    set_nn();
                }
            }
        }
#ifdef GDT_AUDIT_FILTER
        else if ((k != iv) & (k != ak)) {
            audit_fp32(k, O, P, false);
        }
#endif
    };
    auto visit_k = [&](int k) { visit(k, load(k)); };
    // Pairs: the lower lane visits its partner's best (the filtered test, with
    // the FP64 decision when undecided, as for any candidate); the upper lane
    // then restarts from no best and visits the winner, which sets its state
    // with the same expressions, so both end identical.
    auto pair_reduce = [&]() {
#if defined(__HIPCC__)
        if constexpr (W > 1) {
            const int pk = __shfl_xor(bk, 1, 64);
            if (sub == 0 && pk >= 0 && pk != bk) visit(pk, load(pk));
            const int win = __shfl(bk, (int)(__lane_id() & ~1u), 64);
            if (sub != 0 && win != bk) {
                bk = -1;
                bd2 = mbb = BP = 0.f;
                K0 = wrap ? -1.f : 1.f;
                Kx = Ky = 0.f;
                Pc = wrap ? 0.f : 3.0e38f;
                Pp = Pc;
                if (win >= 0) visit(win, load(win));
            }
        }
#endif
    };
    // nearest neighbour: the 3x3 block; successor: the 5x5 block, clipped to
    // the columns that may hold points with O >= -BO (horizontal margin BO/|ay|)
    const float ox = vx - (float)(wx - GDT_FRAME) * S.bsx32;
    const float oy = vy - (float)(wy - GDT_FRAME) * S.bsy32;
    {
        const float clm = wrap ? BO * gdt_rcp32(fabsf(ay) * S.bsx32) + 1e-3f : 0.f;
        const GdtClip cl = gdt_left_clip(wrap, ax, ay, ox, oy, S.bsx32, S.bsy32, clm);
        const unsigned nb = gdt_block_rows(L, st.c0, hs, cl, st.tbl, sub, W);
        gdt_scan_tbl(st.tbl, nb, load, visit);
        npred += nb;
        pair_reduce();
    }
    if (bk < 0 && hs < GDT_SCAN) {
        set_region(GDT_SCAN);
        npred += gdt_lds_block(L, wx, wy, -GDT_SCAN, GDT_SCAN, -GDT_SCAN, GDT_SCAN, hs, visit_k);
        hs = GDT_SCAN;
    }

    const float bsm = S.bsmin32;
    bool ok = false, boxed = false;
    int dx0 = 0, dx1 = 0, dy0 = 0, dy1 = 0;
    // certification against the scanned block [-hs,hs]^2 (sets the box too)
    auto certify = [&]() {
        const float cov = (float)hs * bsm + st.covb - 9.5367431640625e-07f * (float)(hs + 2) * bsM;
        ok = boxed = false;
        // one code path for both searches, as in the FP64 path: for the
        // nearest neighbour a = 0 and K = (1,0,0) exactly, so the centre
        // bounds are the origin and R^2 is the best distance plus its error
        const float mb = wrap ? mbb : 0.f;
        const float sab = ma + mb, pab = ma * mb;
        const float EK0 = 4.f * (2.f * eps1 + 2.f * u) * pab + 4.f * eps0 * sab;
        const float EK = 4.f * (3.f * eps1 + 6.f * u) * pab * sab + 8.f * eps0 * sab * sab;
        const float Klo = K0 - EK0, Khi = K0 + EK0;
        if (bk >= 0 && Klo > 0.0625f * K0) {
            const float nxlo = -(Kx + EK), nxhi = -(Kx - EK);
            const float nylo = -(Ky + EK), nyhi = -(Ky - EK);
            const float il = gdt_rcp32(2.f * Klo), ih = gdt_rcp32(2.f * Khi);
            const float qxlo = GDT_FMINF(nxlo * il, nxlo * ih);
            const float qxhi = GDT_FMAXF(nxhi * il, nxhi * ih);
            const float qylo = GDT_FMINF(nylo * il, nylo * ih);
            const float qyhi = GDT_FMAXF(nyhi * il, nyhi * ih);
            const float bE = fmaf(cE, bd2, eabs);
            const float R2 = wrap ? GDT_FMAXF(qxlo * qxlo, qxhi * qxhi) +
                                        GDT_FMAXF(qylo * qylo, qyhi * qyhi)
                                  : bd2 + bE;
            const float rad = gdt_sqrt32(R2) * 1.0001f;
            gdt_circle_cells<float>(qxlo, qxhi, qylo, qyhi, rad, ox, oy, S.ibsx32, S.ibsy32,
                                    1e-2f, dx0, dx1, dy0, dy1);
            const float gx = fabsf(Kx) + EK, gy = fabsf(Ky) + EK;
            const bool within = wrap ? (gx * gx + gy * gy) * 1.001f <= (cov * cov) * (Klo * Klo)
                                     : R2 <= cov * cov * 0.9999f;
            ok = (dx0 >= -hs && dx1 <= hs && dy0 >= -hs && dy1 <= hs) || (cov > 0.f && within);
            boxed = true;
        }
    };
    certify();
    // a box leaving the window: the full block first (it lies in the window)
    if (!ok && hs < GDT_SCAN && !(boxed && gdt_in_window(S, st.c0, dx0, dx1, dy0, dy1))) {
        set_region(GDT_SCAN);
        npred += gdt_lds_block(L, wx, wy, -GDT_SCAN, GDT_SCAN, -GDT_SCAN, GDT_SCAN, hs, visit_k);
        hs = GDT_SCAN;
        certify();
    }
    if (!ok && boxed && gdt_in_window(S, st.c0, dx0, dx1, dy0, dy1)) {
        // the rest of the circle, as in the FP64 path; the clip margin follows
        // the orientation bound of the larger region
        const int h0 = dx0 < dy0 ? dx0 : dy0, h1 = dx1 > dy1 ? dx1 : dy1;
        set_region(-h0 > h1 ? -h0 : h1);
        const GdtClip clx = gdt_left_clip(wrap, ax, ay, ox, oy, S.bsx32, S.bsy32,
                                          wrap ? BO * gdt_rcp32(fabsf(ay) * S.bsx32) + 1e-3f : 0.f);
        for (int dy = dy0; dy <= dy1;) {
            const unsigned ne = gdt_box_rows(L, st.c0, dy, dy1, dx0, dx1, hs, clx, st.tbl, sub, W);
            gdt_scan_tbl(st.tbl, ne, load, visit);
            npred += ne;
        }
        pair_reduce();
        ok = true;
    }
    if (!ok || gdt_group_any<W>(tie)) return STEP_DEFER;
    // the best's displacement (the loop keeps only its entry)
    const gdt_f2 B = L.X32[bk];
    return gdt_tile_advance<float>(L, io, st, wrap, bk, B.x - vx, B.y - vy);
}

// Whole point, one step after another (host emulation and simple callers).
template <int PREC>
GDT_HD inline GdtRingResult gdt_tile_ring(const GdtSet& S, const GdtTileWin& w, const GdtTileLds& L,
                                          int j, const GdtRingIO& io)
{
    GdtRingResult res = {0, 0, 0u, 0u};
    int rc = STEP_RUN;
    uint32_t tbl[6];
    if (PREC == 32) {
        GdtTileLane<float> st;
        gdt_tile_lane_init<float>(w, L, j, io, S, st);
        st.tbl = tbl;
        while (rc == STEP_RUN) rc = gdt_tile_step_f32(S, L, io, st, res.npred, res.nfb);
        res.count = st.count;
    } else {
        GdtTileLane<double> st;
        gdt_tile_lane_init<double>(w, L, j, io, S, st);
        st.tbl = tbl;
        while (rc == STEP_RUN) rc = gdt_tile_step_f64<1>(S, L, io, st, res.npred);
        res.count = st.count;
    }
    if (rc == STEP_DEFER) res.err = ERR_DEFER;
    if (rc == STEP_NOCLOSE) res.err = ERR_NOCLOSE;
    return res;
}
