/*
 * gdt_oracle.c -- CPU oracle (FP64, OpenMP) for the localized periodic
 * Delaunay path.  TEST INFRASTRUCTURE ONLY: see gdt_oracle.h.
 *
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off), which matches
 * the reference's host build flags (CMakeLists.txt:11) with contraction
 * pinned off so FP64 rounding does not depend on the compiler.
 */
#include "gdt_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { double x, y; } d2;

/* Internal polygon scratch per point.  The reference sizes this by MaxSize
 * and overruns it on the CPU (DelaunayGPU.cu:911-918); the oracle keeps a
 * fixed large scratch and reports the final ring size instead. */
#define ORC_SCRATCH 256

/* THRESHOLD / EPSILON (inc/std_include.h:16-17). */
#define ORC_EPS 1e-18

/* ------------------------------------------------------------------ */
/* Box and grid                                                        */
/* ------------------------------------------------------------------ */

/* cellListGPU::setGridSize (cellListGPU.cpp:97-117): the cell count per
 * axis is floor(T/a), bumped to the next even number (T: box side, or box
 * thickness for a general box). */
static void grid_counts(double tx, double ty, double cell_size, orc_grid* g)
{
    g->xsize = (int)floor(tx / cell_size);
    if (g->xsize % 2 == 1) g->xsize += 1;
    g->ysize = (int)floor(ty / cell_size);
    if (g->ysize % 2 == 1) g->ysize += 1;
    if (g->xsize < 2) g->xsize = 2;
    if (g->ysize < 2) g->ysize = 2;
}

void orc_make_grid(double lx, double ly, double cell_size, orc_grid* g)
{
    memset(g, 0, sizeof(*g));
    g->general = 0;
    g->x11 = lx; g->x22 = ly;                      /* setSquare (periodicBoundaries.h:75-84) */
    g->xi11 = 1. / lx; g->xi22 = 1. / ly;
    g->lx = lx; g->ly = ly;
    g->hx = lx * 0.5; g->hy = ly * 0.5;           /* periodicBoundaries.h:82-83 */
    g->ilx = 1. / lx; g->ily = 1. / ly;           /* periodicBoundaries.h:79 */
    grid_counts(lx, ly, cell_size, g);
    g->bsx = lx / g->xsize;
    g->bsy = ly / g->ysize;   /* reference uses lx/xsize here (cellListGPU.cpp:106) */
    g->hcx = g->bsx;
    g->hcy = g->bsy;
    g->lhalf = (g->xsize * g->bsx) * 0.5;
}

/* setGeneral (periodicBoundaries.h:86-101); the cells are the unit-cell grid
 * mapped by the box matrix, parallelograms of heights T/xsize, T/ysize. */
void orc_make_grid_box(const double* box, double cell_size, orc_grid* g)
{
    const double a = box[0], b = box[1], c = box[2], d = box[3];
    if (b == 0. && c == 0. && a > 0. && d > 0.) {
        orc_make_grid(a, d, cell_size, g);
        return;
    }
    memset(g, 0, sizeof(*g));
    g->general = 1;
    g->x11 = a; g->x12 = b; g->x21 = c; g->x22 = d;
    const double det = a * d - b * c, pre = 1.0 / det;
    g->xi11 = pre * d; g->xi12 = -pre * b;
    g->xi21 = -pre * c; g->xi22 = pre * a;
    const double adet = fabs(det);
    const double tx = adet / sqrt(b * b + d * d);   /* distance between the walls along b */
    const double ty = adet / sqrt(a * a + c * c);   /* distance between the walls along a */
    grid_counts(tx, ty, cell_size, g);
    g->hcx = tx / g->xsize;
    g->hcy = ty / g->ysize;
    g->lhalf = 0.5 * (tx < ty ? tx : ty);
}

/* periodicBoundaries::invTrans / Trans (periodicBoundaries.h:103-129). */
static inline d2 inv_trans(const orc_grid* g, d2 p)
{
    d2 u = { g->xi11 * p.x + g->xi12 * p.y, g->xi21 * p.x + g->xi22 * p.y };
    return u;
}
static inline d2 trans(const orc_grid* g, d2 u)
{
    d2 p = { g->x11 * u.x + g->x12 * u.y, g->x21 * u.x + g->x22 * u.y };
    return p;
}

/* periodicBoundaries::minDist (periodicBoundaries.h:154-179): rectangular
 * branch on the coordinates, general branch on unit-cell coordinates. */
static inline d2 min_dist(const orc_grid* g, d2 p1, d2 p2)
{
    if (g->general) {
        const d2 va = inv_trans(g, p1), vb = inv_trans(g, p2);
        d2 u = { va.x - vb.x, va.y - vb.y };
        while (u.x < -0.5) u.x += 1.0;
        while (u.y < -0.5) u.y += 1.0;
        while (u.x > 0.5) u.x -= 1.0;
        while (u.y > 0.5) u.y -= 1.0;
        return trans(g, u);
    }
    d2 r = { p1.x - p2.x, p1.y - p2.y };
    while (r.x < -g->hx) r.x += g->lx;
    while (r.y < -g->hy) r.y += g->ly;
    while (r.x > g->hx) r.x -= g->lx;
    while (r.y > g->hy) r.y -= g->ly;
    return r;
}

/* periodicBoundaries::putInBoxReal (periodicBoundaries.h:131-152): map to
 * unit coordinates, wrap into [0,1), map back. */
static inline d2 put_in_box(const orc_grid* g, d2 p)
{
    if (g->general) {
        d2 u = inv_trans(g, p);
        while (u.x < 0) u.x += 1.0;
        while (u.y < 0) u.y += 1.0;
        while (u.x >= 1.0) u.x -= 1.0;
        while (u.y >= 1.0) u.y -= 1.0;
        return trans(g, u);
    }
    double ux = g->ilx * p.x, uy = g->ily * p.y;
    while (ux < 0) ux += 1.0;
    while (uy < 0) uy += 1.0;
    while (ux >= 1.0) ux -= 1.0;
    while (uy >= 1.0) uy -= 1.0;
    d2 r = { g->lx * ux, g->ly * uy };
    return r;
}

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* Cell of a point inside the box: rectangles by side length
 * (cellListGPU.cpp:314-316), general boxes by unit-cell coordinates. */
static inline void cell_of(const orc_grid* g, d2 p, int* cx, int* cy)
{
    if (g->general) {
        const d2 u = inv_trans(g, p);
        *cx = clampi((int)floor(u.x * g->xsize), 0, g->xsize - 1);
        *cy = clampi((int)floor(u.y * g->ysize), 0, g->ysize - 1);
        return;
    }
    *cx = clampi((int)floor(p.x / g->bsx), 0, g->xsize - 1);
    *cy = clampi((int)floor(p.y / g->bsy), 0, g->ysize - 1);
}

/* positiveModulo (DelaunayGPU.cu:26-31). */
static inline int pmod(int i, int n)
{
    int m = i % n;
    if (m < 0) m += n;
    return m;
}

/* cellListGPU::compute(points) (cellListGPU.cpp:289-330): one pass over the
 * points in index order, appending each to its bucket.  Stored as CSR. */
int orc_cell_list(const double* xy, int n, const orc_grid* g, int* cell_start, int* cell_idx)
{
    int ncell = g->xsize * g->ysize;
    int* fill = (int*)calloc((size_t)ncell, sizeof(int));
    int* key = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    if (!fill || !key) { free(fill); free(key); return -1; }
    memset(cell_start, 0, sizeof(int) * (size_t)(ncell + 1));
    for (int i = 0; i < n; ++i) {
        int cx, cy;
        d2 p = { xy[2 * i], xy[2 * i + 1] };
        cell_of(g, p, &cx, &cy);
        key[i] = cy * g->xsize + cx;            /* Index2D(xsize,ysize)(cx,cy), indexer.h:38-41 */
        cell_start[key[i] + 1]++;
    }
    for (int c = 0; c < ncell; ++c) cell_start[c + 1] += cell_start[c];
    for (int i = 0; i < n; ++i) {
        int c = key[i];
        cell_idx[cell_start[c] + fill[c]++] = i;
    }
    free(fill);
    free(key);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Geometry helpers (inc/functions.h, DelaunayGPU.cu:18-25)            */
/* ------------------------------------------------------------------ */

/* Circumcenter of the circle through the origin, a and b (functions.h:33-42). */
static inline d2 circumcenter0(d2 a, d2 b)
{
    double an = a.x * a.x + a.y * a.y;
    double bn = b.x * b.x + b.y * b.y;
    double s = 0.5 / (a.x * b.y - a.y * b.x);
    d2 c = { s * (an * b.y - a.y * bn), s * (a.x * bn - an * b.x) };
    return c;
}

/* checkCW (DelaunayGPU.cu:22-25): 0 if (a,b,c) turns counter-clockwise. */
static inline int side_cw(double ax, double ay, double bx, double by, double cx, double cy)
{
    return ((ax - bx) * (ay - cy) - (ay - by) * (ax - cx)) > 0 ? 0 : 1;
}

/* Square spiral over cell offsets (DelaunayGPU.cu:410-427): (0,0),(1,0),
 * (1,-1),(0,-1),(-1,-1),(-1,0),(-1,1),(0,1),(1,1),(2,1),... */
static inline void spiral_next(int* dd, int* cc)
{
    int d = *dd, c = *cc;
    if (abs(d) <= abs(c) && (d != c || d >= 0)) {
        if (c >= 0) d += 1; else d -= 1;
    } else {
        if (d >= 0) c -= 1; else c += 1;
    }
    *dd = d; *cc = c;
}

/* ------------------------------------------------------------------ */
/* One-ring construction                                               */
/* ------------------------------------------------------------------ */

typedef struct {
    const d2* pt;
    const orc_grid* g;
    const int* cstart;
    const int* cidx;
} orc_ctx;

typedef struct {
    int n;                 /* polygon size */
    int idx[ORC_SCRATCH];  /* P_idx: neighbour index, -1 for virtual points */
    d2 P[ORC_SCRATCH];     /* neighbour displacement from v */
    d2 Q[ORC_SCRATCH];     /* Voronoi vertex relative to v: circumcentre of (v, P[m], P[m+1]) */
} ring_t;

enum { CLIP_NONE = 0, CLIP_DONE = 1, CLIP_FULL = 2 };

/* Cut the polygon with the bisector half-plane of v and the candidate at
 * displacement disp.  Restates the removal scan and insertion shared by the
 * two phases: DelaunayGPU.cu:461-589 (phase 1) and 770-926 (phase 2).
 *   wstart   first vertex the forward scan looks at (0 in phase 1, jj in phase 2)
 *   wrap_run whether a cut of the last vertex opens a run that continues at
 *            the scan start (always in phase 1, only for jj == 0 in phase 2) */
static int clip_ring(ring_t* r, int jj, d2 disp, int cand, int wstart, int wrap_run)
{
    /* helper point (xx,yy) on the bisector of the origin and disp (461-476) */
    double xx = 0., yy = 0.;
    if (fabs(disp.y) > ORC_EPS) {
        yy = (disp.y * disp.y + disp.x * disp.x) / (2 * disp.y);
        xx = 0;
    } else if (fabs(disp.y) < ORC_EPS) {
        yy = disp.y / 2 + 1;
        xx = disp.x / 2;
    }
    if (fabs(disp.x) < ORC_EPS) {
        yy = disp.y / 2;
        xx = disp.x / 2 + 1;
    }
    const double hx = 0.5 * disp.x, hy = 0.5 * disp.y;
    const int origin_side = side_cw(hx, hy, xx, yy, 0., 0.);
    /* the vertex being tested must lie on the candidate's side (480-482) */
    if (origin_side == side_cw(hx, hy, xx, yy, r->Q[jj].x, r->Q[jj].y)) return CLIP_NONE;

    int removed = 0, j = -1;
    int in_run = 0, first = 1;
    const int last_side = side_cw(hx, hy, xx, yy, r->Q[r->n - 1].x, r->Q[r->n - 1].y);
    if (last_side != origin_side) {
        j = r->n - 1;
        removed++;
        if (wrap_run) in_run = 1;
    }
    int w;
    for (w = wstart; w < r->n - 1; w++) {
        int s = side_cw(hx, hy, xx, yy, r->Q[w].x, r->Q[w].y);
        if (s != origin_side) {
            if (!in_run) {
                if (j < 0 || j > w) j = w;
                removed++;
                in_run = 1;
            } else {
                if (!first) {
                    /* interior vertex of a removed run: drop entry w */
                    for (int p = w; p < r->n - 1; p++) {
                        r->Q[p] = r->Q[p + 1];
                        r->P[p] = r->P[p + 1];
                        r->idx[p] = r->idx[p + 1];
                    }
                    r->n--;
                    if (j > w) j--;
                    w--;
                } else {
                    first = 0;
                }
                removed++;
            }
        } else {
            in_run = 0;
        }
    }
    if (last_side != origin_side && in_run && !first) {
        r->n--;
        if (j > w) j--;
    }
    if (removed == 0) return CLIP_NONE;

    /* insert the candidate after j with the two new Voronoi vertices (562-584, 864-925) */
    int m = (removed > 1) ? (j + 2) % r->n : (j + 1) % r->n;
    d2 q1 = circumcenter0(r->P[j], disp);
    d2 q2 = circumcenter0(disp, r->P[m]);
    if (removed == 1) {
        if (r->n + 1 > ORC_SCRATCH) return CLIP_FULL;
        r->n++;
        for (int p = r->n - 2; p > j; p--) {
            r->Q[p + 1] = r->Q[p];
            r->P[p + 1] = r->P[p];
            r->idx[p + 1] = r->idx[p];
        }
    }
    m = (j + 1) % r->n;
    r->Q[m] = q2;
    r->P[m] = disp;
    r->idx[m] = cand;
    r->Q[j] = q1;
    return CLIP_DONE;
}

/* Is cand excluded from the test of vertex jj (445-449, 759-763)? */
static inline int skip_candidate(const ring_t* r, int cand, int self, int jj)
{
    if (cand == self || cand == r->idx[0]) return 1;
    for (int p = jj; p < r->n; ++p)
        if (cand == r->idx[p]) return 1;
    return 0;
}

static inline int spiral_half_width(const orc_grid* g, double R)
{
    double bs = g->hcx < g->hcy ? g->hcx : g->hcy;
    int cap = (g->xsize < g->ysize ? g->xsize : g->ysize) / 2;
    int w = (int)ceil(R / bs);
    return w < cap ? w : cap;
}

/* Scan the spiral around cell (cx,cy) for the first candidate that cuts
 * vertex jj.  Returns the clip result. */
static int search_and_clip(const orc_ctx* C, ring_t* r, int self, d2 v, int jj,
                           int cx, int cy, int wstart, int wrap_run)
{
    const orc_grid* g = C->g;
    const d2 cq = r->Q[jj];
    const double R = sqrt(cq.x * cq.x + cq.y * cq.y);
    const double rr = R * R;
    const int side = 2 * spiral_half_width(g, R) + 1;
    int dd = 0, cc = 0;
    for (int s = 0; s < side * side; ++s) {
        int x = pmod(cx + dd, g->xsize);
        int y = pmod(cy + cc, g->ysize);
        spiral_next(&dd, &cc);
        int bin = y * g->xsize + x;
        for (int a = C->cstart[bin]; a < C->cstart[bin + 1]; ++a) {
            int cand = C->cidx[a];
            if (skip_candidate(r, cand, self, jj)) continue;
            d2 disp = min_dist(g, C->pt[cand], v);
            d2 t = min_dist(g, disp, cq);
            if (t.x * t.x + t.y * t.y > rr) continue;
            int res = clip_ring(r, jj, disp, cand, wstart, wrap_run);
            if (res != CLIP_NONE) return res;
        }
    }
    return CLIP_NONE;
}

/* Phase 1, virtual_voronoi_calc_function (DelaunayGPU.cu:318-608): start
 * from four virtual points on a square of half-diagonal ~L/2 and cut each
 * vertex once with the first cutting point found around v's own cell. */
static int candidate_polygon(const orc_ctx* C, int k, ring_t* r)
{
    const orc_grid* g = C->g;
    const double Lmax = g->lhalf;
    const double LL = Lmax / 1.414213562373095 - ORC_EPS;
    r->n = 4;
    r->P[0].x = LL;  r->P[0].y = LL;
    r->P[1].x = -LL; r->P[1].y = LL;
    r->P[2].x = -LL; r->P[2].y = -LL;
    r->P[3].x = LL;  r->P[3].y = -LL;
    for (int m = 0; m < 4; ++m) {
        r->idx[m] = -1;
        r->Q[m] = circumcenter0(r->P[m], r->P[(m + 1) % 4]);
    }
    const d2 v = C->pt[k];
    const d2 vb = put_in_box(g, v);
    int cx, cy;
    cell_of(g, vb, &cx, &cy);
    for (int jj = 0; jj < r->n; ++jj) {
        int res = search_and_clip(C, r, k, v, jj, cx, cy, 0, 1);
        if (res == CLIP_FULL) return -1;
    }
    return 0;
}

/* Phase 2, get_oneRing_function (DelaunayGPU.cu:680-943): for every vertex,
 * search the cells covering its circle (centred on the vertex) for a point
 * strictly inside; on a cut, re-test the same vertex. */
static int refine_ring(const orc_ctx* C, int k, ring_t* r)
{
    const orc_grid* g = C->g;
    const d2 v = C->pt[k];
    for (int jj = 0; jj < r->n; ++jj) {
        d2 c = { v.x + r->Q[jj].x, v.y + r->Q[jj].y };
        c = put_in_box(g, c);
        int cx, cy;
        cell_of(g, c, &cx, &cy);
        int res = search_and_clip(C, r, k, v, jj, cx, cy, jj, jj == 0);
        if (res == CLIP_FULL) return -1;
        if (res == CLIP_DONE) jj--;
    }
    return 0;
}

typedef struct {
    orc_grid g;
    int* cstart;
    int* cidx;
} orc_cells;

static int build_cells(const double* xy, int n, const double* box, double cell_size, orc_cells* cl)
{
    orc_make_grid_box(box, cell_size, &cl->g);
    int ncell = cl->g.xsize * cl->g.ysize;
    cl->cstart = (int*)malloc(sizeof(int) * (size_t)(ncell + 1));
    cl->cidx = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    if (!cl->cstart || !cl->cidx) return -1;
    return orc_cell_list(xy, n, &cl->g, cl->cstart, cl->cidx);
}

static void free_cells(orc_cells* cl)
{
    free(cl->cstart);
    free(cl->cidx);
}

/* Triangulate the points selected by sel (all when sel == NULL). */
static int triangulate_selected(const double* xy, int n, const orc_cells* cl, const int* sel,
                                int max_nbrs, int* nbrs, int* counts, int* required, int nthreads)
{
    orc_ctx C = { (const d2*)xy, &cl->g, cl->cstart, cl->cidx };
    int worst = 0, err = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel reduction(max : worst) reduction(min : err)
    {
        ring_t* r = (ring_t*)malloc(sizeof(ring_t));
#pragma omp for schedule(dynamic, 256)
        for (int k = 0; k < n; ++k) {
            if (sel && sel[k] < 0) continue;
            if (!r || candidate_polygon(&C, k, r) != 0 || refine_ring(&C, k, r) != 0) {
                err = -3;
                continue;
            }
            if (r->n > worst) worst = r->n;
            counts[k] = r->n;
            int m = r->n < max_nbrs ? r->n : max_nbrs;
            for (int q = 0; q < m; ++q) nbrs[(size_t)k * max_nbrs + q] = r->idx[q];
        }
        free(r);
    }
    if (err < 0) return err;
    if (required) *required = worst;
    return worst > max_nbrs ? 1 : 0;
}

int orc_triangulate(const double* xy, int n, double lx, double ly, double cell_size,
                    int max_nbrs, int* nbrs, int* counts, int* required, int nthreads)
{
    const double box[4] = { lx, 0., 0., ly };
    return orc_triangulate_box(xy, n, box, cell_size, max_nbrs, nbrs, counts, required, nthreads);
}

int orc_triangulate_box(const double* xy, int n, const double* box, double cell_size,
                        int max_nbrs, int* nbrs, int* counts, int* required, int nthreads)
{
    if (n <= 0) return -1;
    orc_cells cl;
    if (build_cells(xy, n, box, cell_size, &cl) != 0) { free_cells(&cl); return -4; }
    int rc = triangulate_selected(xy, n, &cl, NULL, max_nbrs, nbrs, counts, required, nthreads);
    free_cells(&cl);
    return rc;
}

/* getCircumcirclesCPU (DelaunayGPU.cpp:673-706). */
int orc_triangles(const int* nbrs, const int* counts, int n, int stride, int* tris, int cap,
                  long long* total_nbrs)
{
    long long total = 0;
    int t = 0;
    for (int i = 0; i < n; ++i) {
        int c = counts[i];
        total += c;
        if (c <= 0) continue;
        int prev = nbrs[(size_t)i * stride + c - 1];
        for (int m = 0; m < c; ++m) {
            int cur = nbrs[(size_t)i * stride + m];
            if (i < prev && i < cur) {
                if (t >= cap) return -1;
                tris[3 * t] = i;
                tris[3 * t + 1] = prev;
                tris[3 * t + 2] = cur;
                t++;
            }
            prev = cur;
        }
    }
    if (total_nbrs) *total_nbrs = total;
    return t;
}

/* test_circumcircle_kernel_function (DelaunayGPU.cu:192-284): spiral over
 * the cells around the circumcentre; the first point strictly inside the
 * circle is flagged together with the three vertices. */
int orc_test_circumcircles(const double* xy, int n, double lx, double ly, double cell_size,
                           const int* tris, int ntri, int* repair, int nthreads)
{
    const double box[4] = { lx, 0., 0., ly };
    return orc_test_circumcircles_box(xy, n, box, cell_size, tris, ntri, repair, nthreads);
}

int orc_test_circumcircles_box(const double* xy, int n, const double* box, double cell_size,
                               const int* tris, int ntri, int* repair, int nthreads)
{
    orc_cells cl;
    if (build_cells(xy, n, box, cell_size, &cl) != 0) { free_cells(&cl); return -4; }
    const orc_grid* g = &cl.g;
    const d2* pt = (const d2*)xy;
    for (int i = 0; i < n; ++i) repair[i] = -1;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(dynamic, 1024)
    for (int t = 0; t < ntri; ++t) {
        const int a = tris[3 * t], b = tris[3 * t + 1], c = tris[3 * t + 2];
        const d2 v = pt[a];
        const d2 p1 = min_dist(g, pt[b], v);
        const d2 p2 = min_dist(g, pt[c], v);
        const d2 q = circumcenter0(p1, p2);
        const double R = sqrt(q.x * q.x + q.y * q.y);     /* Circumcircle, functions.h:185-189 */
        d2 qb = { v.x + q.x, v.y + q.y };
        qb = put_in_box(g, qb);
        int cx, cy;
        cell_of(g, qb, &cx, &cy);
        /* the test caps the half-width by max(xsize,ysize)/2 (DelaunayGPU.cu:227-228) */
        const double bs = g->hcx < g->hcy ? g->hcx : g->hcy;
        int cap = (g->xsize > g->ysize ? g->xsize : g->ysize) / 2;
        int w = (int)ceil(R / bs);
        if (w > cap) w = cap;
        const int side = 2 * w + 1;
        const double rad2 = R * R;
        int bad = 0, dd = 0, cc = 0;
        for (int s = 0; s < side * side && !bad; ++s) {
            int x = pmod(cx + dd, g->xsize);
            int y = pmod(cy + cc, g->ysize);
            spiral_next(&dd, &cc);
            int bin = y * g->xsize + x;
            for (int k = cl.cstart[bin]; k < cl.cstart[bin + 1]; ++k) {
                int p = cl.cidx[k];
                if (p == a || p == b || p == c) continue;
                d2 d = min_dist(g, pt[p], v);
                double ex = d.x - q.x, ey = d.y - q.y;
                if (ex * ex + ey * ey < rad2) {
                    repair[p] = p;
                    bad = 1;
                    break;
                }
            }
        }
        if (bad) {
            repair[a] = a;
            repair[b] = b;
            repair[c] = c;
        }
    }
    free_cells(&cl);
    int nf = 0;
    for (int i = 0; i < n; ++i) nf += repair[i] >= 0;
    return nf;
}

int orc_repair(const double* xy, int n, double lx, double ly, double cell_size, const int* repair,
               int max_nbrs, int* nbrs, int* counts, int* required, int nthreads)
{
    const double box[4] = { lx, 0., 0., ly };
    return orc_repair_box(xy, n, box, cell_size, repair, max_nbrs, nbrs, counts, required,
                          nthreads);
}

int orc_repair_box(const double* xy, int n, const double* box, double cell_size,
                   const int* repair, int max_nbrs, int* nbrs, int* counts, int* required,
                   int nthreads)
{
    orc_cells cl;
    if (build_cells(xy, n, box, cell_size, &cl) != 0) { free_cells(&cl); return -4; }
    int rc = triangulate_selected(xy, n, &cl, repair, max_nbrs, nbrs, counts, required, nthreads);
    free_cells(&cl);
    return rc;
}

int orc_test_and_repair(const double* xy, int n, double lx, double ly, double cell_size,
                        int max_nbrs, int* nbrs, int* counts, int* required, int* nflagged,
                        int nthreads)
{
    const double box[4] = { lx, 0., 0., ly };
    return orc_test_and_repair_box(xy, n, box, cell_size, max_nbrs, nbrs, counts, required,
                                   nflagged, nthreads);
}

int orc_test_and_repair_box(const double* xy, int n, const double* box, double cell_size,
                            int max_nbrs, int* nbrs, int* counts, int* required, int* nflagged,
                            int nthreads)
{
    int cap = 2 * n;
    int* tris = (int*)malloc(sizeof(int) * 3 * (size_t)cap);
    int* rep = (int*)malloc(sizeof(int) * (size_t)n);
    if (!tris || !rep) { free(tris); free(rep); return -4; }
    long long total = 0;
    int nt = orc_triangles(nbrs, counts, n, max_nbrs, tris, cap, &total);
    if (nt != 2 * n || total != 6LL * n) { free(tris); free(rep); return -2; }
    int nf = orc_test_circumcircles_box(xy, n, box, cell_size, tris, nt, rep, nthreads);
    if (nflagged) *nflagged = nf;
    int rc = 0;
    if (nf > 0)
        rc = orc_repair_box(xy, n, box, cell_size, rep, max_nbrs, nbrs, counts, required, nthreads);
    free(tris);
    free(rep);
    return rc;
}

/* ------------------------------------------------------------------ */
/* Voronoi energy and gradient (no reference counterpart)              */
/* ------------------------------------------------------------------ */

static inline double dmin(double a, double b) { return a < b ? a : b; }

/* Jacobians of c = circumcenter0(a, b) with respect to a and b (2x2,
 * row-major: J[r][s] = d c_r / d x_s). */
static void circ_jac(d2 a, d2 b, double Ja[4], double Jb[4])
{
    double aa = a.x * a.x + a.y * a.y, bb = b.x * b.x + b.y * b.y;
    double D = a.x * b.y - a.y * b.x;
    double Nx = aa * b.y - bb * a.y, Ny = bb * a.x - aa * b.x;
    double i2 = 1.0 / (2.0 * D * D);
    /* dN/da, dD/da */
    double dNx_ax = 2 * a.x * b.y, dNx_ay = 2 * a.y * b.y - bb;
    double dNy_ax = bb - 2 * a.x * b.x, dNy_ay = -2 * a.y * b.x;
    double dD_ax = b.y, dD_ay = -b.x;
    Ja[0] = (dNx_ax * D - Nx * dD_ax) * i2;
    Ja[1] = (dNx_ay * D - Nx * dD_ay) * i2;
    Ja[2] = (dNy_ax * D - Ny * dD_ax) * i2;
    Ja[3] = (dNy_ay * D - Ny * dD_ay) * i2;
    double dNx_bx = -2 * b.x * a.y, dNx_by = aa - 2 * b.y * a.y;
    double dNy_bx = 2 * b.x * a.x - aa, dNy_by = 2 * b.y * a.x;
    double dD_bx = -a.y, dD_by = a.x;
    Jb[0] = (dNx_bx * D - Nx * dD_bx) * i2;
    Jb[1] = (dNx_by * D - Nx * dD_by) * i2;
    Jb[2] = (dNy_bx * D - Ny * dD_bx) * i2;
    Jb[3] = (dNy_by * D - Ny * dD_by) * i2;
}

double orc_voronoi_energy(const double* xy, int n, double lx, double ly, const int* nbrs,
                          const int* counts, int stride, double a0, double p0, double* area,
                          double* perim, double* grad, int nthreads)
{
    const double box[4] = { lx, 0., 0., ly };
    return orc_voronoi_energy_box(xy, n, box, nbrs, counts, stride, a0, p0, area, perim, grad,
                                  nthreads);
}

double orc_voronoi_energy_box(const double* xy, int n, const double* box, const int* nbrs,
                              const int* counts, int stride, double a0, double p0, double* area,
                              double* perim, double* grad, int nthreads)
{
    orc_grid g;
    orc_make_grid_box(box, dmin(fabs(box[0]), fabs(box[3])), &g);   /* only minimum images are used */
    const d2* pt = (const d2*)xy;
    double* G = (double*)malloc(sizeof(double) * 2 * (size_t)n * stride);  /* dE_c/dH_l */
    double* Ei = (double*)malloc(sizeof(double) * (size_t)n);
    if (!G || !Ei) { free(G); free(Ei); return NAN; }
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
#pragma omp parallel for schedule(static)
    for (int c = 0; c < n; ++c) {
        int k = counts[c];
        d2 h[ORC_SCRATCH];
        for (int m = 0; m < k; ++m) {
            d2 a = min_dist(&g, pt[nbrs[(size_t)c * stride + m]], pt[c]);
            d2 b = min_dist(&g, pt[nbrs[(size_t)c * stride + (m + 1) % k]], pt[c]);
            h[m] = circumcenter0(a, b);
        }
        double A = 0, P = 0;
        for (int m = 0; m < k; ++m) {
            d2 p = h[m], q = h[(m + 1) % k];
            A += 0.5 * (p.x * q.y - q.x * p.y);
            P += sqrt((q.x - p.x) * (q.x - p.x) + (q.y - p.y) * (q.y - p.y));
        }
        if (area) area[c] = A;
        if (perim) perim[c] = P;
        Ei[c] = (A - a0) * (A - a0) + (P - p0) * (P - p0);
        double ca = 2 * (A - a0), cp = 2 * (P - p0);
        for (int m = 0; m < k; ++m) {
            d2 hp = h[(m + k - 1) % k], hm = h[m], hn = h[(m + 1) % k];
            double gx = ca * 0.5 * (hn.y - hp.y);
            double gy = ca * 0.5 * (hp.x - hn.x);
            double lp = sqrt((hm.x - hp.x) * (hm.x - hp.x) + (hm.y - hp.y) * (hm.y - hp.y));
            double ln = sqrt((hn.x - hm.x) * (hn.x - hm.x) + (hn.y - hm.y) * (hn.y - hm.y));
            gx += cp * ((hm.x - hp.x) / lp - (hn.x - hm.x) / ln);
            gy += cp * ((hm.y - hp.y) / lp - (hn.y - hm.y) / ln);
            G[2 * ((size_t)c * stride + m)] = gx;
            G[2 * ((size_t)c * stride + m) + 1] = gy;
        }
    }
    double E = 0;
    for (int c = 0; c < n; ++c) E += Ei[c];
    if (grad) {
#pragma omp parallel for schedule(static)
        for (int p = 0; p < n; ++p) {
            int k = counts[p];
            double gx = 0, gy = 0;
            for (int m = 0; m < k; ++m) {
                int j = nbrs[(size_t)p * stride + m];
                int l = nbrs[(size_t)p * stride + (m + 1) % k];
                /* dE/dH for the vertex (p, j, l): own cell slot m, cell j slot
                 * (pos of p in j) - 1, cell l slot (pos of p in l). */
                double sx = G[2 * ((size_t)p * stride + m)], sy = G[2 * ((size_t)p * stride + m) + 1];
                int kj = counts[j], kl = counts[l], pj = -1, pl = -1;
                for (int q = 0; q < kj; ++q) if (nbrs[(size_t)j * stride + q] == p) { pj = q; break; }
                for (int q = 0; q < kl; ++q) if (nbrs[(size_t)l * stride + q] == p) { pl = q; break; }
                if (pj >= 0) {
                    size_t s = (size_t)j * stride + (pj + kj - 1) % kj;
                    sx += G[2 * s]; sy += G[2 * s + 1];
                }
                if (pl >= 0) {
                    size_t s = (size_t)l * stride + pl;
                    sx += G[2 * s]; sy += G[2 * s + 1];
                }
                d2 a = min_dist(&g, pt[j], pt[p]);
                d2 b = min_dist(&g, pt[l], pt[p]);
                double Ja[4], Jb[4];
                circ_jac(a, b, Ja, Jb);
                /* H = x_p + c(a,b), a = x_j - x_p, b = x_l - x_p: dH/dx_p = I - Ja - Jb */
                double J0 = 1 - Ja[0] - Jb[0], J1 = -Ja[1] - Jb[1];
                double J2 = -Ja[2] - Jb[2], J3 = 1 - Ja[3] - Jb[3];
                gx += sx * J0 + sy * J2;
                gy += sx * J1 + sy * J3;
            }
            grad[2 * p] = gx;
            grad[2 * p + 1] = gy;
        }
    }
    free(G);
    free(Ei);
    return E;
}
