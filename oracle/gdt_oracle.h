/*
 * gdt_oracle.h -- CPU oracle for the localized periodic Delaunay path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code.  It is used by tests/ as the parity checker, by
 * tests/golden/make_golden.py to write golden fixtures, and by the
 * cpu_baseline leg of bench.py as the timed CPU path.
 *
 * The code restates, in plain C + OpenMP and FP64, the reference's CPU path:
 *   cellListGPU::setGridSize / compute(points)  (src/utilities/cellListGPU.cpp:97-117, 289-330)
 *   virtual_voronoi_calc_function               (src/model/DelaunayGPU.cu:318-608)
 *   get_oneRing_function                        (src/model/DelaunayGPU.cu:680-943)
 *   test_circumcircle_kernel_function           (src/model/DelaunayGPU.cu:192-284)
 *   DelaunayGPU::getCircumcirclesCPU            (src/model/DelaunayGPU.cpp:673-706)
 *   DelaunayGPU::testAndRepairDelaunayTriangulation (src/model/DelaunayGPU.cpp:567-602, 87-115)
 * plus an FP64 Voronoi energy and its gradient, which has no reference
 * counterpart (pinned by finite differences and an independent autograd
 * computation in tests/).
 *
 * Deliberate deviations from the reference (each is a defect listed in the
 * reference's own behaviour, not a change of algorithm):
 *   - rectangular boxes bin y with Ly/ysize instead of Lx/xsize
 *     (reference: cellListGPU.cpp:106 uses the x cell size for y);
 *   - the spiral half-width uses ceil(R/min(bsx,bsy)) capped at
 *     min(xsize,ysize)/2 (identical to the reference for square boxes);
 *   - ring overflow is always detected and reports the FINAL ring size
 *     (reference: DelaunayGPU.cu:874-879 reports MaxSize+1 and only on GPU);
 *   - cells are clamped to [0,xsize-1] x [0,ysize-1].
 *
 * Degenerate input (four cocircular points, equidistant neighbours) is
 * outside the reference's contract (README.md:55) and outside this oracle's:
 * the device routines decide exact ties by a symbolic perturbation, and the
 * tests check those cases against exact integer arithmetic
 * (tests/exact_delaunay.py), not against this code.
 */
#ifndef GDT_ORACLE_H
#define GDT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Box and cell grid.  Rectangular boxes (periodicBoundaries::setSquare,
 * periodicBoundaries.h:75-84) and general ones (setGeneral, 86-101), whose
 * matrix has the lattice vectors as columns: a = (x11, x21), b = (x12, x22);
 * a point's unit-cell coordinates are u = X^-1 p (invTrans, 117-129).  The
 * grid follows cellListGPU::setGridSize (cellListGPU.cpp:97-117): the cell
 * count per axis is floor(T/a) bumped to the next even number, with T the
 * box side (rectangles) or the box thickness across the walls spanned by the
 * other lattice vector (general boxes, whose cells are the unit-cell grid
 * mapped by X). */
typedef struct {
    int general;                    /* 0: rectangle lx x ly */
    double x11, x12, x21, x22;      /* box matrix */
    double xi11, xi12, xi21, xi22;  /* its inverse */
    double lx, ly;      /* rectangle side lengths */
    double hx, hy;      /* half box */
    double ilx, ily;    /* inverse box lengths (periodicBoundaries.h:79) */
    int xsize, ysize;   /* number of cells */
    double bsx, bsy;    /* cell side lengths (rectangles) */
    double hcx, hcy;    /* cell heights: distance between opposite cell walls */
    double lhalf;       /* half box width: size of the initial virtual polygon */
} orc_grid;

void orc_make_grid(double lx, double ly, double cell_size, orc_grid* g);
/* box = {x11, x12, x21, x22}, invertible; a diagonal matrix with positive
 * entries gives the rectangular grid. */
void orc_make_grid_box(const double* box, double cell_size, orc_grid* g);

/* Bin points into cells in point-index order (CSR form of the reference's
 * padded cell_sizes/idxs arrays).  cell_start has xsize*ysize+1 entries,
 * cell_idx has n entries.  Returns 0, or -1 on allocation failure. */
int orc_cell_list(const double* xy, int n, const orc_grid* g,
                  int* cell_start, int* cell_idx);

/* Full triangulation (globalDelaunayTriangulation in CPU mode).
 * nbrs is [n][max_nbrs]; counts is [n].  Returns 0 on success, 1 if some ring
 * needs more than max_nbrs slots (then *required holds the largest ring
 * size and the rows of the too-long rings are truncated), <0 on error. */
int orc_triangulate(const double* xy, int n, double lx, double ly,
                    double cell_size, int max_nbrs, int* nbrs, int* counts,
                    int* required, int nthreads);

/* Triangle list (getCircumcirclesCPU): triangles (i, n_{m-1}, n_m) owned by
 * their smallest vertex, in point order.  tris has room for 3*cap ints.
 * Returns the number of triangles, or -1 if cap is too small.
 * *total_nbrs receives sum(counts). */
int orc_triangles(const int* nbrs, const int* counts, int n, int stride,
                  int* tris, int cap, long long* total_nbrs);

/* Empty-circle test of every triangle; repair[i] = i for flagged points,
 * -1 otherwise (testTriangulationCPU). Returns the number of flagged points. */
int orc_test_circumcircles(const double* xy, int n, double lx, double ly,
                           double cell_size, const int* tris, int ntri,
                           int* repair, int nthreads);

/* Re-triangulate only points with repair[i] >= 0 (locallyRepair...).
 * Same return convention as orc_triangulate. */
int orc_repair(const double* xy, int n, double lx, double ly, double cell_size,
               const int* repair, int max_nbrs, int* nbrs, int* counts,
               int* required, int nthreads);

/* testAndRepairDelaunayTriangulation in CPU mode.  Returns as orc_repair,
 * or -2 when the Euler check (sum counts = 6n, triangles = 2n) fails.
 * *nflagged receives the number of repaired points. */
int orc_test_and_repair(const double* xy, int n, double lx, double ly,
                        double cell_size, int max_nbrs, int* nbrs, int* counts,
                        int* required, int* nflagged, int nthreads);

/* Voronoi cell area and perimeter with the topology given by the rings, and
 * the energy E = sum_i (A_i - a0)^2 + (P_i - p0)^2 and its gradient dE/dx
 * (grad is [n][2]).  Returns E. */
double orc_voronoi_energy(const double* xy, int n, double lx, double ly,
                          const int* nbrs, const int* counts, int stride,
                          double a0, double p0, double* area, double* perim,
                          double* grad, int nthreads);

/* The same functions for a general box = {x11, x12, x21, x22}: minimum
 * images and wrapping follow periodicBoundaries' general branch
 * (periodicBoundaries.h:117-178), cells are binned in unit-cell coordinates,
 * and spiral half-widths count cell heights.  A diagonal box matrix with
 * positive entries reduces to the rectangular functions above, bit for bit. */
int orc_triangulate_box(const double* xy, int n, const double* box, double cell_size,
                        int max_nbrs, int* nbrs, int* counts, int* required, int nthreads);
int orc_test_circumcircles_box(const double* xy, int n, const double* box, double cell_size,
                               const int* tris, int ntri, int* repair, int nthreads);
int orc_repair_box(const double* xy, int n, const double* box, double cell_size,
                   const int* repair, int max_nbrs, int* nbrs, int* counts, int* required,
                   int nthreads);
int orc_test_and_repair_box(const double* xy, int n, const double* box, double cell_size,
                            int max_nbrs, int* nbrs, int* counts, int* required, int* nflagged,
                            int nthreads);
double orc_voronoi_energy_box(const double* xy, int n, const double* box, const int* nbrs,
                              const int* counts, int stride, double a0, double p0, double* area,
                              double* perim, double* grad, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
