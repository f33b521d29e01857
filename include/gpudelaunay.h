/*
 * gpudelaunay.h -- C ABI of the gfx950 localized periodic Delaunay library
 * (libgpudelaunay.so).
 *
 * Every entry point is extern "C", takes plain device pointers and sizes,
 * never throws, and returns a gdt_status.  Device pointers are HIP device
 * allocations; `stream` is a hipStream_t passed as void* (NULL = the legacy
 * default stream, the reference's only stream).
 *
 * Layouts (as in the reference, DelaunayGPU.cpp:47, indexer.h:38-41):
 *   points   double[2*n]            x0,y0,x1,y1,...  (double2 array)
 *   nbrs     int[n*stride]          neighbour m of point i at i*stride+m,
 *                                    counter-clockwise, counts[i] entries
 *   counts   int[n]
 *   tris     int[3*ntri]            (i, j, k) owned by their smallest index
 *   repair   int[n]                 i if point i must be re-triangulated, else -1
 *
 * Output rings start at each point's nearest neighbour; the reference leaves
 * the start unspecified, so compare rings after rotating each to start at its
 * smallest index (the reference's own comparison, triangulation.cpp:108-119).
 */
#ifndef GPUDELAUNAY_H
#define GPUDELAUNAY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    GDT_OK = 0,
    GDT_OVERFLOW = 1,          /* a ring needs more than `stride` slots; *required set */
    GDT_ERR_INVALID = -1,      /* bad argument */
    GDT_ERR_DEVICE = -2,       /* HIP runtime error */
    GDT_ERR_NONGENERIC = -3,   /* a ring did not close / no neighbour: degenerate input */
    GDT_ERR_EULER = -4,        /* sum(counts) != 6n or triangles != 2n */
    GDT_ERR_SEARCH = -5,       /* search radius exhausted (box too small for the point set) */
    GDT_ERR_NOMEM = -6,
    GDT_ERR_SCRATCH = -7       /* input too large for gdt_triangulate_noalloc's fixed scratch */
} gdt_status;

typedef enum {
    GDT_PREC_FP64 = 64,        /* all predicates in FP64 */
    GDT_PREC_FP32F = 32,       /* FP32 with certified error bound, FP64 fallback */
    GDT_PREC_EXACT = 128       /* every decision exact (FP64 where its error bound
                                  certifies it, else expansion arithmetic; exact ties
                                  by the symbolic perturbation): the per-point
                                  kernels, rectangular boxes, caller coordinates
                                  within four box lengths of the box */
} gdt_precision;

typedef struct gdt_context gdt_context;

/* Tunables (all have defaults; the autotuner writes them). */
typedef struct {
    int block;          /* threads per block of the one-ring kernel (64..1024) */
    int tiled;          /* cell order: 1 = 32x32-cell tiles (LDS tile kernel, default),
                           2 = the same tiles, Morton order inside, 0 = row-major cells
                           (direct per-point kernel) */
    double cell_scale;  /* multiplies the caller's cell size (1.0 = as given) */
    int profile;        /* 1: time the cell list and one-ring kernels with HIP events */
    int tile_shift;     /* LDS tile of the tiled orders: 2^tile_shift cells a side, 4 (16x16)
                           or 5 (32x32); 0 (default) = automatic: 16x16 when the 32x32
                           tiles would not outnumber the compute units, else 32x32 */
    int lanes;          /* lanes per point of the global triangulation's tile kernel, 1, 2
                           or 4 (4: FP64; the FP32-filtered kernel takes 2 then), or 64:
                           one wavefront per point whose lanes search from every candidate
                           of the point's cell block at once (FP64 predicates in either
                           precision, the same rings; 256-thread blocks; for small grids);
                           0 (default) = automatic: 64 with 16x16 tiles that number at most
                           a quarter of the compute units and at most 1.25 points per cell,
                           2 with 16x16 tiles that do not outnumber the compute units,
                           else 1 */
    int repair_lanes;   /* lanes per point of the repairs' tile kernel, 1, 2 or 4 (as
                           lanes); 0 (default) = 2 */
} gdt_config;

/* Statistics of the last triangulate / repair call. */
typedef struct {
    unsigned long long predicates;   /* predicate evaluations */
    unsigned long long fallbacks;    /* FP64 fallbacks of the filtered path */
    unsigned long long neighbours;   /* sum of ring lengths written */
    int max_ring;                    /* largest ring */
    int flagged;                     /* points flagged by the last circumcircle test */
    int xsize, ysize;                /* cell grid of the last single-set call */
    double bsx, bsy;                 /* cell sides */
    float ms_cells;                  /* last cell-list build (profile = 1), ms */
    float ms_ring;                   /* last one-ring kernel (profile = 1), ms */
    int deferred;                    /* points the LDS tile kernel handed back */
    float ms_defer;                  /* kernels of the handed-back points (profile = 1), ms */
    /* running totals over every global triangulation timed with profile = 1
     * since the context was created (read twice and subtract to average a
     * loop without reading the statistics inside it) */
    unsigned long long profiled_calls;
    double ms_cells_total, ms_ring_total, ms_defer_total;
    int tile_shift;                  /* tile edge of the last global triangulation (0: row-major) */
    int lanes;                       /* its lanes per point (as launched) */
} gdt_stats;

/* ---- context ---------------------------------------------------------- */

/* Create a context on HIP device `device` (scratch buffers are owned by the
 * context and grow on demand).  Replaces DelaunayGPU::initialize
 * (DelaunayGPU.cpp:16-37) for the device-side state. */
int gdt_create(gdt_context** ctx, int device);
void gdt_destroy(gdt_context* ctx);
int gdt_set_config(gdt_context* ctx, const gdt_config* cfg);
int gdt_get_config(const gdt_context* ctx, gdt_config* cfg);
int gdt_get_stats(const gdt_context* ctx, gdt_stats* out);
const char* gdt_status_string(int status);

/* ---- cell list (cellListGPU) ------------------------------------------ */

/* Bin n points of a periodic lx x ly box into cells of side ~cell_size
 * (counting sort; within a cell points keep index order).
 * Replaces cellListGPU::setGridSize + computeGPU(points)
 * (cellListGPU.cpp:97-117, 404-460; cellListGPU.cu:20-105). */
int gdt_celllist_build(gdt_context* ctx, const double* d_xy, int n, double lx, double ly,
                       double cell_size, void* stream);

/* Grid of the last cell list and the largest cell occupancy (synchronises). */
int gdt_celllist_info(gdt_context* ctx, int* xsize, int* ysize, double* bsx, double* bsy,
                      int* max_occupancy, void* stream);

/* Present the last cell list in the reference layout: cell_sizes[xsize*ysize]
 * (row-major cells, Index2D(xsize,ysize)) and idxs[cell*nmax + k], nmax >=
 * max_occupancy (cellListGPU.h:93-102). */
int gdt_celllist_export(gdt_context* ctx, unsigned int* d_cell_sizes, int* d_idxs, int nmax,
                        void* stream);

/* ---- global triangulation (DelaunayGPU::globalDelaunayTriangulation) -- */

/* Cell list + one-ring of every point.  On GDT_OVERFLOW the caller grows
 * stride to *required and calls again (the reference's safetyMode loop,
 * DelaunayGPU.cpp:137-156, which this reports in one pass).
 * Replaces gpu_voronoi_calc + gpu_get_neighbors (DelaunayGPU.cuh:20-78;
 * DelaunayGPU.cu:1116-1317). */
int gdt_triangulate(gdt_context* ctx, const double* d_xy, int n, double lx, double ly,
                    double cell_size, int stride, int* d_nbrs, int* d_counts, int precision,
                    int* required, void* stream);

/* The same global triangulation without any allocation, for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc); it has no counterpart in the
 * reference.  No context: every scratch buffer lives in one fixed 96 MiB
 * device array of the library (each GPU holds its own copy from the library's
 * first use on it until the process ends), enough for 2^20 points at density
 * 1 with cell size 1 (68 bytes per point, 8 per cell key and the cell-list
 * bucket tables); larger inputs give GDT_ERR_SCRATCH.  Default configuration
 * (32x32-cell tiles), the current device, statuses as gdt_triangulate.  Calls
 * that overlap on one GPU share that array and race: serialise them. */
int gdt_triangulate_noalloc(const double* d_xy, int n, double lx, double ly, double cell_size,
                            int stride, int* d_nbrs, int* d_counts, int precision, int* required,
                            void* stream);

/* ---- test and repair (DelaunayGPU::testAndRepairDelaunayTriangulation) */

/* Owned triangles of a ring set, in point order (deterministic).
 * Replaces gpu_get_circumcircles (DelaunayGPU.cuh:80-86; DelaunayGPU.cu:1008-1028,
 * 1319-1343).  *ntri and *total_nbrs (sum of counts) are read back. */
int gdt_triangles(gdt_context* ctx, const int* d_nbrs, const int* d_counts, int n, int stride,
                  int* d_tris, int cap, int* ntri, long long* total_nbrs, void* stream);

/* Empty-circle test of ntri triangles against the last cell list (built for
 * the same d_xy): repair[v] = v for the three vertices of every non-empty
 * triangle and for the first point found inside it, else -1.
 * Replaces gpu_test_circumcircles (DelaunayGPU.cuh:88-104; DelaunayGPU.cu:192-313). */
int gdt_test_circumcircles(gdt_context* ctx, const double* d_xy, int n, const int* d_tris,
                           int ntri, int* d_repair, int* nflagged, void* stream);

/* Re-triangulate the points with repair[i] >= 0 against the last cell list.
 * Replaces gpu_voronoi_calc_no_sort + gpu_get_neighbors_no_sort
 * (DelaunayGPU.cuh:40-60, 106-129; DelaunayGPU.cu:1036-1114, 1191-1256). */
int gdt_repair(gdt_context* ctx, const double* d_xy, int n, const int* d_repair, int stride,
               int* d_nbrs, int* d_counts, int precision, int* required, void* stream);

/* Triangles -> cell list -> test -> repair, all on the device; *nflagged is
 * the number of repaired points.  Input rings must use `stride`.
 * Replaces DelaunayGPU::testAndRepairDelaunayTriangulation (DelaunayGPU.cpp:567-602). */
int gdt_test_and_repair(gdt_context* ctx, const double* d_xy, int n, double lx, double ly,
                        double cell_size, int stride, int* d_nbrs, int* d_counts, int precision,
                        int* nflagged, int* required, void* stream);

/* ---- batches of independent sets (extension) ---------------------------
 *
 * nsets periodic sets concatenated in d_xy: set s is the caller points
 * [offsets[s], offsets[s+1]) in a boxes[2s] x boxes[2s+1] box (offsets and
 * boxes are host arrays, offsets[0] = 0).  One launch per kernel covers every
 * set; rows of set s hold indices into the concatenated array (its own range).
 * The reference has no batch API: a caller runs one DelaunayGPU per set
 * (DelaunayGPU.cpp:512-565, 567-602 per set).  Status as the single-set calls;
 * *required is the longest ring over all sets. */
int gdt_triangulate_batch(gdt_context* ctx, const double* d_xy, int nsets, const int* offsets,
                          const double* boxes, double cell_size, int stride, int* d_nbrs,
                          int* d_counts, int precision, int* required, void* stream);
int gdt_test_and_repair_batch(gdt_context* ctx, const double* d_xy, int nsets, const int* offsets,
                              const double* boxes, double cell_size, int stride, int* d_nbrs,
                              int* d_counts, int precision, int* nflagged, int* required,
                              void* stream);

/* ---- Brownian test-and-repair loop (extension) --------------------------
 *
 * nsteps steps on a batch whose rings are a triangulation of d_xy: each step
 * moves every point by sigma * N(0,1) per coordinate (counter-based noise of
 * (seed, step, set, point); wrapped into its box), then runs test-and-repair
 * (triangles -> cell list -> empty-circle test -> repair) entirely on the
 * device.  use_graph = 1 captures one step in a hipGraph and replays it.
 *
 * The first step that fails -- a repaired ring longer than `stride` (its row
 * is truncated), an error flag, or input rings that are not a triangulation
 * -- halts the loop ON THE DEVICE at the end of that step: every kernel of
 * every later step exits at once, so no step ever runs on truncated rows.
 * The failing step's displacement is applied and counted in steps_done
 * (first_bad_step = step0 + steps_done - 1); its rings are invalid and must
 * be rebuilt by gdt_triangulate_batch of the current positions, which
 * completes that step.  The host reads the device state every check_every
 * steps and returns at the first check after the halt (GDT_OVERFLOW with
 * report.required; GDT_ERR_NONGENERIC; GDT_ERR_EULER).  Replaces the
 * reference driver's loop of host displacement + testAndRepair per step
 * (triangulation.cpp:349-415; DelaunayGPU.cpp:567-602).  A null stream runs on
 * a stream of the context (synchronised with the device first). */
typedef struct {
    int steps_done;         /* steps whose displacement was applied */
    int first_bad_step;     /* the step that halted the loop, -1 none */
    int required;           /* longest ring seen */
    int max_flagged;        /* most points repaired in one step */
    int skipped_steps;      /* replays enqueued after the halt, which did nothing */
    long long flagged_total;/* points repaired, summed over the steps */
} gdt_md_report;

int gdt_md_run(gdt_context* ctx, double* d_xy, int nsets, const int* offsets, const double* boxes,
               double cell_size, int stride, int* d_nbrs, int* d_counts, int precision,
               double sigma, unsigned long long seed, int step0, int nsteps, int use_graph,
               int check_every, gdt_md_report* report, void* stream);

/* ---- Voronoi energy (extension; no reference counterpart) --------------
 *
 * For the tessellation dual to a ring table: cell areas A_i, perimeters P_i
 * and E = sum_i (A_i - a0)^2 + (P_i - p0)^2, then the adjoint
 * grad[i] = gscale * dE/dx_i (FP64 output in both precisions; gscale is a
 * device scalar, null = 1).  The Voronoi vertices are the circumcentres the
 * reference keeps per point in GPUVoroCur (DelaunayGPU.cu:921-922,
 * DelaunayGPU.h:57-58).  Rows that are not rings of a triangulation give NaN.
 * Asynchronous: no host synchronisation; `area` and `perim` of the forward
 * call are inputs of the adjoint. */
/* Voronoi vertex m of every point relative to it, d_vertices[i*stride + m]
 * (double2 per slot; the GPUVoroCur array of the reference, filled there by
 * get_oneRing_function, DelaunayGPU.cu:921-922); invalid rows give NaN. */
int gdt_voronoi_vertices(gdt_context* ctx, const double* d_xy, int n, double lx, double ly,
                         const int* d_nbrs, const int* d_counts, int stride, double* d_vertices,
                         void* stream);
int gdt_voronoi_forward(gdt_context* ctx, const double* d_xy, int n, double lx, double ly,
                        const int* d_nbrs, const int* d_counts, int stride, double a0, double p0,
                        int precision, double* d_area, double* d_perim, double* d_energy,
                        void* stream);
int gdt_voronoi_adjoint(gdt_context* ctx, const double* d_xy, int n, double lx, double ly,
                        const int* d_nbrs, const int* d_counts, int stride, double a0, double p0,
                        int precision, const double* d_area, const double* d_perim,
                        const double* d_gscale, double* d_grad, void* stream);

/* Batched form: nsets sets concatenated in d_xy, set s = points
 * [offsets[s], offsets[s+1]) in an boxes[2s] x boxes[2s+1] box (host arrays);
 * ring entries are indices into the concatenated array; d_energy[s] and
 * d_gscale[s] per set. */
int gdt_voronoi_forward_batch(gdt_context* ctx, const double* d_xy, int nsets, const int* offsets,
                              const double* boxes, const int* d_nbrs, const int* d_counts,
                              int stride, double a0, double p0, int precision, double* d_area,
                              double* d_perim, double* d_energy, void* stream);
int gdt_voronoi_adjoint_batch(gdt_context* ctx, const double* d_xy, int nsets, const int* offsets,
                              const double* boxes, const int* d_nbrs, const int* d_counts,
                              int stride, double a0, double p0, int precision,
                              const double* d_area, const double* d_perim, const double* d_gscale,
                              double* d_grad, void* stream);

/* ---- CPU mode (DelaunayGPU::setGPUcompute(false)) ----------------------
 *
 * The same per-point routines as the kernels, compiled for the host and run
 * over `nthreads` OpenMP threads (<= 0: the OpenMP default; the reference's
 * setOMPthreads).  Host pointers, no context and no HIP call; the layouts,
 * statuses and rings are those of the device calls.  The cell list uses
 * row-major cells (the reference's own order). */

/* Cell list on the host: *xsize, *ysize and *max_occupancy always; with
 * cell_sizes non-null also cell_sizes[xsize*ysize] and idxs[cell*nmax + k]
 * in the reference layout (GDT_OVERFLOW when nmax < *max_occupancy).
 * Replaces cellListGPU::compute(points) (cellListGPU.cpp:289-330). */
int gdt_host_celllist(const double* xy, int n, double lx, double ly, double cell_size,
                      unsigned int* cell_sizes, int* idxs, int nmax, int* xsize, int* ysize,
                      int* max_occupancy);

/* Global triangulation on the host (statuses as gdt_triangulate).
 * Replaces Voronoi_Calc_CPU + get_neighbors_CPU (DelaunayGPU.cpp:159-343;
 * the OpenMP branches of gpu_voronoi_calc / gpu_get_neighbors,
 * DelaunayGPU.cu:1116-1317). */
int gdt_host_triangulate(const double* xy, int n, double lx, double ly, double cell_size,
                         int stride, int* nbrs, int* counts, int precision, int nthreads,
                         int* required);

/* Test and repair on the host (statuses as gdt_test_and_repair).
 * Replaces testAndRepairDelaunayTriangulation with GPUcompute == false
 * (DelaunayGPU.cpp:567-706; DelaunayGPU.cu:1346-1409 CPU branch). */
int gdt_host_test_and_repair(const double* xy, int n, double lx, double ly, double cell_size,
                             int stride, int* nbrs, int* counts, int precision, int nthreads,
                             int* nflagged, int* required);

/* Voronoi vertices relative to each point, vertices[2*(i*stride + m)] (the
 * reference's GPUVoroCur, filled by get_oneRing_function,
 * DelaunayGPU.cu:921-922); rows that are not rings give NaN. */
int gdt_host_voronoi_vertices(const double* xy, int n, double lx, double ly, const int* nbrs,
                              const int* counts, int stride, double* vertices, int nthreads);

/* ---- general (triclinic) boxes -----------------------------------------
 *
 * box = {x11, x12, x21, x22}, the box matrix of periodicBoundaries::setGeneral
 * (periodicBoundaries.h:86-101): its columns are the lattice vectors
 * a = (x11, x21), b = (x12, x22); any invertible matrix.  The reference keeps
 * general boxes in periodicBoundaries only; its cell list, and so DelaunayGPU,
 * handles rectangles (cellListGPU.cpp:97-117).  Here every call of the path
 * takes them: points are binned in the unit-cell coordinates of the
 * Lagrange-Gauss reduced basis (a', b') of the box's lattice (the same torus;
 * parallelogram cells spanned by a'/xsize and b'/ysize, the counts given by
 * setGridSize's rule applied to the unit cell's thickness across each pair of
 * walls), displacements are cell-relative as for rectangles, and minimum
 * images round unit-cell coordinates (periodicBoundaries::minDist,
 * periodicBoundaries.h:165-178).  Points may lie anywhere (they are wrapped);
 * the Brownian loop wraps moved points into the caller's own box
 * (putInBoxReal, periodicBoundaries.h:131-152).  General boxes run the
 * row-major cell order and FP64 predicates whatever the configuration and
 * precision arguments (the rings are the same); gdt_celllist_info then
 * reports the cell heights as bsx, bsy.  A matrix with x12 = x21 = 0 and a
 * positive diagonal runs the rectangular path.  Batched forms take four
 * doubles per set in `boxes`; otherwise as the rectangular calls. */
int gdt_celllist_build_general(gdt_context* ctx, const double* d_xy, int n, const double* box,
                               double cell_size, void* stream);
int gdt_triangulate_general(gdt_context* ctx, const double* d_xy, int nsets, const int* offsets,
                            const double* boxes, double cell_size, int stride, int* d_nbrs,
                            int* d_counts, int precision, int* required, void* stream);
int gdt_test_and_repair_general(gdt_context* ctx, const double* d_xy, int nsets, const int* offsets,
                                const double* boxes, double cell_size, int stride, int* d_nbrs,
                                int* d_counts, int precision, int* nflagged, int* required,
                                void* stream);
int gdt_md_run_general(gdt_context* ctx, double* d_xy, int nsets, const int* offsets,
                       const double* boxes, double cell_size, int stride, int* d_nbrs,
                       int* d_counts, int precision, double sigma, unsigned long long seed,
                       int step0, int nsteps, int use_graph, int check_every,
                       gdt_md_report* report, void* stream);
int gdt_voronoi_vertices_general(gdt_context* ctx, const double* d_xy, int n, const double* box,
                                 const int* d_nbrs, const int* d_counts, int stride,
                                 double* d_vertices, void* stream);
int gdt_voronoi_forward_general(gdt_context* ctx, const double* d_xy, int nsets, const int* offsets,
                                const double* boxes, const int* d_nbrs, const int* d_counts,
                                int stride, double a0, double p0, int precision, double* d_area,
                                double* d_perim, double* d_energy, void* stream);
int gdt_voronoi_adjoint_general(gdt_context* ctx, const double* d_xy, int nsets, const int* offsets,
                                const double* boxes, const int* d_nbrs, const int* d_counts,
                                int stride, double a0, double p0, int precision,
                                const double* d_area, const double* d_perim,
                                const double* d_gscale, double* d_grad, void* stream);
int gdt_host_celllist_general(const double* xy, int n, const double* box, double cell_size,
                              unsigned int* cell_sizes, int* idxs, int nmax, int* xsize,
                              int* ysize, int* max_occupancy);
int gdt_host_triangulate_general(const double* xy, int n, const double* box, double cell_size,
                                 int stride, int* nbrs, int* counts, int precision, int nthreads,
                                 int* required);
int gdt_host_test_and_repair_general(const double* xy, int n, const double* box, double cell_size,
                                     int stride, int* nbrs, int* counts, int precision,
                                     int nthreads, int* nflagged, int* required);
int gdt_host_voronoi_vertices_general(const double* xy, int n, const double* box,
                                      const int* nbrs, const int* counts, int stride,
                                      double* vertices, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
