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
    GDT_ERR_NOMEM = -6
} gdt_status;

typedef enum {
    GDT_PREC_FP64 = 64,        /* all predicates in FP64 */
    GDT_PREC_FP32F = 32        /* FP32 with certified error bound, FP64 fallback */
} gdt_precision;

typedef struct gdt_context gdt_context;

/* Tunables (all have defaults; the autotuner writes them). */
typedef struct {
    int block;          /* threads per block of the one-ring kernel (64..1024) */
    int tiled;          /* 1: cells stored in 8x8 tiles (Morton-like locality), 0: row-major */
    double cell_scale;  /* multiplies the caller's cell size (1.0 = as given) */
    int profile;        /* 1: time the cell list and one-ring kernels with HIP events */
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
    int deferred;                    /* points the LDS tile kernel handed to the direct kernel */
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

#ifdef __cplusplus
}
#endif
#endif
