// gdt_kernels.h -- host-side launchers of the gfx950 kernels (one per .hip file).
#pragma once

#include <hip/hip_runtime.h>
#include "gdt_point.h"

// One point of the cell-list build in bucket order (one 32-byte store).
struct GdtKbRec {
    gdt_d2 off;         // offset inside the cell
    int idx;            // caller index
    uint32_t key;       // cell key
    uint32_t cell;      // cx | cy << 16
    uint32_t pad;
};

// Arrays of the cell list (cell order), owned by the context.
struct GdtCells {
    uint32_t* pkey;     // [N] cell key of each caller-order point (single-level path)
    uint32_t* prank;    // [N] arrival rank inside its cell (single-level path)
    uint32_t* ccount;   // [K] points per key
    uint32_t* cstart;   // [K+1] exclusive scan of ccount
    int* perm;          // [N] slot -> caller index
    int* slot_of;       // [N] caller index -> slot (not written when null)
    gdt_d2* off64;      // [N]
    gdt_f2* off32;      // [N]
    uint32_t* scell;    // [N]
    uint32_t* scan_tmp; // scan scratch (gdt_scan_tmp_words(max(K, N+1)) words)
    GdtKbRec* bkt;      // [N] the points in bucket order
    uint32_t* bmat;     // [gdt_bucket_words(N, K)] per (bucket, block) counts
    uint32_t* bscan;    // [gdt_bucket_words(N, K)] their exclusive scan
    const int* halt;    // device word: when non-null and set, the kernels of the
                        // build (and of the tests that take these arrays) exit at once
};

// Words of the (bucket, block) count table of the cell-list build.
size_t gdt_bucket_words(size_t n, size_t k);
// True when n points over k keys take the two-level build, which needs only
// bkt, bmat, bscan and the slot arrays (no pkey, prank or scan scratch).
bool gdt_cells_two_level(size_t n, size_t k);

size_t gdt_scan_tmp_words(size_t n);
// out[0..n) = exclusive prefix sums of in, *total = sum (device pointer).
hipError_t gdt_scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* total,
                        uint32_t* tmp, hipStream_t s);

hipError_t gdt_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);
// *dst = *src, on the stream (one word).
hipError_t gdt_copy_u32(const uint32_t* src, uint32_t* dst, hipStream_t s);
// Zero nst status words and nct counters (nst, nct <= 64).
hipError_t gdt_status_reset(int* st, int nst, unsigned long long* ct, int nct, hipStream_t s);
// Copy them to host memory hst/hct (device-visible pinned pointers), then
// store seq to *hseq (system scope, after the copies are visible).
hipError_t gdt_status_publish(const int* st, int nst, const unsigned long long* ct, int nct,
                              int* hst, unsigned long long* hct, unsigned int* hseq,
                              unsigned int seq, hipStream_t s);
hipError_t gdt_fill_i32(int* p, int v, size_t n, hipStream_t s);

// Cell list over all sets.  pchunks: blocks of 256 points, kchunks: blocks of 256 keys.
// With status non-null its ST_WORDS words and the CT_WORDS counters are
// zeroed on the way (as gdt_status_reset, without a launch of its own).
hipError_t gdt_build_cells(const GdtSet* d_sets, int nsets, const GdtChunk* pchunks, int npch,
                           const GdtChunk* kchunks, int nkch, int total_n, int total_keys,
                           const double* d_xy, GdtCells& c, hipStream_t s,
                           int* status = nullptr, unsigned long long* counters = nullptr);


hipError_t gdt_export_cells(const GdtSet* d_sets, const GdtSet& h_set, const GdtCells& c,
                            unsigned int* cell_sizes, int* idxs, int nmax, hipStream_t s);
hipError_t gdt_max_occupancy(const uint32_t* ccount, int total_keys, int* d_out, hipStream_t s);

// One-ring of every point (chunks of `block` points) or of the worklist slots
// [skip, *worklen) (max_work bounds *worklen).  With `defer`, points whose
// search the grid's search radius cannot certify are appended there (count
// in status[ST_NDEFER], zero on entry) for gdt_launch_deferred.
hipError_t gdt_launch_one_ring(int precision, const GdtSet* d_sets, int nsets,
                               const GdtChunk* chunks, int nch, int block, const int* worklist,
                               const uint32_t* worklen, int skip, int max_work,
                               const GdtRingIO& io, int* counts, int* status,
                               unsigned long long* counters, int* defer, hipStream_t s);

// Deferred points: the first GDT_WAVE_CAP one wavefront each (GDT_WAVE_BLOCKS
// blocks of 256 threads), the rest one lane each.
#define GDT_WAVE_CAP 8192
#define GDT_WAVE_BLOCKS 512
// Deferred points a tile block finishes itself (the rest go to the list).
#define GDT_TILE_DEFER 64

// Deferred points of the tile kernel (defer[0..status[ST_NDEFER]), count read
// on the device): the wavefront kernel, then one lane each beyond GDT_WAVE_CAP.
hipError_t gdt_launch_deferred(int precision, const GdtSet* d_sets, int nsets, int total_n,
                               const GdtRingIO& io, int* counts, int* status,
                               unsigned long long* counters, int* defer, hipStream_t s);

// The LDS-window kernels, compiled once per tile edge into namespaces ts4
// (16x16-cell tiles) and ts5 (32x32); gdt_tiles::* picks by the sets' shift.
#define GDT_TS_NS GDT_TS_NS_(GDT_TILE_SHIFT)
#define GDT_TS_NS_(x) GDT_TS_NS__(x)
#define GDT_TS_NS__(x) ts##x
#define GDT_DECLARE_TILE_LAUNCHERS(ns)                                                            \
    namespace ns {                                                                                \
    hipError_t gdt_launch_one_ring_tile(int precision, const GdtSet* d_sets, int nsets,          \
                                        const GdtChunk* tiles, int ntiles, int block, int lanes, \
                                        int total_n, const int* list, const uint32_t* lscan,     \
                                        const int* flags, const int* halt,                       \
                                        const GdtRingIO& io, int* counts, int* status,           \
                                        unsigned long long* counters, int* defer,                \
                                        hipEvent_t after_tile, hipStream_t s);                   \
    hipError_t gdt_launch_one_ring_star(const GdtSet* d_sets, int nsets, const GdtChunk* tiles,  \
                                        int ntiles, int nsplit, int block, int total_n,          \
                                        const GdtRingIO& io, int* counts, int* status,           \
                                        unsigned long long* counters, int* defer,                \
                                        hipEvent_t after_tile, hipStream_t s);                   \
    hipError_t gdt_launch_ring_test_tile(const GdtSet* d_sets, const GdtChunk* tiles, int ntiles,\
                                         const int* nbrs, const int* counts, int stride,         \
                                         int total_n, const double* d_xy, const GdtCells& c,     \
                                         int* repair, unsigned long long* partials,              \
                                         uint32_t* ntri, unsigned long long* counters,           \
                                         int* status, hipStream_t s);                            \
    }

// Triangles owned by their smallest vertex, written at toff (deterministic);
// counters[CT_INSUM] += sum of counts (partials holds npch words).  perm null:
// owners in caller order (toff over caller indices); otherwise in cell-slot
// order, owner of slot r = perm[r] (toff over slots).
hipError_t gdt_launch_triangles(const GdtSet* d_sets, const GdtChunk* pchunks, int npch,
                                int total_n, const int* perm, const int* nbrs, const int* counts,
                                int stride, uint32_t* tcount, uint32_t* toff, int* tris, int cap,
                                uint32_t* scan_tmp, unsigned long long* partials,
                                unsigned long long* counters, int* status, const int* halt,
                                hipStream_t s);

// Empty-circle test of *ntri triangles; repair must be -1 on entry.
hipError_t gdt_launch_test(const GdtSet* d_sets, int nsets, const int* tris, const uint32_t* ntri,
                           int max_tri, int total_n, const double* d_xy, const GdtCells& c, int* repair,
                           hipStream_t s);


// One-ring of every point (list and flags null), of the slots of a
// slot-ordered list (lscan its exclusive scan over slots), or of the points
// flagged in a caller-order array (flags[i] >= 0; each tile finds its own
// flagged slots, status[ST_NFLAGGED] += their count, which must start at zero;
// nothing with *halt set), one block per tile of the tiled cell order,
// `lanes` (1, 2 or 4) lanes per point.  Points the tile kernel cannot finish
// from LDS are deferred (gdt_launch_deferred over `defer`, size total_n, count
// in status[ST_NDEFER]).
//
// The same rings of every point with one wavefront per point (the star
// kernel, FP64 predicates): the lanes run the search steps from every
// candidate of the point's block at once and the ring is read off their
// results; blocks (tile, 0..nsplit-1) share each tile's points, `block`
// threads (a multiple of 64) each.
//
// Enumeration and empty-circle test in one pass over the rings, one block per
// tile (tiled cell order): triangles owned by their smallest caller index,
// flags as gdt_launch_test, counters[CT_INSUM] += sum of counts, *ntri =
// number of owned triangles; partials holds 2 * ntiles words.
GDT_DECLARE_TILE_LAUNCHERS(ts4)
GDT_DECLARE_TILE_LAUNCHERS(ts5)
static_assert(GDT_TSHIFT_MIN == 4 && GDT_TSHIFT_MAX == 5, "a launcher per supported tile edge");

namespace gdt_tiles {
inline hipError_t one_ring(int tshift, int precision, const GdtSet* d_sets, int nsets,
                           const GdtChunk* tiles, int ntiles, int block, int lanes, int total_n,
                           const int* list, const uint32_t* lscan, const int* flags,
                           const int* halt, const GdtRingIO& io, int* counts, int* status,
                           unsigned long long* counters, int* defer, hipEvent_t after_tile,
                           hipStream_t s)
{
    return (tshift == 4 ? ts4::gdt_launch_one_ring_tile : ts5::gdt_launch_one_ring_tile)(
        precision, d_sets, nsets, tiles, ntiles, block, lanes, total_n, list, lscan, flags, halt,
        io, counts, status, counters, defer, after_tile, s);
}
inline hipError_t one_ring_star(int tshift, const GdtSet* d_sets, int nsets,
                                const GdtChunk* tiles, int ntiles, int nsplit, int block,
                                int total_n, const GdtRingIO& io, int* counts, int* status,
                                unsigned long long* counters, int* defer, hipEvent_t after_tile,
                                hipStream_t s)
{
    return (tshift == 4 ? ts4::gdt_launch_one_ring_star : ts5::gdt_launch_one_ring_star)(
        d_sets, nsets, tiles, ntiles, nsplit, block, total_n, io, counts, status, counters, defer,
        after_tile, s);
}
inline hipError_t ring_test(int tshift, const GdtSet* d_sets, const GdtChunk* tiles, int ntiles,
                            const int* nbrs, const int* counts, int stride, int total_n,
                            const double* d_xy, const GdtCells& c, int* repair,
                            unsigned long long* partials, uint32_t* ntri,
                            unsigned long long* counters, int* status, hipStream_t s)
{
    return (tshift == 4 ? ts4::gdt_launch_ring_test_tile : ts5::gdt_launch_ring_test_tile)(
        d_sets, tiles, ntiles, nbrs, counts, stride, total_n, d_xy, c, repair, partials, ntri,
        counters, status, s);
}
}  // namespace gdt_tiles

// Slots of the flagged points, in slot order -> worklist; markscan (total_n+1
// words) = exclusive scan of the flags with the count at markscan[total_n];
// count -> *worklen (device).  With *halt set (halt non-null) no point is
// marked, so the repair that reads the list does nothing.
hipError_t gdt_launch_compact(const int* repair, const int* perm, int total_n, uint32_t* mark,
                              uint32_t* markscan, int* worklist, uint32_t* worklen,
                              uint32_t* scan_tmp, const int* halt, hipStream_t s);

// Voronoi energy (per-set sums in energy[nsets]; partial holds
// nsets * GDT_SUM_PARTS doubles, ecell total_n) and its adjoint (G holds
// 2 * total_n * stride doubles; gscale[nsets] scales each set, null = 1).
#define GDT_SUM_PARTS 256
hipError_t gdt_launch_voronoi_vertices(const GdtSet* d_sets, int nsets, int total_n,
                                       const double* xy, const int* nbrs, const int* counts,
                                       int stride, double* out, hipStream_t s);
hipError_t gdt_launch_voronoi_forward(int precision, const GdtSet* d_sets, int nsets, int total_n,
                                      const double* xy, const int* nbrs, const int* counts,
                                      int stride, double a0, double p0, double* area, double* perim,
                                      double* ecell, double* partial, double* energy,
                                      hipStream_t s);
hipError_t gdt_launch_voronoi_adjoint(int precision, const GdtSet* d_sets, int nsets, int total_n,
                                      const double* xy, const int* nbrs, const int* counts,
                                      int stride, double a0, double p0, const double* area,
                                      const double* perim, const double* gscale, double* G,
                                      double* grad, hipStream_t s);

// Brownian test-and-repair loop: state kept on the device across steps.
// The first step that fails (an error flag, a ring longer than the row
// stride, or an input that is not a triangulation) sets `halted`; every kernel
// of every later step then exits at once, so no step runs on truncated rows
// and the positions stay those of the failing step (its displacement done).
struct GdtMdState {
    int step;                         // step index of the next displacement
    int errflags;                     // OR of the steps' ERR_* flags
    int euler;                        // 1 if a step started from a non-triangulation
    int required;                     // longest ring seen
    int first_bad;                    // the failing step, -1 none
    int max_flagged;                  // most points repaired in one step
    int halted;                       // 1 from the failing step's end on
    int pad;
    unsigned long long flagged_total; // points repaired, summed over steps
};
hipError_t gdt_launch_md_init(GdtMdState* md, int step0, hipStream_t s);
hipError_t gdt_launch_md_displace(const GdtSet* d_sets, const GdtChunk* pchunks, int npch,
                                  double* xy, double sigma, unsigned long long seed,
                                  const GdtMdState* md, hipStream_t s);
hipError_t gdt_launch_md_step_end(const int* status, const unsigned long long* counters,
                                  const uint32_t* ntri, int total_n, int stride, GdtMdState* md,
                                  hipStream_t s);
