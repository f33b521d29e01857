// hostpath.cpp -- CPU mode of the library (DelaunayGPU::setGPUcompute(false)).
//
// The reference runs its per-point functions either as CUDA kernels or in
// OpenMP loops on the host (the `GPUcompute == false` branches of
// DelaunayGPU.cu:1036-1409 and DelaunayGPU.cpp:159-343, 567-706, with
// cellListGPU::compute binning on the host, cellListGPU.cpp:289-330).  This
// file does the same with this library's own routines: the per-point one-ring
// (gdt_point.h, and for points whose search reaches the grid's search radius
// the wavefront routine of gdt_wave.h), the triangle circle test
// (gdt_circle.h) and the Voronoi vertices (gdt_vor.h) are the code the
// kernels run, compiled here for the
// host (g++, -ffp-contract=off like the kernels) and called from OpenMP loops.
// Host pointers only; no HIP call is made.
#include "gdt_circle.h"
#include "gdt_exact_ring.h"
#include "gdt_vor.h"
#include "gdt_wave.h"
#include "gpudelaunay.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <new>
#include <vector>

namespace {

// The cell list in the kernels' layout (row-major cell keys): slot arrays in
// cell order, caller index order inside each cell.
struct HostCells {
    GdtSet S;
    std::vector<uint32_t> cstart;
    std::vector<int> perm;
    std::vector<gdt_d2> off64;
    std::vector<gdt_f2> off32;
    std::vector<uint32_t> scell;
    const double* xy = nullptr;     // the caller's coordinates (the exact mode reads them)

    GdtRingIO io(int* nbrs, int stride) const
    {
        GdtRingIO r;
        r.cstart = cstart.data();
        r.perm = perm.data();
        r.off64 = off64.data();
        r.off32 = off32.data();
        r.scell = scell.data();
        r.nbrs = nbrs;
        r.stride = stride;
        r.xy = reinterpret_cast<const gdt_d2*>(xy);
        return r;
    }
};

int nthreads_of(int nthreads) { return nthreads > 0 ? nthreads : omp_get_max_threads(); }

// A usable box {x11, x12, x21, x22}: finite and invertible.
bool box_ok(const double* b)
{
    for (int k = 0; k < 4; ++k)
        if (!std::isfinite(b[k])) return false;
    const double det = b[0] * b[3] - b[1] * b[2];
    return std::isfinite(det) && std::fabs(det) > 0.;
}

bool rect_ok(double lx, double ly) { return lx > 0. && ly > 0.; }

// Same checks and the same binning as the device build (gdt_make_set_box,
// gdt_bin_point), so both modes see the same cells.
int build_cells(HostCells& h, const double* xy, int n, const double* box, double a, int nt)
{
    if (n < 4 || !box_ok(box) || !(a > 0.) || !std::isfinite(a)) return GDT_ERR_INVALID;
    h.S = gdt_make_set_box(0, n, 0, box, a, 0);
    h.xy = xy;
    const GdtSet& S = h.S;
    if (S.xsize > 65535 || S.ysize > 65535 || S.xsize < 3 || S.ysize < 3) return GDT_ERR_INVALID;
    const int K = S.nkeys;
    try {
        std::vector<int> key((size_t)n);
        std::vector<uint32_t> cnt((size_t)K, 0u);
        h.cstart.assign((size_t)K + 1, 0u);
        h.perm.resize((size_t)n);
        h.off64.resize((size_t)n);
        h.off32.resize((size_t)n);
        h.scell.resize((size_t)n);
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int i = 0; i < n; ++i) {
            int cx, cy;
            double ox, oy;
            gdt_bin_point(S, xy[2 * (size_t)i], xy[2 * (size_t)i + 1], cx, cy, ox, oy);
            key[(size_t)i] = gdt_cell_key(S, cx, cy);
        }
        for (int i = 0; i < n; ++i) ++cnt[(size_t)key[(size_t)i]];
        for (int k = 0; k < K; ++k) h.cstart[(size_t)k + 1] = h.cstart[(size_t)k] + cnt[(size_t)k];
        std::vector<uint32_t> fill(h.cstart.begin(), h.cstart.end() - 1);
        for (int i = 0; i < n; ++i) h.perm[fill[(size_t)key[(size_t)i]]++] = i;
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int j = 0; j < n; ++j) {
            const int i = h.perm[(size_t)j];
            int cx, cy;
            double ox, oy;
            gdt_bin_point(S, xy[2 * (size_t)i], xy[2 * (size_t)i + 1], cx, cy, ox, oy);
            h.off64[(size_t)j].x = ox;
            h.off64[(size_t)j].y = oy;
            h.off32[(size_t)j].x = (float)ox;
            h.off32[(size_t)j].y = (float)oy;
            h.scell[(size_t)j] = (uint32_t)cx | ((uint32_t)cy << 16);
        }
    } catch (const std::bad_alloc&) {
        return GDT_ERR_NOMEM;
    }
    return GDT_OK;
}

int status_of(int err, int required, int stride)
{
    if (err & ERR_RANGE) return GDT_ERR_INVALID;
    if (err & (ERR_NOCLOSE | ERR_NONEIGHBOUR)) return GDT_ERR_NONGENERIC;
    if (err & ERR_SEARCH) return GDT_ERR_SEARCH;
    return required > stride ? GDT_OVERFLOW : GDT_OK;
}

// Rings of the slots `slots` (all slots when null), the kernels' per-point
// routine in the requested precision; a point whose search reaches the grid's
// search radius (boxes of a few cells) is redone by the routine the device
// path hands such points to (gdt_wave.h, its 64 lanes emulated), as the
// device's deferred points are (the exact mode, as on the device, reports it).  Returns the OR of the error bits and the
// longest ring in *maxc.
int rings_of(const HostCells& h, const int* slots, int nslots, int* nbrs, int* counts, int stride,
             int precision, int nt, int* maxc)
{
    const GdtRingIO io = h.io(nbrs, stride);
    // the FP32 filter's bounds are for rectangles (the exact mode reports a general box)
    if (h.S.general && precision != GDT_PREC_EXACT) precision = GDT_PREC_FP64;
    int mx = 0, err = 0;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 256) reduction(max : mx) reduction(| : err)
    for (int t = 0; t < nslots; ++t) {
        const int j = slots ? slots[t] : t;
        GdtRingResult r = precision == GDT_PREC_FP32F   ? gdt_one_ring_f32(h.S, j, io)
                          : precision == GDT_PREC_EXACT ? gdt_one_ring_exact(h.S, j, io)
                                                        : gdt_one_ring_f64(h.S, j, io);
        if ((r.err & ERR_SEARCH) && precision != GDT_PREC_EXACT) {
            GdtWaveHost w;
            r = gdt_wave_ring<true>(h.S, j, io, w, true);
        }
        counts[h.perm[(size_t)j]] = r.count;
        mx = std::max(mx, r.count);
        err |= r.err;
    }
    *maxc = mx;
    return err;
}

}  // namespace

extern "C" {

int gdt_host_celllist_general(const double* xy, int n, const double* box, double cell_size,
                              unsigned int* cell_sizes, int* idxs, int nmax, int* xsize,
                              int* ysize, int* max_occupancy)
{
    if (!xy || !box || (cell_sizes && (!idxs || nmax < 1))) return GDT_ERR_INVALID;
    HostCells h;
    const int rc = build_cells(h, xy, n, box, cell_size, nthreads_of(0));
    if (rc != GDT_OK) return rc;
    const GdtSet& S = h.S;
    int occ = 0;
    for (int k = 0; k < S.nkeys; ++k)
        occ = std::max(occ, (int)(h.cstart[(size_t)k + 1] - h.cstart[(size_t)k]));
    if (xsize) *xsize = S.xsize;
    if (ysize) *ysize = S.ysize;
    if (max_occupancy) *max_occupancy = occ;
    if (!cell_sizes) return GDT_OK;
    if (occ > nmax) return GDT_OVERFLOW;
    // reference layout: cell_sizes[cy*xsize+cx], idxs[cell*nmax + k], zero padded
    // (cellListGPU.h:93-102); row-major keys are exactly those cell indices
    for (int k = 0; k < S.nkeys; ++k) {
        const uint32_t b = h.cstart[(size_t)k], e = h.cstart[(size_t)k + 1];
        cell_sizes[k] = e - b;
        int* row = idxs + (size_t)k * (size_t)nmax;
        for (int m = 0; m < nmax; ++m) row[m] = m < (int)(e - b) ? h.perm[(size_t)b + m] : 0;
    }
    return GDT_OK;
}

int gdt_host_celllist(const double* xy, int n, double lx, double ly, double cell_size,
                      unsigned int* cell_sizes, int* idxs, int nmax, int* xsize, int* ysize,
                      int* max_occupancy)
{
    if (!rect_ok(lx, ly)) return GDT_ERR_INVALID;
    const double box[4] = {lx, 0., 0., ly};
    return gdt_host_celllist_general(xy, n, box, cell_size, cell_sizes, idxs, nmax, xsize, ysize,
                                     max_occupancy);
}

int gdt_host_triangulate_general(const double* xy, int n, const double* box, double cell_size,
                                 int stride, int* nbrs, int* counts, int precision, int nthreads,
                                 int* required)
{
    if (!xy || !box || !nbrs || !counts || stride < 3) return GDT_ERR_INVALID;
    if (precision != GDT_PREC_FP64 && precision != GDT_PREC_FP32F && precision != GDT_PREC_EXACT)
        return GDT_ERR_INVALID;
    const int nt = nthreads_of(nthreads);
    HostCells h;
    int rc = build_cells(h, xy, n, box, cell_size, nt);
    if (rc != GDT_OK) return rc;
    int mx = 0;
    const int err = rings_of(h, nullptr, n, nbrs, counts, stride, precision, nt, &mx);
    if (required) *required = mx;
    return status_of(err, mx, stride);
}

int gdt_host_triangulate(const double* xy, int n, double lx, double ly, double cell_size,
                         int stride, int* nbrs, int* counts, int precision, int nthreads,
                         int* required)
{
    if (!rect_ok(lx, ly)) return GDT_ERR_INVALID;
    const double box[4] = {lx, 0., 0., ly};
    return gdt_host_triangulate_general(xy, n, box, cell_size, stride, nbrs, counts, precision,
                                        nthreads, required);
}

int gdt_host_test_and_repair_general(const double* xy, int n, const double* box, double cell_size,
                                     int stride, int* nbrs, int* counts, int precision,
                                     int nthreads, int* nflagged, int* required)
{
    if (!xy || !box || !nbrs || !counts || stride < 3) return GDT_ERR_INVALID;
    if (precision != GDT_PREC_FP64 && precision != GDT_PREC_FP32F && precision != GDT_PREC_EXACT)
        return GDT_ERR_INVALID;
    const int nt = nthreads_of(nthreads);
    if (nflagged) *nflagged = 0;
    if (required) *required = 0;
    // 1. cell list of the moved points (DelaunayGPU.cpp:580-586)
    HostCells h;
    int rc = build_cells(h, xy, n, box, cell_size, nt);
    if (rc != GDT_OK) return rc;
    const GdtSet& S = h.S;
    const gdt_d2* p = reinterpret_cast<const gdt_d2*>(xy);
    const double ilx = 1.0 / S.lx, ily = 1.0 / S.ly;
    std::vector<unsigned char> flag;
    try {
        flag.assign((size_t)n, 0);
    } catch (const std::bad_alloc&) {
        return GDT_ERR_NOMEM;
    }
    // 2./3. triangles owned by their smallest vertex (getCircumcirclesCPU,
    // DelaunayGPU.cpp:673-706) and their empty-circle test (testTriangulation,
    // DelaunayGPU.cpp:607-648); the Euler counts of the input rings
    long long sum = 0, owned = 0;
    int malformed = 0;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 256) reduction(+ : sum, owned) reduction(| : malformed)
    for (int i = 0; i < n; ++i) {
        int c = counts[i];
        sum += c > 0 ? c : 0;
        if (c < 3 || c > stride) {
            malformed = 1;
            c = c > stride ? stride : (c < 0 ? 0 : c);
        }
        if (c == 0) continue;
        const int* row = nbrs + (size_t)i * (size_t)stride;
        int prev = row[c - 1];
        for (int m = 0; m < c; ++m) {
            const int cur = row[m];
            if (i < prev && i < cur) {
                ++owned;
                if ((unsigned)prev < (unsigned)n && (unsigned)cur < (unsigned)n) {
                    const GdtTriCircle cc = tri_circle(S, ilx, ily, p, i, prev, cur);
                    int intruder = -1;
                    bool bad = cc.bad;
                    if (!bad) {
                        bool covered;
                        intruder = tri_scan_global(S, p, i, cc, prev, cur, h.cstart.data(),
                                                   h.perm.data(), h.off64.data(), covered);
                        bad = intruder >= 0 || !covered;
                    }
                    if (bad) {
                        __atomic_store_n(&flag[(size_t)i], 1, __ATOMIC_RELAXED);
                        __atomic_store_n(&flag[(size_t)prev], 1, __ATOMIC_RELAXED);
                        __atomic_store_n(&flag[(size_t)cur], 1, __ATOMIC_RELAXED);
                        if (intruder >= 0) __atomic_store_n(&flag[(size_t)intruder], 1, __ATOMIC_RELAXED);
                    }
                }
            }
            prev = cur;
        }
    }
    if (malformed || sum != 6LL * n || owned != 2LL * n) return GDT_ERR_EULER;
    // 4. repair the flagged points, in slot order (locallyRepairDelaunayTriangulation,
    // DelaunayGPU.cpp:87-115)
    std::vector<int> slots;
    for (int j = 0; j < n; ++j)
        if (flag[(size_t)h.perm[(size_t)j]]) slots.push_back(j);
    if (nflagged) *nflagged = (int)slots.size();
    int mx = 0;
    const int err = rings_of(h, slots.data(), (int)slots.size(), nbrs, counts, stride, precision,
                             nt, &mx);
    if (required) *required = mx;
    return status_of(err, mx, stride);
}

int gdt_host_test_and_repair(const double* xy, int n, double lx, double ly, double cell_size,
                             int stride, int* nbrs, int* counts, int precision, int nthreads,
                             int* nflagged, int* required)
{
    if (!rect_ok(lx, ly)) return GDT_ERR_INVALID;
    const double box[4] = {lx, 0., 0., ly};
    return gdt_host_test_and_repair_general(xy, n, box, cell_size, stride, nbrs, counts, precision,
                                            nthreads, nflagged, required);
}

int gdt_host_voronoi_vertices_general(const double* xy, int n, const double* box,
                                      const int* nbrs, const int* counts, int stride,
                                      double* vertices, int nthreads)
{
    if (!xy || !box || !nbrs || !counts || !vertices || stride < 3 || n < 3 || !box_ok(box))
        return GDT_ERR_INVALID;
    using namespace gdt_vor;
    GdtSet S{};
    if (gdt_box_is_rect(box)) {
        S.lx = box[0];
        S.ly = box[3];
    } else {
        S = gdt_make_set_box(0, n, 0, box, std::numeric_limits<double>::infinity(), 0);
    }
    const int nt = nthreads_of(nthreads);
#pragma omp parallel for num_threads(nt) schedule(static)
    for (int i = 0; i < n; ++i) {
        const int k = counts[i];
        const int* row = nbrs + (size_t)i * (size_t)stride;
        double* o = vertices + 2 * (size_t)i * (size_t)stride;
        if (!vor_row_ok(row, k, stride, n)) {
            for (int m = 0; m < 2 * stride; ++m) o[m] = std::nan("");
            continue;
        }
        const V2<double> d0 = vor_round<double>(vor_disp(S, xy, i, row[0]));
        V2<double> a = d0;
        for (int m = 0; m < k; ++m) {
            const V2<double> b = m + 1 < k ? vor_round<double>(vor_disp(S, xy, i, row[m + 1])) : d0;
            const V2<double> c = vor_circ<double>(a, b);
            o[2 * m] = c.x;
            o[2 * m + 1] = c.y;
            a = b;
        }
    }
    return GDT_OK;
}

int gdt_host_voronoi_vertices(const double* xy, int n, double lx, double ly, const int* nbrs,
                              const int* counts, int stride, double* vertices, int nthreads)
{
    if (!rect_ok(lx, ly)) return GDT_ERR_INVALID;
    const double box[4] = {lx, 0., 0., ly};
    return gdt_host_voronoi_vertices_general(xy, n, box, nbrs, counts, stride, vertices, nthreads);
}

}  // extern "C"
