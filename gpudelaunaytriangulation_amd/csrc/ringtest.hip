// ringtest.hip -- triangle enumeration and empty-circle test of one tile of
// cells per block, against the tile's window staged in LDS (the test of
// test-and-repair and of the Brownian step in tiled cell order).
//
// Compiled once per supported tile edge (GDT_TILE_SHIFT, GDT_WIN_MAXPTS from
// the compile line) into namespace ts<shift>, like tile.hip.  Replaces
// gpu_get_circumcircles_kernel + gpu_test_circumcircles_kernel
// (DelaunayGPU.cu:1008-1028, 290-313) for tiled cell orders.
#include "gdt_circle.h"
#include "gdt_kernels.h"
#include "gdt_stage.h"

namespace GDT_TS_NS {

// Test of triangle (i, a, b), owned by the point in slot si of the tile whose
// window L holds (when `fits`): the window cells of the circle's conservative
// bounding box are scanned in LDS, and a circle whose box leaves the window, or
// a window too full for LDS, takes the global scan.  Flags the repair rows.
__device__ inline void tri_test_in_tile(const GdtSet& S, double ilx, double ily,
                                        const GdtTileWin& w, bool fits,
                                        const GdtTileLds& L, const gdt_d2* __restrict__ xy,
                                        const GdtRingIO& io, int i, int a, int b, int si,
                                        int* __restrict__ repair)
{
    const uint32_t* s_cs = L.cs;
    const int* s_gstart = L.gstart;
    const gdt_d2* s_X64 = L.X64;
    const int* s_orig = L.orig;
    {
        const GdtTriCircle c = tri_circle(S, ilx, ily, xy, i, a, b);
        if (c.bad) {
            tri_flag(repair, i, a, b, -1);
            return;
        }
        bool done = false, bad = false;
        int intruder = -1;
        if (fits) {
            // the owner's window cell and frame coordinate; the circle's box
            const uint32_t pc = io.scell[si];
            const int wx = (int)(pc & 0xffffu) - w.gx0, wy = (int)(pc >> 16) - w.gy0;
            const int cv = wy * GDT_WIN + wx;
            const gdt_d2 Xv = s_X64[s_cs[cv] + (si - s_gstart[cv])];
            const double cx = Xv.x + c.qx, cy = Xv.y + c.qy;
            const double R = gdt_sqrt64(c.R2) * (1.0 + 1e-6);
            const int x0 = (int)floor((cx - R) * S.ibsx - 1e-3) + GDT_FRAME;
            const int x1 = (int)floor((cx + R) * S.ibsx + 1e-3) + GDT_FRAME;
            const int y0 = (int)floor((cy - R) * S.ibsy - 1e-3) + GDT_FRAME;
            const int y1 = (int)floor((cy + R) * S.ibsy + 1e-3) + GDT_FRAME;
            if (x0 >= 0 && x1 < GDT_WIN && y0 >= 0 && y1 < GDT_WIN && x1 - x0 <= 2 * S.rmax &&
                y1 - y0 <= 2 * S.rmax) {
                for (int y = y0; y <= y1 && intruder < 0; ++y) {
                    const uint32_t e = s_cs[y * GDT_WIN + x1 + 1];
                    for (uint32_t k = s_cs[y * GDT_WIN + x0]; k < e; ++k) {
                        const gdt_d2 X = s_X64[k];
                        const double ex = X.x - cx, ey = X.y - cy;
                        if (ex * ex + ey * ey < c.lim) {
                            const int o = s_orig[k];
                            if (o != i && o != a && o != b) { intruder = o; break; }
                        }
                    }
                }
                bad = intruder >= 0;
                done = true;
            }
        }
        if (!done) {
            bool covered;
            intruder = tri_scan_global(S, xy, i, c, a, b, io.cstart, io.perm, io.off64, covered);
            bad = intruder >= 0 || !covered;
        }
        if (bad) tri_flag(repair, i, a, b, intruder);
    }
}

// Enumeration and test in one pass over the rings (tiled cell order), one
// block per tile: the block stages the tile's window, then, for a block's
// worth of the tile's points at a time, lists in LDS the triangles each point
// owns -- those whose other two vertices have larger caller indices, the
// ownership rule of k_tri_count -- and tests the listed triangles with all
// lanes (tri_test_in_tile); when the list fills up, the lanes go on from where
// they stopped in further lists.  partials[2 tile] = the tile's ring entries,
// partials[2 tile + 1] = its owned triangles (for the Euler check).
// (the list holds the owned triangles of a block's worth of points, about two each)
#define GDT_RT_THREADS (GDT_TILE_CELLS < 512 ? GDT_TILE_CELLS : 512)
#define GDT_RT_LCAP (3 * GDT_RT_THREADS)
__global__ __launch_bounds__(GDT_RT_THREADS) void k_ring_test_tile(const GdtSet* __restrict__ sets,
                                                        const GdtChunk* __restrict__ tiles,
                                                        const int* __restrict__ nbrs,
                                                        const int* __restrict__ counts, int stride,
                                                        int total_n, const gdt_d2* __restrict__ xy,
                                                        GdtRingIO io,
                                                        int* __restrict__ repair,
                                                        unsigned long long* __restrict__ partials,
                                                        int* __restrict__ status,
                                                        const int* __restrict__ halt)
{
    if (halt && *halt) return;
    __shared__ uint32_t s_cs[GDT_WIN_CELLS + 1];
    __shared__ int s_gstart[GDT_WIN_CELLS];
    __shared__ gdt_d2 s_X64[GDT_WIN_MAXPTS];
    __shared__ int s_orig[GDT_WIN_MAXPTS];
    __shared__ uint32_t s_wsum[16];
    __shared__ int4 s_tri[GDT_RT_LCAP];
    __shared__ unsigned s_tn;
    __shared__ unsigned long long s_red[2][8];

    const int tid = (int)threadIdx.x, nt = (int)blockDim.x;
    const GdtChunk ch = tiles[blockIdx.x];
    const GdtSet S = sets[ch.set];
    const double ilx = 1.0 / S.lx, ily = 1.0 / S.ly;
    const int k0 = S.c0 + ch.start * GDT_TILE_CELLS;
    const int T0 = (int)io.cstart[k0], T1 = (int)io.cstart[k0 + GDT_TILE_CELLS];
    if (T0 >= T1) {     // block-uniform: an empty tile
        if (tid == 0) {
            partials[2 * (size_t)blockIdx.x] = 0ull;
            partials[2 * (size_t)blockIdx.x + 1] = 0ull;
        }
        return;
    }
    const GdtTileWin w = gdt_tile_window(S, ch.start);
    GdtTileLds L;
    L.cs = s_cs; L.gstart = s_gstart; L.X64 = s_X64; L.X32 = nullptr; L.orig = s_orig;
    const bool fits = gdt_stage_window<false>(S, w, io, L, s_wsum);

    unsigned long long tot = 0, own = 0;
    int bad = 0;
    for (int r0 = T0; r0 < T1; r0 += nt) {
        // this round's point of the lane: its ring, walked from entry m
        const int r = r0 + tid;
        int i = 0, c = 0, m = 0, prev = 0;
        const int* row = nbrs;
        if (r < T1) {
            i = io.perm[r];
            c = counts[i];
            tot += c > 0 ? (unsigned long long)c : 0ull;
            if (c > stride || c < 3) { bad = 1; c = c > stride ? stride : (c < 0 ? 0 : c); }
            row = nbrs + (size_t)i * stride;
            if (c > 0) prev = row[c - 1];
        }
        bool pending = m < c;
        for (;;) {
            if (tid == 0) s_tn = 0u;
            __syncthreads();
            if (pending) {
                for (; m < c; ++m) {
                    const int cur = row[m];
                    if (i < prev && i < cur) {
                        if ((unsigned)prev < (unsigned)total_n && (unsigned)cur < (unsigned)total_n) {
                            const unsigned pos = atomicAdd(&s_tn, 1u);
                            if (pos >= GDT_RT_LCAP) break;     // full: this one goes in the next list
                            s_tri[pos] = make_int4(i, prev, cur, r);
                        }
                        ++own;
                    }
                    prev = cur;
                }
                pending = m < c;
            }
            __syncthreads();
            const int nl = min((int)s_tn, GDT_RT_LCAP);
            for (int t = tid; t < nl; t += nt) {
                const int4 e = s_tri[t];
                tri_test_in_tile(S, ilx, ily, w, fits, L, xy, io, e.x, e.y, e.z, e.w, repair);
            }
            // also the barrier before the list is rewritten
            if (!__syncthreads_or(pending)) break;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        tot += __shfl_xor(tot, o, 64);
        own += __shfl_xor(own, o, 64);
        bad |= __shfl_xor(bad, o, 64);
    }
    if ((tid & 63) == 0) {
        s_red[0][tid >> 6] = tot;
        s_red[1][tid >> 6] = own;
        if (bad) atomicOr(&status[ST_ERRFLAGS], ERR_NOCLOSE);
    }
    __syncthreads();
    if (tid == 0) {
        unsigned long long a = 0, b = 0;
        for (int v = 0; v < (nt >> 6); ++v) {
            a += s_red[0][v];
            b += s_red[1][v];
        }
        partials[2 * (size_t)blockIdx.x] = a;
        partials[2 * (size_t)blockIdx.x + 1] = b;
    }
}

// One block: counters[CT_INSUM] += sum of the tiles' ring entries, *ntri = sum
// of their owned triangles (fixed order).
__global__ __launch_bounds__(256) void k_sum_ring_partials(const unsigned long long* __restrict__ partials,
                                                           int ntiles,
                                                           unsigned long long* __restrict__ counters,
                                                           uint32_t* __restrict__ ntri)
{
    __shared__ unsigned long long s[2][256];
    unsigned long long a = 0, b = 0;
    for (int t = (int)threadIdx.x; t < ntiles; t += (int)blockDim.x) {
        a += partials[2 * (size_t)t];
        b += partials[2 * (size_t)t + 1];
    }
    s[0][threadIdx.x] = a;
    s[1][threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s[0][threadIdx.x] += s[0][threadIdx.x + w];
            s[1][threadIdx.x] += s[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        counters[CT_INSUM] += s[0][0];
        *ntri = (uint32_t)s[1][0];
    }
}

hipError_t gdt_launch_ring_test_tile(const GdtSet* d_sets, const GdtChunk* tiles, int ntiles,
                                     const int* nbrs, const int* counts, int stride, int total_n,
                                     const double* d_xy, const GdtCells& c, int* repair,
                                     unsigned long long* partials, uint32_t* ntri,
                                     unsigned long long* counters, int* status, hipStream_t s)
{
    if (ntiles <= 0) return hipSuccess;
    GdtRingIO io;
    io.cstart = c.cstart;
    io.perm = c.perm;
    io.off64 = c.off64;
    io.off32 = c.off32;
    io.scell = c.scell;
    io.nbrs = nullptr;
    io.stride = 0;
    io.xy = nullptr;
    hipLaunchKernelGGL(k_ring_test_tile, dim3(ntiles), dim3(GDT_RT_THREADS), 0, s, d_sets, tiles, nbrs,
                       counts, stride, total_n, reinterpret_cast<const gdt_d2*>(d_xy), io,
                       repair, partials, status, c.halt);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sum_ring_partials, dim3(1), dim3(256), 0, s, partials, ntiles, counters,
                       ntri);
    return hipGetLastError();
}


}  // namespace GDT_TS_NS
