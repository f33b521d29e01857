// testrepair.hip -- triangle enumeration, empty-circle test and flag
// compaction for the device-only test-and-repair pipeline.
//
// Replaces gpu_get_circumcircles_kernel (DelaunayGPU.cu:1008-1028), which
// appends triangles through one global atomic counter (order varies run to
// run), by a count + exclusive scan + write (deterministic order), and
// gpu_test_circumcircles_kernel (DelaunayGPU.cu:290-313).
#include "gdt_circle.h"
#include "gdt_kernels.h"

// Owned triangles per point, the block's sum of ring lengths into
// partials[block] (summed by k_sum_partials: no contended atomics) and the
// malformed-row flag.  Item r = p0 + local index is the point itself (perm
// null: caller order) or the cell-ordered slot r holding point perm[r].
__global__ __launch_bounds__(GDT_BLOCK) void k_tri_count(const GdtSet* __restrict__ sets,
                                                         const GdtChunk* __restrict__ chunks,
                                                         const int* __restrict__ perm,
                                                         const int* __restrict__ nbrs,
                                                         const int* __restrict__ counts, int stride,
                                                         uint32_t* __restrict__ tcount,
                                                         unsigned long long* __restrict__ partials,
                                                         int* __restrict__ status,
                                                         const int* __restrict__ halt)
{
    if (halt && *halt) return;
    __shared__ unsigned long long s_tot[GDT_BLOCK / 64];
    const GdtChunk ch = chunks[blockIdx.x];
    const GdtSet S = sets[ch.set];
    const int li = ch.start + (int)threadIdx.x;
    unsigned cnt = 0, tot = 0;
    int bad = 0;
    if (li < S.n) {
        const int r = S.p0 + li;
        const int i = perm ? perm[r] : r;
        int c = counts[i];
        tot = c > 0 ? (unsigned)c : 0u;
        if (c > stride || c < 3) { bad = 1; c = c > stride ? stride : (c < 0 ? 0 : c); }
        if (c > 0) {
            const int* row = nbrs + (size_t)i * stride;
            int prev = row[c - 1];
            for (int m = 0; m < c; ++m) {
                const int cur = row[m];
                cnt += (i < prev && i < cur) ? 1u : 0u;
                prev = cur;
            }
        }
        tcount[r] = cnt;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        tot += __shfl_xor(tot, o, 64);
        bad |= __shfl_xor(bad, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        s_tot[threadIdx.x >> 6] = tot;
        if (bad) atomicOr(&status[ST_ERRFLAGS], ERR_NOCLOSE);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s_tot[w];
        partials[blockIdx.x] = t;
    }
}

// One block: counters[CT_INSUM] += sum of partials[0..n) (fixed order).
__global__ __launch_bounds__(256) void k_sum_partials(const unsigned long long* __restrict__ partials,
                                                      int n, unsigned long long* __restrict__ counters)
{
    __shared__ unsigned long long s[256];
    unsigned long long acc = 0;
    for (int i = (int)threadIdx.x; i < n; i += (int)blockDim.x) acc += partials[i];
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) counters[CT_INSUM] += s[0];
}

__global__ __launch_bounds__(GDT_BLOCK) void k_tri_write(const GdtSet* __restrict__ sets,
                                                         const GdtChunk* __restrict__ chunks,
                                                         const int* __restrict__ perm,
                                                         const int* __restrict__ nbrs,
                                                         const int* __restrict__ counts, int stride,
                                                         const uint32_t* __restrict__ toff,
                                                         int* __restrict__ tris, int cap,
                                                         const int* __restrict__ halt)
{
    if (halt && *halt) return;
    const GdtChunk ch = chunks[blockIdx.x];
    const GdtSet S = sets[ch.set];
    const int li = ch.start + (int)threadIdx.x;
    if (li >= S.n) return;
    const int r = S.p0 + li;
    const int i = perm ? perm[r] : r;
    int c = counts[i];
    if (c > stride) c = stride;
    if (c <= 0) return;
    const int* row = nbrs + (size_t)i * stride;
    uint32_t t = toff[r];
    int prev = row[c - 1];
    for (int m = 0; m < c; ++m) {
        const int cur = row[m];
        if (i < prev && i < cur) {
            if (t < (uint32_t)cap) {
                tris[3 * (size_t)t] = i;
                tris[3 * (size_t)t + 1] = prev;
                tris[3 * (size_t)t + 2] = cur;
            }
            ++t;
        }
        prev = cur;
    }
}

hipError_t gdt_launch_triangles(const GdtSet* d_sets, const GdtChunk* pchunks, int npch,
                                int total_n, const int* perm, const int* nbrs, const int* counts,
                                int stride, uint32_t* tcount, uint32_t* toff, int* tris, int cap,
                                uint32_t* scan_tmp, unsigned long long* partials,
                                unsigned long long* counters, int* status, const int* halt,
                                hipStream_t s)
{
    hipLaunchKernelGGL(k_tri_count, dim3(npch), dim3(GDT_BLOCK), 0, s, d_sets, pchunks, perm, nbrs,
                       counts, stride, tcount, partials, status, halt);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, partials, npch, counters);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = gdt_scan_u32(tcount, toff, (size_t)total_n, toff + total_n, scan_tmp, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_tri_write, dim3(npch), dim3(GDT_BLOCK), 0, s, d_sets, pchunks, perm, nbrs,
                       counts, stride, toff, tris, cap, halt);
    return hipGetLastError();
}

// The circle of a triangle and its scan over the cell list (tri_circle,
// tri_scan_global, tri_flag) are shared with the CPU mode: gdt_circle.h.

__global__ __launch_bounds__(GDT_BLOCK) void k_test(const GdtSet* __restrict__ sets, int nsets,
                                                    const int* __restrict__ tris,
                                                    const uint32_t* __restrict__ ntri, int cap,
                                                    int total_n, const gdt_d2* __restrict__ xy,
                                                    const uint32_t* __restrict__ cstart,
                                                    const int* __restrict__ perm,
                                                    const gdt_d2* __restrict__ off64,
                                                    int* __restrict__ repair,
                                                    const int* __restrict__ halt)
{
    if (halt && *halt) return;
    const int nt = min((int)*ntri, cap);
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
        const int i = tris[3 * (size_t)t], a = tris[3 * (size_t)t + 1], b = tris[3 * (size_t)t + 2];
        if ((unsigned)i >= (unsigned)total_n || (unsigned)a >= (unsigned)total_n ||
            (unsigned)b >= (unsigned)total_n)
            continue;   // malformed ring entry: the Euler check reports it
        const GdtSet S = sets[gdt_set_of(sets, nsets, i)];
        const GdtTriCircle c = tri_circle(S, 1.0 / S.lx, 1.0 / S.ly, xy, i, a, b);
        bool bad = c.bad;
        int intruder = -1;
        if (!bad) {
            bool covered;
            intruder = tri_scan_global(S, xy, i, c, a, b, cstart, perm, off64, covered);
            bad = intruder >= 0 || !covered;
        }
        if (bad) tri_flag(repair, i, a, b, intruder);
    }
}

hipError_t gdt_launch_test(const GdtSet* d_sets, int nsets, const int* tris, const uint32_t* ntri,
                           int max_tri, int total_n, const double* d_xy, const GdtCells& c, int* repair,
                           hipStream_t s)
{
    if (max_tri <= 0) return hipSuccess;
    int grid = (max_tri + GDT_BLOCK - 1) / GDT_BLOCK;
    if (grid > 8192) grid = 8192;
    hipLaunchKernelGGL(k_test, dim3(grid), dim3(GDT_BLOCK), 0, s, d_sets, nsets, tris, ntri,
                       max_tri, total_n, reinterpret_cast<const gdt_d2*>(d_xy), c.cstart, c.perm, c.off64,
                       repair, c.halt);
    return hipGetLastError();
}

// The enumeration and test of a tile in one pass over the rings (the LDS
// window kernel, compiled once per tile edge) is in ringtest.hip.

__global__ __launch_bounds__(GDT_BLOCK) void k_mark(const int* __restrict__ repair,
                                                    const int* __restrict__ perm, int n,
                                                    uint32_t* __restrict__ mark,
                                                    const int* __restrict__ halt)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool off = halt && *halt;
    if (s < n) mark[s] = !off && repair[perm[s]] >= 0 ? 1u : 0u;
}

// Flagged slots, in slot order, to worklist[pos[s]]; *worklen = pos[n].
__global__ __launch_bounds__(GDT_BLOCK) void k_compact(const uint32_t* __restrict__ mark, int n,
                                                       const uint32_t* __restrict__ pos,
                                                       int* __restrict__ worklist,
                                                       uint32_t* __restrict__ worklen)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n && mark[s]) worklist[pos[s]] = s;
    if (s == 0) *worklen = pos[n];
}

hipError_t gdt_launch_compact(const int* repair, const int* perm, int total_n, uint32_t* mark,
                              uint32_t* markscan, int* worklist, uint32_t* worklen,
                              uint32_t* scan_tmp, const int* halt, hipStream_t s)
{
    const int g = (total_n + GDT_BLOCK - 1) / GDT_BLOCK;
    hipLaunchKernelGGL(k_mark, dim3(g), dim3(GDT_BLOCK), 0, s, repair, perm, total_n, mark, halt);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    e = gdt_scan_u32(mark, markscan, (size_t)total_n, markscan + total_n, scan_tmp, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_compact, dim3(g), dim3(GDT_BLOCK), 0, s, mark, total_n, markscan,
                       worklist, worklen);
    return hipGetLastError();
}
