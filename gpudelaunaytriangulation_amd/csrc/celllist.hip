// celllist.hip -- counting-sort cell list for gfx950.
//
// Replaces the reference's padded-bucket cell list (cellListGPU.cu:20-58 with
// its host regrow loop, cellListGPU.cpp:404-460): no per-bucket capacity, no
// host readback.  Steps, all on one stream:
//   1. bin: cell key of every point, atomic arrival rank inside its cell;
//   2. exclusive scan of the per-cell counts (wave64 shuffle scan, two levels);
//   3. scatter the caller indices into cell order;
//   4. per cell: sort its few indices ascending (the reference's CPU binning
//      keeps index order inside a bucket, cellListGPU.cpp:306-326), then write
//      each slot's FP64/FP32 offsets inside the cell and the inverse permutation.
#include "gdt_kernels.h"

#define SCAN_ITEMS 8
#define SCAN_TILE (GDT_BLOCK * SCAN_ITEMS)

__global__ void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_i32(int* __restrict__ p, int v, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

static int fill_grid(size_t n)
{
    size_t b = (n + GDT_BLOCK - 1) / GDT_BLOCK;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return (int)b;
}

hipError_t gdt_fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_u32, dim3(fill_grid(n)), dim3(GDT_BLOCK), 0, s, p, v, n);
    return hipGetLastError();
}

hipError_t gdt_fill_i32(int* p, int v, size_t n, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_i32, dim3(fill_grid(n)), dim3(GDT_BLOCK), 0, s, p, v, n);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Exclusive scan of uint32 (per 2048-item tile: thread-serial over 8 items,
// wave64 shuffle scan, cross-wave via LDS; tiles combined recursively).
// ---------------------------------------------------------------------------
__device__ inline uint32_t block_exclusive(uint32_t v, uint32_t* lds, uint32_t& total)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) lds[wid] = x;
    __syncthreads();
    uint32_t woff = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < GDT_BLOCK / 64; ++w) {
        const uint32_t t = lds[w];
        if (w < wid) woff += t;
        tot += t;
    }
    total = tot;
    return woff + x - v;
}

__global__ __launch_bounds__(GDT_BLOCK) void k_scan_tiles(const uint32_t* __restrict__ in,
                                                          uint32_t* __restrict__ out, size_t n,
                                                          uint32_t* __restrict__ tile_sums)
{
    __shared__ uint32_t lds[GDT_BLOCK / 64];
    const size_t base = blockIdx.x * (size_t)SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t sum = 0;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; ++q) {
        v[q] = (base + q < n) ? in[base + q] : 0u;
        sum += v[q];
    }
    uint32_t total;
    uint32_t run = block_exclusive(sum, lds, total);
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; ++q) {
        if (base + q < n) out[base + q] = run;
        run += v[q];
    }
    if (threadIdx.x == 0) tile_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(GDT_BLOCK) void k_scan_add(uint32_t* __restrict__ out, size_t n,
                                                        const uint32_t* __restrict__ tile_off)
{
    const uint32_t add = tile_off[blockIdx.x];
    const size_t base = blockIdx.x * (size_t)SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
#pragma unroll
    for (int q = 0; q < SCAN_ITEMS; ++q)
        if (base + q < n) out[base + q] += add;
}

__global__ void k_copy_u32(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = *src;
}

size_t gdt_scan_tmp_words(size_t n)
{
    size_t words = 0;
    while (true) {
        const size_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
        words += 2 * tiles + 64;
        if (tiles <= 1) break;
        n = tiles;
    }
    return words;
}

hipError_t gdt_scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* total,
                        uint32_t* tmp, hipStream_t s)
{
    if (n == 0) return gdt_fill_u32(total, 0u, 1, s);
    const size_t tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    uint32_t* sums = tmp;
    uint32_t* offs = tmp + tiles + 32;
    hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)tiles), dim3(GDT_BLOCK), 0, s, in, out, n, sums);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (tiles == 1) {
        hipLaunchKernelGGL(k_copy_u32, dim3(1), dim3(64), 0, s, sums, total);
        return hipGetLastError();
    }
    e = gdt_scan_u32(sums, offs, tiles, total, tmp + 2 * tiles + 64, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scan_add, dim3((unsigned)tiles), dim3(GDT_BLOCK), 0, s, out, n, offs);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Cell list kernels
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(GDT_BLOCK) void k_bin(const GdtSet* __restrict__ sets,
                                                   const GdtChunk* __restrict__ chunks,
                                                   const gdt_d2* __restrict__ xy,
                                                   uint32_t* __restrict__ pkey,
                                                   uint32_t* __restrict__ prank,
                                                   uint32_t* __restrict__ ccount)
{
    const GdtChunk ch = chunks[blockIdx.x];
    const GdtSet S = sets[ch.set];
    const int li = ch.start + (int)threadIdx.x;
    if (li >= S.n) return;
    const int i = S.p0 + li;
    const gdt_d2 p = xy[i];
    int cx, cy;
    double ox, oy;
    gdt_bin_point(S, p.x, p.y, cx, cy, ox, oy);
    const uint32_t key = (uint32_t)(S.c0 + gdt_cell_key(S, cx, cy));
    pkey[i] = key;
    prank[i] = atomicAdd(&ccount[key], 1u);
}

__global__ __launch_bounds__(GDT_BLOCK) void k_scatter(const GdtSet* __restrict__ sets,
                                                       const GdtChunk* __restrict__ chunks,
                                                       const uint32_t* __restrict__ pkey,
                                                       const uint32_t* __restrict__ prank,
                                                       const uint32_t* __restrict__ cstart,
                                                       int* __restrict__ perm)
{
    const GdtChunk ch = chunks[blockIdx.x];
    const GdtSet S = sets[ch.set];
    const int li = ch.start + (int)threadIdx.x;
    if (li >= S.n) return;
    const int i = S.p0 + li;
    perm[cstart[pkey[i]] + prank[i]] = i;
}

__global__ __launch_bounds__(GDT_BLOCK) void k_finalize(const GdtSet* __restrict__ sets,
                                                        const GdtChunk* __restrict__ chunks,
                                                        const uint32_t* __restrict__ cstart,
                                                        const gdt_d2* __restrict__ xy,
                                                        int* __restrict__ perm,
                                                        int* __restrict__ slot_of,
                                                        gdt_d2* __restrict__ off64,
                                                        gdt_f2* __restrict__ off32,
                                                        uint32_t* __restrict__ scell)
{
    const GdtChunk ch = chunks[blockIdx.x];
    const GdtSet S = sets[ch.set];
    const int lk = ch.start + (int)threadIdx.x;
    if (lk >= S.nkeys) return;
    const int key = S.c0 + lk;
    const uint32_t b = cstart[key], e = cstart[key + 1];
    if (b == e) return;
    // insertion sort of the cell's caller indices (cells hold a few points)
    for (uint32_t s = b + 1; s < e; ++s) {
        const int v = perm[s];
        uint32_t t = s;
        while (t > b && perm[t - 1] > v) {
            perm[t] = perm[t - 1];
            --t;
        }
        perm[t] = v;
    }
    for (uint32_t s = b; s < e; ++s) {
        const int i = perm[s];
        const gdt_d2 p = xy[i];
        int cx, cy;
        double ox, oy;
        gdt_bin_point(S, p.x, p.y, cx, cy, ox, oy);
        gdt_d2 o64;
        o64.x = ox;
        o64.y = oy;
        gdt_f2 o32;
        o32.x = (float)ox;
        o32.y = (float)oy;
        off64[s] = o64;
        off32[s] = o32;
        scell[s] = (uint32_t)cx | ((uint32_t)cy << 16);
        slot_of[i] = (int)s;
    }
}

hipError_t gdt_build_cells(const GdtSet* d_sets, const GdtChunk* pchunks, int npch,
                           const GdtChunk* kchunks, int nkch, int total_n, int total_keys,
                           const double* d_xy, GdtCells& c, hipStream_t s)
{
    hipError_t e = gdt_fill_u32(c.ccount, 0u, (size_t)total_keys, s);
    if (e != hipSuccess) return e;
    const gdt_d2* xy = reinterpret_cast<const gdt_d2*>(d_xy);
    hipLaunchKernelGGL(k_bin, dim3(npch), dim3(GDT_BLOCK), 0, s, d_sets, pchunks, xy, c.pkey,
                       c.prank, c.ccount);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    e = gdt_scan_u32(c.ccount, c.cstart, (size_t)total_keys, c.cstart + total_keys, c.scan_tmp, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scatter, dim3(npch), dim3(GDT_BLOCK), 0, s, d_sets, pchunks, c.pkey,
                       c.prank, c.cstart, c.perm);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_finalize, dim3(nkch), dim3(GDT_BLOCK), 0, s, d_sets, kchunks, c.cstart,
                       xy, c.perm, c.slot_of, c.off64, c.off32, c.scell);
    (void)total_n;
    return hipGetLastError();
}

// Reference layout: cell_sizes[cy*xsize+cx], idxs[cell*nmax + k] (zero padded).
__global__ __launch_bounds__(GDT_BLOCK) void k_export(const GdtSet* __restrict__ sets,
                                                      const uint32_t* __restrict__ cstart,
                                                      const int* __restrict__ perm,
                                                      unsigned int* __restrict__ cell_sizes,
                                                      int* __restrict__ idxs, int nmax)
{
    const GdtSet S = sets[0];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= S.xsize * S.ysize) return;
    const int cx = c % S.xsize, cy = c / S.xsize;
    const int key = S.c0 + gdt_cell_key(S, cx, cy);
    const uint32_t b = cstart[key], e = cstart[key + 1];
    cell_sizes[c] = e - b;
    for (int k = 0; k < nmax; ++k)
        idxs[(size_t)c * nmax + k] = (b + (uint32_t)k < e) ? perm[b + k] : 0;
}

hipError_t gdt_export_cells(const GdtSet* d_sets, const GdtSet& h_set, const GdtCells& c,
                            unsigned int* cell_sizes, int* idxs, int nmax, hipStream_t s)
{
    const int nc = h_set.xsize * h_set.ysize;
    hipLaunchKernelGGL(k_export, dim3((nc + GDT_BLOCK - 1) / GDT_BLOCK), dim3(GDT_BLOCK), 0, s,
                       d_sets, c.cstart, c.perm, cell_sizes, idxs, nmax);
    return hipGetLastError();
}

__global__ __launch_bounds__(GDT_BLOCK) void k_max_u32(const uint32_t* __restrict__ v, int n,
                                                       int* __restrict__ out)
{
    int m = 0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        m = max(m, (int)v[i]);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

hipError_t gdt_max_occupancy(const uint32_t* ccount, int total_keys, int* d_out, hipStream_t s)
{
    hipError_t e = gdt_fill_i32(d_out, 0, 1, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_max_u32, dim3(fill_grid((size_t)total_keys)), dim3(GDT_BLOCK), 0, s,
                       ccount, total_keys, d_out);
    return hipGetLastError();
}
