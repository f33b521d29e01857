// celllist.hip -- counting-sort cell list for gfx950.
//
// Replaces the reference's padded-bucket cell list (cellListGPU.cu:20-58 with
// its host regrow loop, cellListGPU.cpp:404-460): no per-bucket capacity, no
// host readback.  Slot order is cell-key order, and inside a cell ascending
// caller index (the reference's CPU binning keeps index order inside a bucket,
// cellListGPU.cpp:306-326).
//
// Two-level counting sort in four kernels, every pass reading or writing in
// runs:
//   1. k_kb_count: blocks of consecutive points histogram their key buckets
//      (2^B consecutive cell keys, B >= 10) in LDS; one count per (bucket,
//      block);
//   2. k_kb_rowscan: per bucket, the exclusive scan of its block counts and
//      the bucket total;
//   3. k_kb_scatter: each block scans the bucket totals in LDS (block 0 also
//      stores them), bins its points again and appends one 32-byte record
//      (offset, caller index, key, packed cell) to its run of each bucket;
//   4. k_kb_finalize: one block per bucket counting-sorts its points by cell
//      in LDS, orders each cell by caller index, and writes cstart/ccount and
//      the slot arrays (perm, slot_of, off64, off32, scell) of its key range
//      in slot order; a bucket above KB_CAP points (dense clusters) takes the
//      same steps through global memory.
// Up to 16,384 points over at most 64 buckets the four passes run as one
// launch (k_kb_small), which writes the same slots.
// More than 2^25 keys (beyond the LDS histograms) take the single-level path:
// per-point atomic arrival ranks, scan, scatter, per-cell sort.
#include "gdt_kernels.h"

#include <algorithm>

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

// Status words and counters: zeroed, and copied to pinned host memory, by one
// small kernel each (one launch instead of two fills or two copies).
__global__ void k_status_reset(int* __restrict__ st, int nst, unsigned long long* __restrict__ ct,
                               int nct)
{
    const int t = threadIdx.x;
    if (t < nst) st[t] = 0;
    if (t < nct) ct[t] = 0ull;
}

__global__ void k_status_publish(const int* __restrict__ st, int nst,
                                 const unsigned long long* __restrict__ ct, int nct,
                                 int* __restrict__ hst, unsigned long long* __restrict__ hct,
                                 unsigned int* __restrict__ hseq, unsigned int seq)
{
    const int t = threadIdx.x;
    if (t < nst) hst[t] = st[t];
    if (t < nct) hct[t] = ct[t];
    __threadfence_system();
    __syncthreads();
    if (t == 0) __hip_atomic_store(hseq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

hipError_t gdt_status_reset(int* st, int nst, unsigned long long* ct, int nct, hipStream_t s)
{
    hipLaunchKernelGGL(k_status_reset, dim3(1), dim3(64), 0, s, st, nst, ct, nct);
    return hipGetLastError();
}

hipError_t gdt_status_publish(const int* st, int nst, const unsigned long long* ct, int nct,
                              int* hst, unsigned long long* hct, unsigned int* hseq,
                              unsigned int seq, hipStream_t s)
{
    hipLaunchKernelGGL(k_status_publish, dim3(1), dim3(64), 0, s, st, nst, ct, nct, hst, hct,
                       hseq, seq);
    return hipGetLastError();
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

hipError_t gdt_copy_u32(const uint32_t* src, uint32_t* dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_copy_u32, dim3(1), dim3(64), 0, s, src, dst);
    return hipGetLastError();
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
                                                   uint32_t* __restrict__ ccount,
                                                   const int* __restrict__ halt)
{
    if (halt && *halt) return;
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
                                                       int* __restrict__ perm,
                                                       const int* __restrict__ halt)
{
    if (halt && *halt) return;
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
                                                        uint32_t* __restrict__ scell,
                                                        const int* __restrict__ halt)
{
    if (halt && *halt) return;
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
        if (slot_of) slot_of[i] = (int)s;
    }
}

// ---------------------------------------------------------------------------
// Two-level build
// ---------------------------------------------------------------------------
#define KB_THREADS 1024
#define KB_NB 256            // point blocks of passes 1 and 3 (at most)
#define KB_MAXBUCKETS 8192   // LDS histogram bins of passes 1 and 3 (32 KB)
#define KB_MAXSHIFT 12       // 2^B cells per bucket in LDS (pass 4, at most 32 KB)
#define KB_CAP 4096          // points of a bucket sorted in LDS (4 per thread)

struct KbPlan {
    int B;                   // bucket = key >> B
    int nkb;                 // buckets
    int nb;                  // point blocks
    int per;                 // points per block
};

static KbPlan kb_plan(size_t n, size_t k)
{
    KbPlan p;
    p.B = 10;
    while (((k + ((size_t)1 << p.B) - 1) >> p.B) > KB_MAXBUCKETS) ++p.B;
    p.nkb = (int)((k + ((size_t)1 << p.B) - 1) >> p.B);
    p.nb = (int)std::min<size_t>(KB_NB, (n + KB_THREADS - 1) / KB_THREADS);
    if (p.nb < 1) p.nb = 1;
    p.per = (int)((n + p.nb - 1) / p.nb);
    return p;
}

// bscan layout: [nkb * nb] row-exclusive block offsets, [nkb] bucket totals,
// [nkb + 1] bucket starts.
size_t gdt_bucket_words(size_t n, size_t k)
{
    const KbPlan p = kb_plan(n, k);
    return (size_t)p.nkb * p.nb + 2 * (size_t)p.nkb + 2;
}

bool gdt_cells_two_level(size_t n, size_t k)
{
    return kb_plan(n, k).B <= KB_MAXSHIFT;
}

// A lane's current set while it walks increasing caller indices (the sets are
// contiguous index ranges in order): one binary search for its first point,
// then a step forward whenever a point lies past the set's end.
struct KbSetCursor {
    const GdtSet* sets;
    int nsets, k, end;
    GdtSet S;

    __device__ void start(const GdtSet* __restrict__ all, int count, int i)
    {
        sets = all;
        nsets = count;
        k = gdt_set_of(all, count, i);
        S = all[k];
        end = S.p0 + S.n;
    }
    __device__ void advance(int i)
    {
        while (i >= end && k + 1 < nsets) {
            S = sets[++k];
            end = S.p0 + S.n;
        }
    }
};

// Cell of caller point i at p in set S.
__device__ inline void kb_bin_in(const GdtSet& S, const gdt_d2 p, uint32_t& key, uint32_t& packed,
                                 double& ox, double& oy)
{
    int cx, cy;
    gdt_bin_point(S, p.x, p.y, cx, cy, ox, oy);
    key = (uint32_t)(S.c0 + gdt_cell_key(S, cx, cy));
    packed = (uint32_t)cx | ((uint32_t)cy << 16);
}

// Points a lane reads before it bins any (their loads overlap).
#define KB_BATCH 4

// Pass 1: bmat[bucket * nb + block] = points of the block in the bucket.
__global__ __launch_bounds__(KB_THREADS) void k_kb_count(const GdtSet* __restrict__ sets, int nsets,
                                                         const gdt_d2* __restrict__ xy, int n,
                                                         KbPlan pl, uint32_t* __restrict__ bmat,
                                                         const int* __restrict__ halt,
                                                         int* __restrict__ st_zero,
                                                         unsigned long long* __restrict__ ct_zero)
{
    if (st_zero && blockIdx.x == 0) {
        if (threadIdx.x < ST_WORDS) st_zero[threadIdx.x] = 0;
        if (threadIdx.x < CT_WORDS) ct_zero[threadIdx.x] = 0ull;
    }
    if (halt && *halt) return;
    extern __shared__ uint32_t hist[];
    const int tid = threadIdx.x, b = blockIdx.x;
    for (int q = tid; q < pl.nkb; q += KB_THREADS) hist[q] = 0u;
    __syncthreads();
    const int i0 = b * pl.per, i1 = min(n, i0 + pl.per);
    if (i0 + tid < i1) {
        KbSetCursor cur;
        cur.start(sets, nsets, i0 + tid);
        for (int ib = i0 + tid; ib < i1; ib += KB_BATCH * KB_THREADS) {
            gdt_d2 p[KB_BATCH];
#pragma unroll
            for (int u = 0; u < KB_BATCH; ++u) p[u] = xy[min(ib + u * KB_THREADS, i1 - 1)];
#pragma unroll
            for (int u = 0; u < KB_BATCH; ++u) {
                const int i = ib + u * KB_THREADS;
                if (i < i1) {
                    cur.advance(i);
                    uint32_t key, packed;
                    double ox, oy;
                    kb_bin_in(cur.S, p[u], key, packed, ox, oy);
                    atomicAdd(&hist[key >> pl.B], 1u);
                }
            }
        }
    }
    __syncthreads();
    for (int q = tid; q < pl.nkb; q += KB_THREADS) bmat[(size_t)q * pl.nb + b] = hist[q];
}

// Pass 2: per bucket (one block each, KB_NB threads), the exclusive scan of
// its row of block counts, and the bucket total.
__global__ __launch_bounds__(KB_NB) void k_kb_rowscan(KbPlan pl, const uint32_t* __restrict__ bmat,
                                                      uint32_t* __restrict__ bscan)
{
    __shared__ uint32_t wsum[KB_NB / 64];
    const int tid = threadIdx.x, bin = blockIdx.x;
    const size_t row = (size_t)bin * pl.nb;
    const uint32_t v = tid < pl.nb ? bmat[row + tid] : 0u;
    const int lane = tid & 63, wid = tid >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    uint32_t woff = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < KB_NB / 64; ++w) {
        const uint32_t t = wsum[w];
        if (w < wid) woff += t;
        tot += t;
    }
    if (tid < pl.nb) bscan[row + tid] = woff + x - v;
    if (tid == 0) bscan[(size_t)pl.nkb * pl.nb + bin] = tot;
}

// Exclusive scan of a[0..m) in LDS by the block (m <= KB_THREADS * 8); a[m] = total.
__device__ inline void kb_scan_lds(uint32_t* a, int m, uint32_t* wsum)
{
    const int tid = threadIdx.x;
    const int ipt = (m + KB_THREADS - 1) / KB_THREADS;
    const int b = tid * ipt, e = min(b + ipt, m);
    uint32_t sum = 0;
    for (int q = b; q < e; ++q) sum += a[q];
    const int lane = tid & 63, wid = tid >> 6;
    uint32_t x = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    uint32_t woff = 0, tot = 0;
    for (int w = 0; w < KB_THREADS / 64; ++w) {
        const uint32_t t = wsum[w];
        if (w < wid) woff += t;
        tot += t;
    }
    uint32_t run = woff + x - sum;
    for (int q = b; q < e; ++q) {
        const uint32_t v = a[q];
        a[q] = run;
        run += v;
    }
    if (tid == 0) a[m] = tot;
    __syncthreads();
}

// Pass 3: bucket starts (scan of the totals), then each point appended to its
// (bucket, block) run.
__global__ __launch_bounds__(KB_THREADS) void k_kb_scatter(const GdtSet* __restrict__ sets, int nsets,
                                                           const gdt_d2* __restrict__ xy, int n,
                                                           KbPlan pl, uint32_t* __restrict__ bscan,
                                                           GdtKbRec* __restrict__ bkt,
                                                           const int* __restrict__ halt)
{
    if (halt && *halt) return;
    extern __shared__ uint32_t cursor[];                  // [nkb + 1]
    __shared__ uint32_t wsum[KB_THREADS / 64];
    const int tid = threadIdx.x, b = blockIdx.x;
    const size_t m = (size_t)pl.nkb * pl.nb;
    for (int q = tid; q < pl.nkb; q += KB_THREADS) cursor[q] = bscan[m + q];
    __syncthreads();
    kb_scan_lds(cursor, pl.nkb, wsum);
    if (b == 0)
        for (int q = tid; q <= pl.nkb; q += KB_THREADS) bscan[m + pl.nkb + q] = cursor[q];
    for (int q = tid; q < pl.nkb; q += KB_THREADS) cursor[q] += bscan[(size_t)q * pl.nb + b];
    __syncthreads();
    const int i0 = b * pl.per, i1 = min(n, i0 + pl.per);
    if (i0 + tid < i1) {
        KbSetCursor cur;
        cur.start(sets, nsets, i0 + tid);
        for (int ib = i0 + tid; ib < i1; ib += KB_BATCH * KB_THREADS) {
            gdt_d2 p[KB_BATCH];
#pragma unroll
            for (int u = 0; u < KB_BATCH; ++u) p[u] = xy[min(ib + u * KB_THREADS, i1 - 1)];
#pragma unroll
            for (int u = 0; u < KB_BATCH; ++u) {
                const int i = ib + u * KB_THREADS;
                if (i < i1) {
                    cur.advance(i);
                    GdtKbRec r;
                    kb_bin_in(cur.S, p[u], r.key, r.cell, r.off.x, r.off.y);
                    r.idx = i;
                    r.pad = 0u;
                    bkt[atomicAdd(&cursor[r.key >> pl.B], 1u)] = r;
                }
            }
        }
    }
}

// cstart/ccount of the bucket's keys from the LDS scan cst[0..nk].
__device__ inline void kb_write_starts(const uint32_t* cst, int nk, uint32_t k0, uint32_t base,
                                       uint32_t* __restrict__ cstart, uint32_t* __restrict__ ccount)
{
    for (int q = threadIdx.x; q < nk; q += KB_THREADS) {
        cstart[k0 + q] = base + cst[q];
        ccount[k0 + q] = cst[q + 1] - cst[q];
    }
}

// Slot arrays of slot `slot` from bucket record r.
__device__ inline void kb_write_slot(GdtCells& c, uint32_t slot, const GdtKbRec& r)
{
    gdt_f2 o32;
    o32.x = (float)r.off.x;
    o32.y = (float)r.off.y;
    c.perm[slot] = r.idx;
    if (c.slot_of) c.slot_of[r.idx] = (int)slot;
    c.off64[slot] = r.off;
    c.off32[slot] = o32;
    c.scell[slot] = r.cell;
}

// Pass 4: one block per bucket.  Up to KB_CAP points are ranked in LDS (cell
// by LDS counting sort, then caller index inside the cell) and the slot
// arrays written in slot order, one slot per thread; a larger bucket goes
// through global memory (perm holds bucket positions until the per-cell
// insertion sort is done).
__global__ __launch_bounds__(KB_THREADS) void k_kb_finalize(KbPlan pl, int total_keys, int total_n,
                                                            const uint32_t* __restrict__ bscan,
                                                            const GdtKbRec* __restrict__ bkt,
                                                            GdtCells c)
{
    if (c.halt && *c.halt) return;
    extern __shared__ uint32_t lds[];
    uint32_t* cst = lds;                                  // [2^B + 1] cell starts
    __shared__ uint32_t wsum[KB_THREADS / 64];
    const int tid = threadIdx.x, bin = blockIdx.x;
    const uint32_t* bstart = bscan + (size_t)pl.nkb * pl.nb + pl.nkb;
    const uint32_t k0 = (uint32_t)bin << pl.B;
    const int nk = min((int)(1 << pl.B), total_keys - (int)k0);
    const uint32_t base = bstart[bin];
    const int cnt = (int)(bstart[bin + 1] - base);
    if (bin == pl.nkb - 1 && tid == 0) c.cstart[total_keys] = (uint32_t)total_n;
    for (int q = tid; q <= nk; q += KB_THREADS) cst[q] = 0u;
    __syncthreads();
    if (cnt <= KB_CAP) {                                  // block-uniform
        int* idxs = reinterpret_cast<int*>(lds + (1 << pl.B) + 1);   // [KB_CAP] by position
        uint32_t* ord = lds + (1 << pl.B) + 1 + KB_CAP;   // [KB_CAP] bucket position by slot
        constexpr int PER = KB_CAP / KB_THREADS;
        uint32_t lc[PER], r0[PER];
        int idx[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int q = tid + u * KB_THREADS;
            lc[u] = 0u;
            r0[u] = 0u;
            idx[u] = 0;
            if (q < cnt) {
                lc[u] = bkt[base + q].key - k0;
                idx[u] = bkt[base + q].idx;
                r0[u] = atomicAdd(&cst[lc[u]], 1u);
            }
        }
        __syncthreads();
        kb_scan_lds(cst, nk, wsum);
        kb_write_starts(cst, nk, k0, base, c.cstart, c.ccount);
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int q = tid + u * KB_THREADS;
            if (q < cnt) idxs[cst[lc[u]] + r0[u]] = idx[u];
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < PER; ++u) {
            const int q = tid + u * KB_THREADS;
            if (q < cnt) {
                // rank by caller index inside the cell
                const uint32_t c0 = cst[lc[u]], c1 = cst[lc[u] + 1];
                uint32_t r = 0;
                for (uint32_t t = c0; t < c1; ++t) r += (uint32_t)(idxs[t] < idx[u]);
                ord[c0 + r] = (uint32_t)q;
            }
        }
        __syncthreads();
        for (int t = tid; t < cnt; t += KB_THREADS) kb_write_slot(c, base + t, bkt[base + ord[t]]);
        return;
    }
    uint32_t* cur = lds + (1 << pl.B) + 1;                // [2^B]
    for (int q = tid; q < cnt; q += KB_THREADS) atomicAdd(&cst[bkt[base + q].key - k0], 1u);
    __syncthreads();
    kb_scan_lds(cst, nk, wsum);
    kb_write_starts(cst, nk, k0, base, c.cstart, c.ccount);
    for (int q = tid; q < nk; q += KB_THREADS) cur[q] = cst[q];
    __syncthreads();
    for (int q = tid; q < cnt; q += KB_THREADS) {
        const uint32_t pos = atomicAdd(&cur[bkt[base + q].key - k0], 1u);
        c.perm[base + pos] = (int)(base + q);
    }
    __syncthreads();
    for (int q = tid; q < nk; q += KB_THREADS) {
        const uint32_t b = base + cst[q], e = base + cst[q + 1];
        for (uint32_t s = b + 1; s < e; ++s) {
            const int v = c.perm[s];
            const int vi = bkt[v].idx;
            uint32_t t = s;
            while (t > b && bkt[c.perm[t - 1]].idx > vi) {
                c.perm[t] = c.perm[t - 1];
                --t;
            }
            c.perm[t] = v;
        }
    }
    __syncthreads();
    for (int q = tid; q < cnt; q += KB_THREADS) {
        const uint32_t slot = base + (uint32_t)q;
        kb_write_slot(c, slot, bkt[c.perm[slot]]);
    }
}

// Small inputs (at most KB_SMALL_N points over at most KB_SMALL_NKB buckets of
// 2^10 keys): the four passes in one launch, one block per bucket.  Each block
// bins every point (16 per thread at most, read from L2), counts those of the
// buckets before its own (its first slot), counting-sorts its own by cell in
// LDS, orders each cell by caller index and writes cstart/ccount and the slot
// arrays: the slots of the four-pass build, without its three launch gaps.
// Measured per call (4,096 / 8,192 / 12,288 / 16,384 points): 0.071 / 0.089 /
// 0.114 / 0.122 ms against 0.083 / 0.099 / 0.122 / 0.127 ms with the four
// passes.
#define KB_SMALL_N 16384
#define KB_SMALL_NKB 64
#define KB_SMALL_B 10
#define KB_SMALL_PER (KB_SMALL_N / KB_THREADS)

__global__ __launch_bounds__(KB_THREADS) void k_kb_small(const GdtSet* __restrict__ sets, int nsets,
                                                         const gdt_d2* __restrict__ xy, int n,
                                                         int total_keys, GdtCells c,
                                                         int* __restrict__ st_zero,
                                                         unsigned long long* __restrict__ ct_zero)
{
    if (st_zero && blockIdx.x == 0) {
        if (threadIdx.x < ST_WORDS) st_zero[threadIdx.x] = 0;
        if (threadIdx.x < CT_WORDS) ct_zero[threadIdx.x] = 0ull;
    }
    if (c.halt && *c.halt) return;
    __shared__ uint32_t cst[(1 << KB_SMALL_B) + 1];      // cell counts, then starts
    __shared__ int idxs[KB_SMALL_N];                     // caller indices by position
    __shared__ uint16_t pcell[KB_SMALL_N];               // their cells in the bucket
    __shared__ uint32_t wsum[KB_THREADS / 64];
    __shared__ uint32_t bsum[KB_THREADS / 64];
    const int tid = threadIdx.x, bin = blockIdx.x;
    const uint32_t k0 = (uint32_t)bin << KB_SMALL_B;
    const int nk = min(1 << KB_SMALL_B, total_keys - (int)k0);
    if (bin == (int)gridDim.x - 1 && tid == 0) c.cstart[total_keys] = (uint32_t)n;
    for (int q = tid; q <= nk; q += KB_THREADS) cst[q] = 0u;
    __syncthreads();
    // pk[u]: (cell in the bucket << 16) | arrival rank in the cell, or ~0
    uint32_t pk[KB_SMALL_PER];
    uint32_t below = 0;
    KbSetCursor cur;
    if (tid < n) cur.start(sets, nsets, tid);
#pragma unroll
    for (int u = 0; u < KB_SMALL_PER; ++u) pk[u] = ~0u;
#pragma unroll
    for (int u0 = 0; u0 < KB_SMALL_PER; u0 += KB_BATCH) {
        if (tid + u0 * KB_THREADS >= n) break;            // no batch of clamped reads
        gdt_d2 p[KB_BATCH];
#pragma unroll
        for (int u = 0; u < KB_BATCH; ++u) p[u] = xy[min(tid + (u0 + u) * KB_THREADS, n - 1)];
#pragma unroll
        for (int u = 0; u < KB_BATCH; ++u) {
            const int i = tid + (u0 + u) * KB_THREADS;
            if (i < n) {
                cur.advance(i);
                uint32_t key, packed;
                double ox, oy;
                kb_bin_in(cur.S, p[u], key, packed, ox, oy);
                const uint32_t kb = key >> KB_SMALL_B;
                if (kb < (uint32_t)bin) ++below;
                else if (kb == (uint32_t)bin)
                    pk[u0 + u] = ((key - k0) << 16) | atomicAdd(&cst[key - k0], 1u);
            }
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) below += __shfl_xor(below, o, 64);
    if ((tid & 63) == 0) bsum[tid >> 6] = below;
    __syncthreads();
    uint32_t base = 0;
#pragma unroll
    for (int w = 0; w < KB_THREADS / 64; ++w) base += bsum[w];
    kb_scan_lds(cst, nk, wsum);
    kb_write_starts(cst, nk, k0, base, c.cstart, c.ccount);
#pragma unroll
    for (int u = 0; u < KB_SMALL_PER; ++u)
        if (pk[u] != ~0u) {
            const uint32_t pos = cst[pk[u] >> 16] + (pk[u] & 0xffffu);
            idxs[pos] = tid + u * KB_THREADS;
            pcell[pos] = (uint16_t)(pk[u] >> 16);
        }
    __syncthreads();
    // the bucket's points by position (about one per thread), each ranked by
    // caller index inside its cell and binned again for its slot record
    const int cnt = (int)cst[nk];
    for (int t = tid; t < cnt; t += KB_THREADS) {
        const int i = idxs[t];
        const uint32_t c0 = cst[pcell[t]], c1 = cst[pcell[t] + 1];
        uint32_t r = 0;
        for (uint32_t q = c0; q < c1; ++q) r += (uint32_t)(idxs[q] < i);
        GdtKbRec rec;
        kb_bin_in(sets[gdt_set_of(sets, nsets, i)], xy[i], rec.key, rec.cell, rec.off.x, rec.off.y);
        rec.idx = i;
        rec.pad = 0u;
        kb_write_slot(c, base + c0 + r, rec);
    }
}

static hipError_t build_cells_single_level(const GdtSet* d_sets, const GdtChunk* pchunks, int npch,
                                           const GdtChunk* kchunks, int nkch, int total_keys,
                                           const gdt_d2* xy, GdtCells& c, hipStream_t s)
{
    hipError_t e = gdt_fill_u32(c.ccount, 0u, (size_t)total_keys, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_bin, dim3(npch), dim3(GDT_BLOCK), 0, s, d_sets, pchunks, xy, c.pkey,
                       c.prank, c.ccount, c.halt);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    e = gdt_scan_u32(c.ccount, c.cstart, (size_t)total_keys, c.cstart + total_keys, c.scan_tmp, s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_scatter, dim3(npch), dim3(GDT_BLOCK), 0, s, d_sets, pchunks, c.pkey,
                       c.prank, c.cstart, c.perm, c.halt);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_finalize, dim3(nkch), dim3(GDT_BLOCK), 0, s, d_sets, kchunks, c.cstart,
                       xy, c.perm, c.slot_of, c.off64, c.off32, c.scell, c.halt);
    return hipGetLastError();
}

hipError_t gdt_build_cells(const GdtSet* d_sets, int nsets, const GdtChunk* pchunks, int npch,
                           const GdtChunk* kchunks, int nkch, int total_n, int total_keys,
                           const double* d_xy, GdtCells& c, hipStream_t s, int* status,
                           unsigned long long* counters)
{
    const gdt_d2* xy = reinterpret_cast<const gdt_d2*>(d_xy);
    const KbPlan pl = kb_plan((size_t)total_n, (size_t)total_keys);
    if (pl.B > KB_MAXSHIFT) {
        if (status) {
            const hipError_t e = gdt_status_reset(status, ST_WORDS, counters, CT_WORDS, s);
            if (e != hipSuccess) return e;
        }
        return build_cells_single_level(d_sets, pchunks, npch, kchunks, nkch, total_keys, xy, c, s);
    }
    if (pl.B == KB_SMALL_B && pl.nkb <= KB_SMALL_NKB && total_n >= 1 && total_n <= KB_SMALL_N) {
        hipLaunchKernelGGL(k_kb_small, dim3(pl.nkb), dim3(KB_THREADS), 0, s, d_sets, nsets, xy,
                           total_n, total_keys, c, status, counters);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_kb_count, dim3(pl.nb), dim3(KB_THREADS), pl.nkb * sizeof(uint32_t), s,
                       d_sets, nsets, xy, total_n, pl, c.bmat, c.halt, status, counters);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_kb_rowscan, dim3(pl.nkb), dim3(KB_NB), 0, s, pl, c.bmat, c.bscan);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    hipLaunchKernelGGL(k_kb_scatter, dim3(pl.nb), dim3(KB_THREADS), (pl.nkb + 1) * sizeof(uint32_t),
                       s, d_sets, nsets, xy, total_n, pl, c.bscan, c.bkt, c.halt);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const size_t cells = ((size_t)1 << pl.B) + 1;
    hipLaunchKernelGGL(k_kb_finalize, dim3(pl.nkb), dim3(KB_THREADS),
                       (cells + 2 * KB_CAP) * sizeof(uint32_t), s, pl, total_keys, total_n, c.bscan,
                       c.bkt, c);
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
