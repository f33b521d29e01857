// onering.hip -- per-point Delaunay one-ring kernels (FP64 and FP32-filtered).
//
// One lane per point.  Lanes of a block take consecutive cell-ordered slots,
// so a wavefront covers a compact patch of cells (8x8-cell tiles by default)
// and the candidate offsets it reads come from the same few cache lines.
// Replaces gpu_voronoi_calc_global_kernel + gpu_get_neighbors_global_kernel
// (DelaunayGPU.cu:649-675, 981-1006) and their masked repair variants
// (DelaunayGPU.cu:611-640, 945-975), which instead launch every point and
// mask most of them out.
#include "gdt_kernels.h"
#include "gdt_stage.h"
#include "gdt_wave.h"

#include <type_traits>

__device__ inline unsigned wave_sum_u32(unsigned v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ inline int wave_max_i32(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline int wave_or_i32(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
    return v;
}

template <int PREC>
__device__ inline void one_ring_lane(bool active, const GdtSet& S, int j, const GdtRingIO& io,
                                     int* __restrict__ counts, int* __restrict__ status,
                                     unsigned long long* __restrict__ counters)
{
    GdtRingResult r = {0, 0, 0u, 0u};
    if (active) {
        r = (PREC == 32) ? gdt_one_ring_f32(S, j, io) : gdt_one_ring_f64(S, j, io);
        counts[io.perm[j]] = r.count;
    }
    // every lane of the wave reaches the reductions below
    const int mx = wave_max_i32(r.count);
    const int er = wave_or_i32(r.err);
    const unsigned np = wave_sum_u32(r.npred);
    const unsigned nf = wave_sum_u32(r.nfb);
    const unsigned ns = wave_sum_u32(active ? (unsigned)r.count : 0u);
    if ((threadIdx.x & 63) == 0) {
        if (mx > 0) atomicMax(&status[ST_REQUIRED], mx);
        if (er) atomicOr(&status[ST_ERRFLAGS], er);
        if (np) atomicAdd(&counters[CT_PRED], (unsigned long long)np);
        if (nf) atomicAdd(&counters[CT_FALLBACK], (unsigned long long)nf);
        if (ns) atomicAdd(&counters[CT_NBRSUM], (unsigned long long)ns);
    }
}

// All points: block b handles chunk b (one set, blockDim consecutive slots).
template <int PREC>
__global__ __launch_bounds__(1024) void k_one_ring(const GdtSet* __restrict__ sets,
                                                   const GdtChunk* __restrict__ chunks,
                                                   GdtRingIO io, int* __restrict__ counts,
                                                   int* __restrict__ status,
                                                   unsigned long long* __restrict__ counters)
{
    const GdtChunk ch = chunks[blockIdx.x];
    const GdtSet S = sets[ch.set];
    const int li = ch.start + (int)threadIdx.x;
    const bool active = li < S.n;
    one_ring_lane<PREC>(active, S, S.p0 + li, io, counts, status, counters);
}

// Worklist of slots (count in device memory): grid-stride, set found per lane.
// Entries [skip, *worklen) of the list are processed.
template <int PREC>
__global__ __launch_bounds__(1024) void k_one_ring_list(const GdtSet* __restrict__ sets, int nsets,
                                                        const int* __restrict__ worklist,
                                                        const uint32_t* __restrict__ worklen,
                                                        int skip, GdtRingIO io,
                                                        int* __restrict__ counts,
                                                        int* __restrict__ status,
                                                        unsigned long long* __restrict__ counters)
{
    const int len = max((int)*worklen - skip, 0);
    worklist += skip;
    const int stride = gridDim.x * blockDim.x;
    const int first = blockIdx.x * blockDim.x + threadIdx.x;
    // uniform trip count per wave so the reductions see every lane
    const int wave_first = first - (int)(threadIdx.x & 63);
    for (int base = wave_first; base < len; base += stride) {
        const int t = base + (int)(threadIdx.x & 63);
        const bool active = t < len;
        int j = 0, s = 0;
        if (active) {
            j = worklist[t];
            s = gdt_set_of(sets, nsets, j);
        }
        const GdtSet S = sets[s];
        one_ring_lane<PREC>(active, S, j, io, counts, status, counters);
    }
}

// Wavefront policy of gdt_wave_ring: every lane scans its share of a ring,
// then a shuffle-down tree leaves the best at lane 0, which is broadcast.
struct GdtWaveDev {
    int lane;
    template <bool TIES, class Fn>
    __device__ GdtWaveBest ring(const GdtWaveBest& start, bool wrap, double ax, double ay,
                                GdtTieIds& t, Fn& fn)
    {
        GdtWaveBest m = start;
        fn(lane, m);
#pragma unroll
        for (int off = GDT_WAVE / 2; off > 0; off >>= 1) {
            GdtWaveBest p;
            p.k = __shfl_down(m.k, off, 64);
            p.x = __shfl_down(m.x, off, 64);
            p.y = __shfl_down(m.y, off, 64);
            p.d2 = __shfl_down(m.d2, off, 64);
            if (lane < off && gdt_wave_better<TIES>(wrap, ax, ay, t, m, p)) m = p;
        }
        GdtWaveBest g;
        g.k = __shfl(m.k, 0, 64);
        g.x = __shfl(m.x, 0, 64);
        g.y = __shfl(m.y, 0, 64);
        g.d2 = __shfl(m.d2, 0, 64);
        if constexpr (!TIES) t.tie = __ballot(t.tie) != 0;
        return g;
    }
};

// Deferred points, one wavefront per point (FP64 in both precision paths):
// list entries [0, min(*worklen, cap)), grid-stride over the waves.
__global__ __launch_bounds__(256) void k_one_ring_wave(const GdtSet* __restrict__ sets, int nsets,
                                                       const int* __restrict__ worklist,
                                                       const uint32_t* __restrict__ worklen,
                                                       int cap, GdtRingIO io,
                                                       int* __restrict__ counts,
                                                       int* __restrict__ status,
                                                       unsigned long long* __restrict__ counters)
{
    const int lane = (int)(threadIdx.x & 63);
    const int wid = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int nw = (int)((gridDim.x * blockDim.x) >> 6);
    const int len = min((int)*worklen, cap);
    GdtWaveDev w;
    w.lane = lane;
    unsigned npred = 0, nsum = 0;
    int mx = 0, er = 0;
    for (int t = wid; t < len; t += nw) {
        const int j = worklist[t];
        const GdtSet S = sets[gdt_set_of(sets, nsets, j)];
        const GdtRingResult r = gdt_wave_ring(S, j, io, w, lane == 0);
        npred += r.npred;
        if (lane == 0) {
            counts[io.perm[j]] = r.count;
            mx = max(mx, r.count);
            er |= r.err;
            nsum += (unsigned)r.count;
        }
    }
    mx = wave_max_i32(mx);
    er = wave_or_i32(er);
    npred = wave_sum_u32(npred);
    nsum = wave_sum_u32(nsum);
    if (lane == 0) {
        if (mx > 0) atomicMax(&status[ST_REQUIRED], mx);
        if (er) atomicOr(&status[ST_ERRFLAGS], er);
        if (npred) atomicAdd(&counters[CT_PRED], (unsigned long long)npred);
        if (nsum) atomicAdd(&counters[CT_NBRSUM], (unsigned long long)nsum);
    }
}

// One block per tile (GDT_TILE x GDT_TILE cells): stage the tile window in LDS, then the lanes
// advance the tile's points one search step at a time.  With `list` null the
// tile's points are all its slots [T0, T1); otherwise they are the slots
// list[lscan[T0] .. lscan[T1]) (a slot-ordered list of the points to redo, with
// lscan its exclusive scan over slots), and a tile with none exits before
// staging.  The points of a window that does not fit the LDS budget, or of a
// grid too small for the scan block, and points with a step whose circle
// leaves the window, are appended to `defer` and handled by k_one_ring_wave
// in the next launch.  With W = 2 (lists of points) two neighbouring lanes
// share each point, each scanning half of every row (a list leaves most lanes
// of a block without a point otherwise).
#ifndef GDT_LIST_LANES
#define GDT_LIST_LANES 2    // lanes per point of the tile kernel over a list of points
#endif

template <int PREC, int W>
__global__ __launch_bounds__(1024) void k_one_ring_tile(const GdtSet* __restrict__ sets, const GdtChunk* __restrict__ tiles,
                     const int* __restrict__ list, const uint32_t* __restrict__ lscan,
                     GdtRingIO io, int* __restrict__ counts, int* __restrict__ status,
                     unsigned long long* __restrict__ counters, int* __restrict__ defer)
{
    __shared__ uint32_t s_cs[GDT_WIN_CELLS + 1];
    __shared__ int s_gstart[GDT_WIN_CELLS];
    __shared__ gdt_d2 s_X64[GDT_WIN_MAXPTS + 1];     // (+1: read by the scan's last prefetch)
    __shared__ gdt_f2 s_X32[PREC == 32 ? GDT_WIN_MAXPTS + 1 : 1];
    __shared__ int s_orig[GDT_WIN_MAXPTS];
    __shared__ uint32_t s_wsum[16];
    __shared__ int s_next;
    __shared__ int s_dlist[GDT_TILE_DEFER];
    __shared__ uint32_t s_dn;
    extern __shared__ uint32_t s_tbl[];     // [5 * blockDim + 1] the lanes' row tables

    const GdtChunk ch = tiles[blockIdx.x];
    const GdtSet S = sets[ch.set];
    const GdtTileWin w = gdt_tile_window(S, ch.start);
    GdtTileLds L;
    L.cs = s_cs; L.gstart = s_gstart; L.X64 = s_X64; L.X32 = s_X32; L.orig = s_orig;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int k0 = S.c0 + ch.start * GDT_TILE_CELLS;
    const int T0 = (int)io.cstart[k0], T1 = (int)io.cstart[k0 + GDT_TILE_CELLS];
    // the tile's items q in [Q0, Q1): point j = list ? list[q] : q
    const int Q0 = list ? (int)lscan[T0] : T0, Q1 = list ? (int)lscan[T1] : T1;
    if (Q0 >= Q1) return;    // block-uniform: nothing to do in this tile

    const bool fits = gdt_stage_window<PREC == 32>(S, w, io, L, s_wsum);

    if (!fits) {
        for (int q = Q0 + tid; q < Q1; q += nt) {
            const uint32_t pos = atomicAdd(reinterpret_cast<unsigned int*>(&status[ST_NDEFER]), 1u);
            defer[pos] = list ? list[q] : q;
        }
        return;
    }
    // Lanes take points of the tile and advance them one search step per
    // iteration; a lane whose point is finished takes the next one, so a
    // wave's iterations track the total number of steps, not the longest ring.
    static_assert(W == 1 || W == 2, "one or two lanes per point");
    if (tid == 0) {
        s_next = Q0 + nt / W;
        s_dn = 0u;
    }
    __syncthreads();
    using Real = typename std::conditional<PREC == 32, float, double>::type;
    GdtTileLane<Real> st;
    st.tbl = s_tbl + 5 * tid;
    const int sub = tid & (W - 1);      // the lane's part of its point's rows
    int q = Q0 + tid / W;
    bool have = q < Q1;
    int j = have ? (list ? list[q] : q) : 0;
    if (have) gdt_tile_lane_init<Real>(w, L, j, io, S, st);
    unsigned npred = 0, nfb = 0, nsum = 0;
    int mx = 0, er = 0;
    while (__ballot(have) != 0) {
        if (have) {
            int rc;
            if constexpr (PREC == 32)
                rc = gdt_tile_step_f32<W>(S, L, io, st, npred, nfb, sub);
            else
                rc = gdt_tile_step_f64<W>(S, L, io, st, npred, sub);
            // (both lanes of a pair end each step in the same state; the first reports)
            if (rc != STEP_RUN && sub == 0) {
                if (rc == STEP_DEFER) {
                    // the block's own waves finish it after the tile (below)
                    const uint32_t dpos = atomicAdd(&s_dn, 1u);
                    if (dpos < GDT_TILE_DEFER) {
                        s_dlist[dpos] = j;
                    } else {
                        const uint32_t pos = atomicAdd(reinterpret_cast<unsigned int*>(&status[ST_NDEFER]), 1u);
                        defer[pos] = j;
                    }
                } else {
                    counts[L.orig[st.iv]] = st.count;
                    mx = max(mx, st.count);
                    nsum += (unsigned)st.count;
                    if (rc == STEP_NOCLOSE) er |= ERR_NOCLOSE;
                }
            }
            if (rc != STEP_RUN) {
                if constexpr (W > 1) {
                    int qq = 0;
                    if (sub == 0) qq = atomicAdd(&s_next, 1);
                    q = __shfl(qq, (tid & 63) & ~(W - 1), 64);
                } else {
                    q = atomicAdd(&s_next, 1);
                }
                have = q < Q1;
                if (have) {
                    j = list ? list[q] : q;
                    gdt_tile_lane_init<Real>(w, L, j, io, S, st);
                }
            }
        }
    }
    // Points whose circle left the window: one wavefront each, read from
    // global memory (gdt_wave_ring), by this block's waves while other blocks
    // still run their tiles.
    __syncthreads();
    const int ndl = min((int)s_dn, GDT_TILE_DEFER);
    if (ndl > 0) {
        if (tid == 0) atomicAdd(&status[ST_NDEFER_TILE], ndl);
        GdtWaveDev wv;
        wv.lane = tid & 63;
        for (int t = tid >> 6; t < ndl; t += nt >> 6) {
            const int jd = s_dlist[t];
            // (built without the tie rule, which would cost this kernel's main
            // loop registers: a point with an exact tie goes to k_one_ring_wave)
            const GdtRingResult r = gdt_wave_ring<false>(S, jd, io, wv, wv.lane == 0);
            npred += r.npred;
            if (r.err & ERR_TIE) {
                if (wv.lane == 0) {
                    const uint32_t pos = atomicAdd(reinterpret_cast<unsigned int*>(&status[ST_NDEFER]), 1u);
                    defer[pos] = jd;
                }
            } else if (wv.lane == 0) {
                counts[io.perm[jd]] = r.count;
                mx = max(mx, r.count);
                er |= r.err;
                nsum += (unsigned)r.count;
            }
        }
    }
    // block totals: one set of atomics per block, not per wave
    mx = wave_max_i32(mx);
    er = wave_or_i32(er);
    npred = wave_sum_u32(npred);
    nfb = wave_sum_u32(nfb);
    nsum = wave_sum_u32(nsum);
    __shared__ unsigned s_red[5][16];
    if ((tid & 63) == 0) {
        const int w = tid >> 6;
        s_red[0][w] = (unsigned)mx;
        s_red[1][w] = (unsigned)er;
        s_red[2][w] = npred;
        s_red[3][w] = nfb;
        s_red[4][w] = nsum;
    }
    __syncthreads();
    if (tid == 0) {
        unsigned long long bp = 0, bf = 0, bs = 0;
        int bm = 0, be = 0;
        for (int w = 0; w < (nt >> 6); ++w) {
            bm = max(bm, (int)s_red[0][w]);
            be |= (int)s_red[1][w];
            bp += s_red[2][w];
            bf += s_red[3][w];
            bs += s_red[4][w];
        }
        if (bm > 0) atomicMax(&status[ST_REQUIRED], bm);
        if (be) atomicOr(&status[ST_ERRFLAGS], be);
        if (bp) atomicAdd(&counters[CT_PRED], bp);
        if (bf) atomicAdd(&counters[CT_FALLBACK], bf);
        if (bs) atomicAdd(&counters[CT_NBRSUM], bs);
    }
}

hipError_t gdt_launch_one_ring_tile(int precision, const GdtSet* d_sets, int nsets,
                                    const GdtChunk* tiles, int ntiles, int block, int total_n,
                                    const int* list, const uint32_t* lscan, const GdtRingIO& io,
                                    int* counts, int* status, unsigned long long* counters,
                                    int* defer, hipEvent_t after_tile, hipStream_t s)
{
    if (ntiles <= 0) return hipSuccess;
    if (block > 1024) block = 1024;
    const size_t tbl = (5 * (size_t)block + 1) * sizeof(uint32_t);
    if (precision == 32 && list)
        hipLaunchKernelGGL((k_one_ring_tile<32, GDT_LIST_LANES>), dim3(ntiles), dim3(block), tbl,
                           s, d_sets, tiles, list, lscan, io, counts, status, counters, defer);
    else if (precision == 32)
        hipLaunchKernelGGL((k_one_ring_tile<32, 1>), dim3(ntiles), dim3(block), tbl, s, d_sets,
                           tiles, list, lscan, io, counts, status, counters, defer);
    else if (list)
        hipLaunchKernelGGL((k_one_ring_tile<64, GDT_LIST_LANES>), dim3(ntiles), dim3(block), tbl, s, d_sets,
                           tiles, list, lscan, io, counts, status, counters, defer);
    else
        hipLaunchKernelGGL((k_one_ring_tile<64, 1>), dim3(ntiles), dim3(block), tbl, s, d_sets,
                           tiles, list, lscan, io, counts, status, counters, defer);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (after_tile && (e = hipEventRecord(after_tile, s)) != hipSuccess) return e;
    // Deferred points (count read on the device): the first GDT_WAVE_CAP one
    // wavefront each, any beyond (dense clusters) one lane each.
    const uint32_t* ndefer = reinterpret_cast<const uint32_t*>(status + ST_NDEFER);
    hipLaunchKernelGGL(k_one_ring_wave, dim3(GDT_WAVE_BLOCKS), dim3(256), 0, s, d_sets, nsets, defer,
                       ndefer, GDT_WAVE_CAP, io, counts, status, counters);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return gdt_launch_one_ring(precision, d_sets, nsets, nullptr, 0, 256, defer, ndefer,
                               GDT_WAVE_CAP, total_n, io, counts, status, counters, s);
}

hipError_t gdt_launch_one_ring(int precision, const GdtSet* d_sets, int nsets,
                               const GdtChunk* chunks, int nch, int block, const int* worklist,
                               const uint32_t* worklen, int skip, int max_work,
                               const GdtRingIO& io, int* counts, int* status,
                               unsigned long long* counters, hipStream_t s)
{
    if (!worklist) {
        if (nch <= 0) return hipSuccess;
        if (precision == 32)
            hipLaunchKernelGGL(k_one_ring<32>, dim3(nch), dim3(block), 0, s, d_sets, chunks, io,
                               counts, status, counters);
        else
            hipLaunchKernelGGL(k_one_ring<64>, dim3(nch), dim3(block), 0, s, d_sets, chunks, io,
                               counts, status, counters);
        return hipGetLastError();
    }
    max_work -= skip;
    if (max_work <= 0) return hipSuccess;
    int grid = (max_work + block - 1) / block;
    if (grid > 4096) grid = 4096;
    if (precision == 32)
        hipLaunchKernelGGL(k_one_ring_list<32>, dim3(grid), dim3(block), 0, s, d_sets, nsets,
                           worklist, worklen, skip, io, counts, status, counters);
    else
        hipLaunchKernelGGL(k_one_ring_list<64>, dim3(grid), dim3(block), 0, s, d_sets, nsets,
                           worklist, worklen, skip, io, counts, status, counters);
    return hipGetLastError();
}
