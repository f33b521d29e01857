// tile.hip -- the LDS-staged one-ring kernel (one block per tile of cells).
//
// Compiled once per supported tile edge: GDT_TILE_SHIFT (4: 16x16 cells, 5:
// 32x32) and the window's LDS point budget GDT_WIN_MAXPTS come from the
// compile line, and everything here lives in namespace ts<shift>, so both
// builds link into one library (gdt_kernels.h declares both launchers).
// Replaces gpu_voronoi_calc_global_kernel + gpu_get_neighbors_global_kernel
// (DelaunayGPU.cu:649-675, 981-1006) and their masked repair variants
// (DelaunayGPU.cu:611-640, 945-975).
#include "gdt_kernels.h"
#include "gdt_stage.h"
#include "gdt_wavedev.h"

#include <type_traits>

namespace GDT_TS_NS {

// One block per tile (GDT_TILE x GDT_TILE cells, this file's compiled tile edge): stage the tile window in LDS, then the lanes
// advance the tile's points one search step at a time.  With `list` and
// `flags` null the tile's points are all its slots [T0, T1); with `list` they
// are the slots list[lscan[T0] .. lscan[T1]) (a slot-ordered list of the
// points to redo, with lscan its exclusive scan over slots); with `flags` (a
// caller-order array, flagged >= 0) they are the tile's flagged slots, which
// the block gathers first into a bit mask in LDS (status[ST_NFLAGGED] += their
// count).  A tile with none exits before staging.  The points of a window that does not fit the LDS budget, or of a
// grid too small for the scan block, and points with a step whose circle
// leaves the window, are appended to `defer` and handled by k_one_ring_wave
// in the next launch.  With W = 2 or 4 (FP64) neighbouring lanes share each
// point, each scanning its part of every row (a list leaves most lanes of a
// block without a point otherwise, and a small grid leaves most of the chip
// idle; the launcher takes W from the caller).

// The tile's flagged slots (flags mode): slot T0 + 64 w + b is flagged when bit
// b of mask[w] is set; pre[w] = flagged slots before word w.
struct TileFlags {
    const unsigned long long* mask;
    const uint32_t* pre;
    int nw;     // words
};

// Slot of the q-th flagged slot (q below the tile's count).
__device__ inline int tile_flagged_slot(const TileFlags& f, int T0, int q)
{
    int lo = 0, hi = f.nw - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((int)f.pre[mid] <= q) lo = mid; else hi = mid - 1;
    }
    unsigned long long m = f.mask[lo];
    for (int r = q - (int)f.pre[lo]; r > 0; --r) m &= m - 1ull;
    return T0 + 64 * lo + __ffsll((long long)m) - 1;
}

template <int PREC, int W>
__global__ __launch_bounds__(1024) void k_one_ring_tile(const GdtSet* __restrict__ sets, const GdtChunk* __restrict__ tiles,
                     const int* __restrict__ list, const uint32_t* __restrict__ lscan,
                     const int* __restrict__ flags, const int* __restrict__ halt,
                     GdtRingIO io, int* __restrict__ counts, int* __restrict__ status,
                     unsigned long long* __restrict__ counters, int* __restrict__ defer)
{
    __shared__ uint32_t s_cs[GDT_WIN_CELLS + 1];
    __shared__ gdt_d2 s_X64[GDT_WIN_MAXPTS + 1];     // (+1: read by the scan's last prefetch)
    __shared__ gdt_f2 s_X32[PREC == 32 ? GDT_WIN_MAXPTS + 1 : 1];
    __shared__ int s_orig[GDT_WIN_MAXPTS];
    __shared__ uint32_t s_wsum[16];
    __shared__ int s_next;
    __shared__ int s_dlist[GDT_TILE_DEFER];
    __shared__ uint32_t s_dn;
    __shared__ unsigned long long s_fmask[GDT_WIN_MAXPTS / 64];
    __shared__ uint32_t s_fpre[GDT_WIN_MAXPTS / 64 + 1];
    extern __shared__ uint32_t s_tbl[];     // [5 * blockDim + 1] the lanes' row tables

    const GdtChunk ch = tiles[blockIdx.x];
    const GdtSet S = sets[ch.set];
    const GdtTileWin w = gdt_tile_window(S, ch.start);
    GdtTileLds L;
    L.cs = s_cs; L.gstart = nullptr; L.X64 = s_X64; L.X32 = s_X32; L.orig = s_orig;
    const int tid = threadIdx.x, nt = blockDim.x;
    const int k0 = S.c0 + ch.start * GDT_TILE_CELLS;
    const int T0 = (int)io.cstart[k0], T1 = (int)io.cstart[k0 + GDT_TILE_CELLS];
    // the tile's items q in [Q0, Q1): point j = item(q)
    int Q0 = list ? (int)lscan[T0] : T0, Q1 = list ? (int)lscan[T1] : T1;
    TileFlags tf;
    tf.mask = s_fmask;
    tf.pre = s_fpre;
    tf.nw = 0;
    if (flags) {
        if (halt && *halt) return;
        const int npt = T1 - T0;
        if (npt > GDT_WIN_MAXPTS) {
            // (more points than a window holds: the window cannot fit either)
            unsigned nf = 0;
            for (int q = T0 + tid; q < T1; q += nt)
                if (flags[io.perm[q]] >= 0) {
                    ++nf;
                    const uint32_t pos = atomicAdd(reinterpret_cast<unsigned int*>(&status[ST_NDEFER]), 1u);
                    defer[pos] = q;
                }
            if (nf) atomicAdd(&status[ST_NFLAGGED], (int)nf);
            return;
        }
        tf.nw = (npt + 63) >> 6;
        for (int t0 = 0; t0 < 64 * tf.nw; t0 += nt) {
            const int t = t0 + tid;
            const bool f = t < npt && flags[io.perm[T0 + t]] >= 0;
            const unsigned long long b = __ballot(f);
            if ((tid & 63) == 0 && (t >> 6) < tf.nw) s_fmask[t >> 6] = b;
        }
        __syncthreads();
        if (tid < 64) {
            const uint32_t c = tid < tf.nw ? (uint32_t)__popcll(s_fmask[tid]) : 0u;
            uint32_t x = c;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t y = __shfl_up(x, d, 64);
                if (tid >= d) x += y;
            }
            if (tid < tf.nw) s_fpre[tid + 1] = x;
            if (tid == 0) s_fpre[0] = 0u;
        }
        __syncthreads();
        Q0 = 0;
        Q1 = (int)s_fpre[tf.nw];
        if (tid == 0 && Q1 > 0) atomicAdd(&status[ST_NFLAGGED], Q1);
    }
    auto item = [&](int q) { return flags ? tile_flagged_slot(tf, T0, q) : (list ? list[q] : q); };
    if (Q0 >= Q1) return;    // block-uniform: nothing to do in this tile

    const bool fits = gdt_stage_window<PREC == 32>(S, w, io, L, s_wsum);

    if (!fits) {
        for (int q = Q0 + tid; q < Q1; q += nt) {
            const uint32_t pos = atomicAdd(reinterpret_cast<unsigned int*>(&status[ST_NDEFER]), 1u);
            defer[pos] = item(q);
        }
        return;
    }
    // Lanes take points of the tile and advance them one search step per
    // iteration; a lane whose point is finished takes the next one, so a
    // wave's iterations track the total number of steps, not the longest ring.
    static_assert(W == 1 || W == 2 || (W == 4 && PREC == 64), "1, 2 or (FP64) 4 lanes per point");
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
    int j = have ? item(q) : 0;
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
                    j = item(q);
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
            // (built without the tie rule and for rectangles only, which would
            // cost this kernel's main loop registers: a point with an exact tie
            // goes to k_one_ring_wave)
            const GdtRingResult r = gdt_wave_ring<false, false>(S, jd, io, wv, wv.lane == 0);
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
                                    const GdtChunk* tiles, int ntiles, int block, int lanes,
                                    int total_n, const int* list, const uint32_t* lscan,
                                    const int* flags, const int* halt,
                                    const GdtRingIO& io, int* counts, int* status,
                                    unsigned long long* counters, int* defer,
                                    hipEvent_t after_tile, hipStream_t s)
{
    if (ntiles <= 0) return hipSuccess;
    if (block > 1024) block = 1024;
    const size_t tbl = (5 * (size_t)block + 1) * sizeof(uint32_t);
    // (the FP32-filtered step shares a point between at most two lanes)
    const bool two = lanes == 2 || (lanes == 4 && precision == 32);
    if (precision == 64 && lanes == 4)
        hipLaunchKernelGGL((k_one_ring_tile<64, 4>), dim3(ntiles), dim3(block), tbl, s, d_sets,
                           tiles, list, lscan, flags, halt, io, counts, status, counters, defer);
    else if (precision == 32 && two)
        hipLaunchKernelGGL((k_one_ring_tile<32, 2>), dim3(ntiles), dim3(block), tbl, s, d_sets,
                           tiles, list, lscan, flags, halt, io, counts, status, counters, defer);
    else if (precision == 32)
        hipLaunchKernelGGL((k_one_ring_tile<32, 1>), dim3(ntiles), dim3(block), tbl, s, d_sets,
                           tiles, list, lscan, flags, halt, io, counts, status, counters, defer);
    else if (two)
        hipLaunchKernelGGL((k_one_ring_tile<64, 2>), dim3(ntiles), dim3(block), tbl, s, d_sets,
                           tiles, list, lscan, flags, halt, io, counts, status, counters, defer);
    else
        hipLaunchKernelGGL((k_one_ring_tile<64, 1>), dim3(ntiles), dim3(block), tbl, s, d_sets,
                           tiles, list, lscan, flags, halt, io, counts, status, counters, defer);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (after_tile && (e = hipEventRecord(after_tile, s)) != hipSuccess) return e;
    return gdt_launch_deferred(precision, d_sets, nsets, total_n, io, counts, status, counters,
                               defer, s);
}

}  // namespace GDT_TS_NS
