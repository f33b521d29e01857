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

// Block totals of a ring kernel (every thread calls it): one set of atomics
// per block, not per wave.
__device__ inline void ring_block_totals(int mx, int er, unsigned npred, unsigned nfb,
                                         unsigned nsum, int* status, unsigned long long* counters)
{
    const int tid = threadIdx.x, nt = blockDim.x;
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
    ring_block_totals(mx, er, npred, nfb, nsum, status, counters);
}

// Words of a star lane's row table: one per cell row of the 7x7 block (the
// uniform scan reads no word beyond its rows; a step's own tables, at most
// five rows and the word read after them, fit in the first six).
#define STAR_TBL (2 * GDT_SCAN + 1)

// One wavefront per point (small grids; FP64 predicates): lane g takes entry g
// of the cell block around the point's cell (7x7 cells, or 5x5 when the 7x7
// block holds more than 64 entries), in the order a step scans it, and runs
// the step's search (gdt_tile_search_f64) over that block with that
// candidate as the current neighbour -- every candidate's successor at once;
// the lane holding the point itself runs the nearest-neighbour search.  The
// ring is then the walk nearest -> successor -> ... read from the lanes: the
// sequential walk, each step's result taken from the lane that ran the same
// search on the same state.  Where the walk reaches a candidate whose search
// the lanes did not certify (its circle leaves the block: those lanes do not
// extend) or one outside the block, lane 0 runs that one search in full; a
// block of more than 64 entries runs the sequential steps in lane 0.  Blocks
// (tile, y) share a tile's points: wave w of block y takes the points
// T0 + y nw + w + m gridDim.y nw.  Deferred points go to `defer` (handled by
// k_one_ring_wave in the next launch).
__global__ __launch_bounds__(1024) void k_one_ring_star(const GdtSet* __restrict__ sets,
                                                        const GdtChunk* __restrict__ tiles,
                                                        GdtRingIO io, int* __restrict__ counts,
                                                        int* __restrict__ status,
                                                        unsigned long long* __restrict__ counters,
                                                        int* __restrict__ defer)
{
    __shared__ uint32_t s_cs[GDT_WIN_CELLS + 1];
    __shared__ gdt_d2 s_X64[GDT_WIN_MAXPTS + 1];
    __shared__ gdt_f2 s_X32[1];
    __shared__ int s_orig[GDT_WIN_MAXPTS];
    __shared__ uint32_t s_wsum[16];
    extern __shared__ uint32_t s_tbl[];     // [STAR_TBL * blockDim + 1] the lanes' row tables

    const GdtChunk ch = tiles[blockIdx.x];
    const GdtSet S = sets[ch.set];
    const GdtTileWin w = gdt_tile_window(S, ch.start);
    GdtTileLds L;
    L.cs = s_cs; L.gstart = nullptr; L.X64 = s_X64; L.X32 = s_X32; L.orig = s_orig;
    const int tid = threadIdx.x, lane = tid & 63, nw = (int)blockDim.x >> 6;
    const int k0 = S.c0 + ch.start * GDT_TILE_CELLS;
    const int T0 = (int)io.cstart[k0], T1 = (int)io.cstart[k0 + GDT_TILE_CELLS];
    const int qstep = (int)gridDim.y * nw;
    const int qb = T0 + (int)blockIdx.y * nw;
    if (qb >= T1) return;    // block-uniform: no point for this block

    const bool fits = gdt_stage_window<false>(S, w, io, L, s_wsum);
    if (!fits) {
        if (lane == 0)
            for (int q = qb + (tid >> 6); q < T1; q += qstep) {
                const uint32_t pos = atomicAdd(reinterpret_cast<unsigned int*>(&status[ST_NDEFER]), 1u);
                defer[pos] = q;
            }
        return;
    }
    unsigned npred = 0, nsum = 0;
    int mx = 0, er = 0;
    for (int q = qb + (tid >> 6); q < T1; q += qstep) {     // wave-uniform
        GdtTileLane<double> st;
        st.tbl = s_tbl + STAR_TBL * tid;
        gdt_tile_lane_init<double>(w, L, q, io, S, st);
        // the lanes' block, as one row table (every lane the same): the
        // (2 GDT_SCAN + 1)^2 cells when they hold at most 64 entries, else
        // the (2 GDT_SCAN0 + 1)^2 cells; and this lane's entry of it
        const GdtClip whole = {0.f, 0.f, 0};
        int ph = GDT_SCAN;
        int tot = (int)gdt_block_rows(L, st.c0, ph, whole, st.tbl);
        if (tot > 64) {
            ph = GDT_SCAN0;
            tot = (int)gdt_block_rows(L, st.c0, ph, whole, st.tbl);
        }
        int kg = -1;
        for (int acc = 0, r = 0; acc < tot; ++r) {      // (the table's non-empty rows)
            const uint32_t t = st.tbl[r];
            const int b = (int)(t & 0xffffu), n = (int)(t >> 16) - b;
            if (lane >= acc && lane < acc + n) kg = b + lane - acc;
            acc += n;
        }
        int bk = -1, rc = STEP_DEFER;
        if (tot <= 64 && kg >= 0) {
            GdtTileLane<double> sl = st;
            if (kg != st.iv) {
                const gdt_d2 X = L.X64[kg];
                sl.ak = kg;
                sl.ax = X.x - st.vx;
                sl.ay = X.y - st.vy;
            }
            double bx, by;
            rc = gdt_tile_search_f64<1>(S, L, sl, npred, 0, false, bk, bx, by, st.tbl,
                                        (unsigned)tot, ph);
        }
        int res = STEP_RUN, count = 0;
        if (tot > 64) {
            if (lane == 0) {
                int r = STEP_RUN;
                while (r == STEP_RUN) r = gdt_tile_step_f64<1>(S, L, io, st, npred);
                res = r;
                count = st.count;
            }
        } else {
            const unsigned long long mv = __ballot(kg == st.iv);
            const int hv = __ffsll((long long)mv) - 1;
            int nn = __shfl(bk, hv, 64), nrc = __shfl(rc, hv, 64);
            if (nrc != STEP_RUN) {
                // not certified over the block: lane 0 runs the full search
                if (lane == 0) {
                    double bx, by;
                    nrc = gdt_tile_search_f64<1>(S, L, st, npred, 0, true, nn, bx, by);
                }
                nrc = __shfl(nrc, 0, 64);
                nn = __shfl(nn, 0, 64);
            }
            if (nrc != STEP_RUN) {
                res = STEP_DEFER;
            } else {
                if (lane == 0) st.row[0] = L.orig[nn];
                count = 1;
                for (int cur = nn;;) {
                    const unsigned long long m = __ballot(kg == cur && rc == STEP_RUN);
                    int nx;
                    if (m) {
                        nx = __shfl(bk, __ffsll((long long)m) - 1, 64);
                    } else {
                        // no lane certified cur's successor: lane 0 runs the
                        // full search from cur (extension included)
                        int r = STEP_DEFER, b = -1;
                        if (lane == 0) {
                            GdtTileLane<double> sl = st;
                            const gdt_d2 X = L.X64[cur];
                            sl.ak = cur;
                            sl.ax = X.x - st.vx;
                            sl.ay = X.y - st.vy;
                            double bx, by;
                            r = gdt_tile_search_f64<1>(S, L, sl, npred, 0, true, b, bx, by);
                        }
                        if (__shfl(r, 0, 64) != STEP_RUN) {
                            res = STEP_DEFER;
                            break;
                        }
                        nx = __shfl(b, 0, 64);
                    }
                    if (nx == nn) {
                        res = STEP_DONE;
                        break;
                    }
                    if (lane == 0 && count < io.stride) st.row[count] = L.orig[nx];
                    if (++count > GDT_HARD_CAP) {
                        res = STEP_NOCLOSE;
                        break;
                    }
                    cur = nx;
                }
            }
        }
        if (lane == 0) {
            if (res == STEP_DEFER) {
                const uint32_t pos = atomicAdd(reinterpret_cast<unsigned int*>(&status[ST_NDEFER]), 1u);
                defer[pos] = q;
            } else {
                counts[L.orig[st.iv]] = count;
                mx = max(mx, count);
                nsum += (unsigned)count;
                if (res == STEP_NOCLOSE) er |= ERR_NOCLOSE;
            }
        }
    }
    ring_block_totals(mx, er, npred, 0u, nsum, status, counters);
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

hipError_t gdt_launch_one_ring_star(const GdtSet* d_sets, int nsets, const GdtChunk* tiles,
                                    int ntiles, int nsplit, int block, int total_n,
                                    const GdtRingIO& io, int* counts, int* status,
                                    unsigned long long* counters, int* defer,
                                    hipEvent_t after_tile, hipStream_t s)
{
    if (ntiles <= 0) return hipSuccess;
    if (block < 64 || block > 1024 || (block & 63) || nsplit < 1 || nsplit > 65535)
        return hipErrorInvalidValue;
    const size_t tbl = (STAR_TBL * (size_t)block + 1) * sizeof(uint32_t);
    hipLaunchKernelGGL(k_one_ring_star, dim3(ntiles, nsplit), dim3(block), tbl, s, d_sets, tiles,
                       io, counts, status, counters, defer);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (after_tile && (e = hipEventRecord(after_tile, s)) != hipSuccess) return e;
    return gdt_launch_deferred(64, d_sets, nsets, total_n, io, counts, status, counters, defer, s);
}

}  // namespace GDT_TS_NS
