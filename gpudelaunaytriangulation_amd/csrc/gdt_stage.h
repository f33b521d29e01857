// gdt_stage.h -- device-only helpers shared by the LDS tile kernels: block
// scan in LDS and the staging of a tile's window (gdt_tile.h layout).
#pragma once

#include "gdt_tile.h"

// Exclusive scan of a[0..n) in LDS by the whole block; a[n] = total.
// wsum holds one word per wave.
__device__ inline void block_scan_lds(uint32_t* a, int n, uint32_t* wsum)
{
    const int nt = blockDim.x, tid = threadIdx.x;
    const int ipt = (n + nt - 1) / nt;
    const int b = tid * ipt, e = min(b + ipt, n);
    uint32_t s = 0;
    for (int i = b; i < e; ++i) s += a[i];
    const int lane = tid & 63, wid = tid >> 6;
    uint32_t x = s;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    uint32_t woff = 0, tot = 0;
    for (int q = 0; q < (nt >> 6); ++q) {
        const uint32_t t = wsum[q];
        if (q < wid) woff += t;
        tot += t;
    }
    uint32_t run = woff + x - s;
    for (int i = b; i < e; ++i) {
        const uint32_t v = a[i];
        a[i] = run;
        run += v;
    }
    if (tid == 0) a[n] = tot;
}

// Stage the window of tile w into L (cell starts, global starts, frame
// coordinates, caller indices; FP32 coordinates too when F32).  Returns
// whether it fits (block-uniform); ends with a barrier.
template <bool F32>
__device__ inline bool gdt_stage_window(const GdtSet& S, const GdtTileWin& w, const GdtRingIO& io,
                                        const GdtTileLds& L, uint32_t* wsum)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int c = tid; c < GDT_WIN_CELLS; c += nt) {
        int gs;
        uint32_t cn;
        gdt_tile_count(S, w, c, io.cstart, gs, cn);
        L.gstart[c] = gs;
        L.cs[c] = cn;
    }
    __syncthreads();
    block_scan_lds(L.cs, GDT_WIN_CELLS, wsum);
    __syncthreads();
    const bool fits = L.cs[GDT_WIN_CELLS] <= GDT_WIN_MAXPTS && S.rmax >= GDT_SCAN;
    if (fits)
        for (int c = tid; c < GDT_WIN_CELLS; c += nt)
            gdt_tile_copy<F32>(S, c, L.gstart[c], L.cs[c], L.cs[c + 1] - L.cs[c], io, L);
    __syncthreads();
    return fits;
}
