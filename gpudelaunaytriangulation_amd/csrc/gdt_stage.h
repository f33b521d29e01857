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

// Global loads a lane issues before it waits for any (staging is a few
// dependent rounds of loads per block: batching them keeps each round to one
// memory latency instead of one per cell or point).
#define GDT_STAGE_BATCH 4

// Stage the window of tile w into L (cell starts, frame coordinates, caller
// indices; FP32 coordinates too when F32; the cells' global starts when
// L.gstart is set).  Returns whether it fits (block-uniform); ends with a
// barrier.  Phase 1 reads the window cells' slot ranges; phase 2 copies the
// points entry-parallel (lane per LDS entry, so consecutive lanes read
// consecutive slots of a cell row), each entry first told its window cell
// through L.orig, which its caller index then overwrites.  Without L.gstart
// an entry's cell start is read again from the cell list (a cached read: the
// block read it in phase 1), which keeps that array out of LDS.
template <bool F32>
__device__ inline bool gdt_stage_window(const GdtSet& S, const GdtTileWin& w, const GdtRingIO& io,
                                        const GdtTileLds& L, uint32_t* wsum)
{
    // (indices past the end are clamped to the last valid one, so every load
    // of a batch is issued without a branch and the batch waits once)
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int c0 = 0; c0 < GDT_WIN_CELLS; c0 += GDT_STAGE_BATCH * nt) {
        uint32_t b[GDT_STAGE_BATCH], e[GDT_STAGE_BATCH];
#pragma unroll
        for (int j = 0; j < GDT_STAGE_BATCH; ++j) {
            const int key = gdt_tile_key(S, w, min(c0 + j * nt + tid, GDT_WIN_CELLS - 1));
            b[j] = io.cstart[key];
            e[j] = io.cstart[key + 1];
        }
#pragma unroll
        for (int j = 0; j < GDT_STAGE_BATCH; ++j) {
            const int c = c0 + j * nt + tid;
            if (c < GDT_WIN_CELLS) {
                if (L.gstart) L.gstart[c] = (int)b[j];
                L.cs[c] = e[j] - b[j];
            }
        }
    }
    __syncthreads();
    block_scan_lds(L.cs, GDT_WIN_CELLS, wsum);
    __syncthreads();
    const int np = (int)L.cs[GDT_WIN_CELLS];
    const bool fits = np <= GDT_WIN_MAXPTS && S.rmax >= GDT_SCAN;
    if (fits && np > 0) {
        for (int c = tid; c < GDT_WIN_CELLS; c += nt)
            for (uint32_t k = L.cs[c], ke = L.cs[c + 1]; k < ke; ++k) L.orig[k] = c;
        __syncthreads();
        for (int k0 = 0; k0 < np; k0 += GDT_STAGE_BATCH * nt) {
            int cell[GDT_STAGE_BATCH], slot[GDT_STAGE_BATCH], p[GDT_STAGE_BATCH];
            gdt_d2 o[GDT_STAGE_BATCH];
#pragma unroll
            for (int j = 0; j < GDT_STAGE_BATCH; ++j) {
                const int k = min(k0 + j * nt + tid, np - 1);
                cell[j] = L.orig[k];
                slot[j] = k - (int)L.cs[cell[j]];
                if (L.gstart) slot[j] += L.gstart[cell[j]];
            }
            if (!L.gstart) {
                uint32_t b[GDT_STAGE_BATCH];
#pragma unroll
                for (int j = 0; j < GDT_STAGE_BATCH; ++j) b[j] = io.cstart[gdt_tile_key(S, w, cell[j])];
#pragma unroll
                for (int j = 0; j < GDT_STAGE_BATCH; ++j) slot[j] += (int)b[j];
            }
#pragma unroll
            for (int j = 0; j < GDT_STAGE_BATCH; ++j) {
                o[j] = io.off64[slot[j]];
                p[j] = io.perm[slot[j]];
            }
            // (every lane read its entries' cells before any is overwritten:
            // a clamped lane reads entry np - 1, which another lane writes)
            __syncthreads();
#pragma unroll
            for (int j = 0; j < GDT_STAGE_BATCH; ++j) {
                const int k = k0 + j * nt + tid;
                if (k < np) {
                    const int fx = cell[j] % GDT_WIN - GDT_FRAME, fy = cell[j] / GDT_WIN - GDT_FRAME;
                    gdt_d2 X;
                    X.x = gdt_frame64(fx, S.bsx, o[j].x);
                    X.y = gdt_frame64(fy, S.bsy, o[j].y);
                    L.X64[k] = X;
                    L.orig[k] = p[j];
                    if (F32) {
                        gdt_f2 Y;
                        Y.x = (float)X.x;
                        Y.y = (float)X.y;
                        L.X32[k] = Y;
                    }
                }
            }
        }
    }
    __syncthreads();
    return fits;
}
