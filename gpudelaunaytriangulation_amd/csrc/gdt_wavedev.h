// gdt_wavedev.h -- device-only wavefront helpers shared by the one-ring
// kernels: wave-wide reductions and the wavefront policy of gdt_wave_ring.
#pragma once

#include "gdt_wave.h"

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

