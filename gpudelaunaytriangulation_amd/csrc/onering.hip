// onering.hip -- per-point Delaunay one-ring kernels (FP64 and FP32-filtered)
// and the wavefront kernel of the points the LDS tile kernel (tile.hip) defers.
//
// One lane per point.  Lanes of a block take consecutive cell-ordered slots,
// so a wavefront covers a compact patch of cells (8x8-cell tiles by default)
// and the candidate offsets it reads come from the same few cache lines.
// Replaces gpu_voronoi_calc_global_kernel + gpu_get_neighbors_global_kernel
// (DelaunayGPU.cu:649-675, 981-1006) and their masked repair variants
// (DelaunayGPU.cu:611-640, 945-975), which instead launch every point and
// mask most of them out.
#include "gdt_exact_ring.h"
#include "gdt_kernels.h"
#include "gdt_wavedev.h"

#include <type_traits>

// A point whose search the grid's search radius cannot certify (boxes of a
// few cells) is appended to `defer` (when given) for the wavefront kernel,
// whose searches may stop at the cell box of their circle; its row is then
// rewritten there.
template <int PREC>
__device__ inline void one_ring_lane(bool active, const GdtSet& S, int j, const GdtRingIO& io,
                                     int* __restrict__ counts, int* __restrict__ status,
                                     unsigned long long* __restrict__ counters,
                                     int* __restrict__ defer)
{
    GdtRingResult r = {0, 0, 0u, 0u};
    if (active) {
        if constexpr (PREC == GDT_EXACT)
            r = gdt_one_ring_exact(S, j, io);
        else
            r = (PREC == 32) ? gdt_one_ring_f32(S, j, io) : gdt_one_ring_f64(S, j, io);
        if (defer && (r.err & ERR_SEARCH)) {
            defer[atomicAdd(reinterpret_cast<unsigned*>(status + ST_NDEFER), 1u)] = j;
            r.count = 0;
            r.err = 0;
        } else {
            counts[io.perm[j]] = r.count;
        }
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
                                                   unsigned long long* __restrict__ counters,
                                                   int* __restrict__ defer)
{
    const GdtChunk ch = chunks[blockIdx.x];
    const GdtSet S = sets[ch.set];
    const int li = ch.start + (int)threadIdx.x;
    const bool active = li < S.n;
    one_ring_lane<PREC>(active, S, S.p0 + li, io, counts, status, counters, defer);
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
                                                        unsigned long long* __restrict__ counters,
                                                        int* __restrict__ defer)
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
        one_ring_lane<PREC>(active, S, j, io, counts, status, counters, defer);
    }
}

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

hipError_t gdt_launch_deferred(int precision, const GdtSet* d_sets, int nsets, int total_n,
                               const GdtRingIO& io, int* counts, int* status,
                               unsigned long long* counters, int* defer, hipStream_t s)
{
    // the count is read on the device: the first GDT_WAVE_CAP one wavefront
    // each, any beyond (dense clusters) one lane each
    const uint32_t* ndefer = reinterpret_cast<const uint32_t*>(status + ST_NDEFER);
    hipLaunchKernelGGL(k_one_ring_wave, dim3(GDT_WAVE_BLOCKS), dim3(256), 0, s, d_sets, nsets, defer,
                       ndefer, GDT_WAVE_CAP, io, counts, status, counters);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return gdt_launch_one_ring(precision, d_sets, nsets, nullptr, 0, 256, defer, ndefer,
                               GDT_WAVE_CAP, total_n, io, counts, status, counters, nullptr, s);
}

hipError_t gdt_launch_one_ring(int precision, const GdtSet* d_sets, int nsets,
                               const GdtChunk* chunks, int nch, int block, const int* worklist,
                               const uint32_t* worklen, int skip, int max_work,
                               const GdtRingIO& io, int* counts, int* status,
                               unsigned long long* counters, int* defer, hipStream_t s)
{
    if (!worklist) {
        if (nch <= 0) return hipSuccess;
        if (precision == 32)
            hipLaunchKernelGGL(k_one_ring<32>, dim3(nch), dim3(block), 0, s, d_sets, chunks, io,
                               counts, status, counters, defer);
        else if (precision == GDT_EXACT)
            hipLaunchKernelGGL(k_one_ring<GDT_EXACT>, dim3(nch), dim3(block), 0, s, d_sets,
                               chunks, io, counts, status, counters, defer);
        else
            hipLaunchKernelGGL(k_one_ring<64>, dim3(nch), dim3(block), 0, s, d_sets, chunks, io,
                               counts, status, counters, defer);
        return hipGetLastError();
    }
    max_work -= skip;
    if (max_work <= 0) return hipSuccess;
    int grid = (max_work + block - 1) / block;
    if (grid > 4096) grid = 4096;
    if (precision == 32)
        hipLaunchKernelGGL(k_one_ring_list<32>, dim3(grid), dim3(block), 0, s, d_sets, nsets,
                           worklist, worklen, skip, io, counts, status, counters, defer);
    else if (precision == GDT_EXACT)
        hipLaunchKernelGGL(k_one_ring_list<GDT_EXACT>, dim3(grid), dim3(block), 0, s, d_sets,
                           nsets, worklist, worklen, skip, io, counts, status, counters, defer);
    else
        hipLaunchKernelGGL(k_one_ring_list<64>, dim3(grid), dim3(block), 0, s, d_sets, nsets,
                           worklist, worklen, skip, io, counts, status, counters, defer);
    return hipGetLastError();
}
