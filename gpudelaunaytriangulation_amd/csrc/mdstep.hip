// mdstep.hip -- kernels of the device-only Brownian test-and-repair loop.
//
// One step: displace every point by sigma * N(0,1) per coordinate (counter
// noise of gdt_rng.h, wrapped into its box) -> triangles of the current rings
// -> cell list -> empty-circle test -> compaction -> repair -> bookkeeping
// (k_md_step_end).  Nothing in a step reads back to the host, so a step is
// captured once in a hipGraph and replayed.  The first step that fails sets
// md->halted at its end, and every kernel of the later replays exits at once
// (this one, the cell list, the ring test, and the compaction, which then
// marks no point, so the repair has nothing to do): the positions keep the
// failing step's displacement and no step runs on truncated rows.  The reference moves the points in
// its driver (triangulation.cpp:349-415) and calls
// testAndRepairDelaunayTriangulation (DelaunayGPU.cpp:567-602) once per step,
// with host synchronisations inside.
#include "gdt_kernels.h"
#include "gdt_rng.h"

__global__ __launch_bounds__(GDT_BLOCK) void k_md_displace(const GdtSet* __restrict__ sets,
                                                           const GdtChunk* __restrict__ chunks,
                                                           double* __restrict__ xy, double sigma,
                                                           unsigned long long seed,
                                                           const GdtMdState* __restrict__ md)
{
    if (md->halted) return;
    const GdtChunk ch = chunks[blockIdx.x];
    const GdtSet S = sets[ch.set];
    const int li = ch.start + (int)threadIdx.x;
    if (li >= S.n) return;
    const uint64_t key = gdt_md_key(seed, (uint64_t)md->step, (uint64_t)ch.set);
    double g0, g1;
    gdt_gauss_pair(key, (uint64_t)li, g0, g1);
    const size_t i = (size_t)(S.p0 + li);
    double x = xy[2 * i] + sigma * g0, y = xy[2 * i + 1] + sigma * g1;
    gdt_wrap_into_box(S, x, y);
    xy[2 * i] = x;
    xy[2 * i + 1] = y;
}

// End of a step (one thread): fold the step's status into the sticky words,
// check the Euler counts of the triangulation the step started from, and
// advance the step counter.  A step that found an error, a repaired ring
// longer than the row stride (its row is truncated) or an input that is not a
// triangulation halts the loop; a halted replay changes nothing.
__global__ void k_md_step_end(const int* __restrict__ status,
                              const unsigned long long* __restrict__ counters,
                              const uint32_t* __restrict__ ntri, int total_n, int stride,
                              GdtMdState* __restrict__ md)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (md->halted) return;
    const int err = status[ST_ERRFLAGS];
    const bool euler = *ntri != 2u * (uint32_t)total_n ||
                       counters[CT_INSUM] != 6ULL * (unsigned long long)total_n;
    const bool over = status[ST_REQUIRED] > stride;
    if (err || euler || over) {
        md->first_bad = md->step;
        md->halted = 1;
    }
    md->errflags |= err;
    if (euler) md->euler = 1;
    md->required = max(md->required, status[ST_REQUIRED]);
    const int nf = status[ST_NFLAGGED];
    md->flagged_total += (unsigned long long)nf;
    md->max_flagged = max(md->max_flagged, nf);
    md->step += 1;
}

__global__ void k_md_init(GdtMdState* __restrict__ md, int step0)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    md->step = step0;
    md->errflags = 0;
    md->euler = 0;
    md->required = 0;
    md->first_bad = -1;
    md->max_flagged = 0;
    md->halted = 0;
    md->pad = 0;
    md->flagged_total = 0ULL;
}

hipError_t gdt_launch_md_init(GdtMdState* md, int step0, hipStream_t s)
{
    hipLaunchKernelGGL(k_md_init, dim3(1), dim3(64), 0, s, md, step0);
    return hipGetLastError();
}

hipError_t gdt_launch_md_displace(const GdtSet* d_sets, const GdtChunk* pchunks, int npch,
                                  double* xy, double sigma, unsigned long long seed,
                                  const GdtMdState* md, hipStream_t s)
{
    hipLaunchKernelGGL(k_md_displace, dim3(npch), dim3(GDT_BLOCK), 0, s, d_sets, pchunks, xy, sigma,
                       seed, md);
    return hipGetLastError();
}

hipError_t gdt_launch_md_step_end(const int* status, const unsigned long long* counters,
                                  const uint32_t* ntri, int total_n, int stride, GdtMdState* md,
                                  hipStream_t s)
{
    hipLaunchKernelGGL(k_md_step_end, dim3(1), dim3(64), 0, s, status, counters, ntri, total_n,
                       stride, md);
    return hipGetLastError();
}
