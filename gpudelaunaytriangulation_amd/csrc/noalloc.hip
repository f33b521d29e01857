// noalloc.hip -- gdt_triangulate_noalloc: the global triangulation of
// gdt_triangulate with every scratch buffer in one fixed device array, for
// callers that cannot allocate device memory (code under HIP graph capture,
// tight serving loops, harnesses that forbid hipMalloc).  It takes no context:
// the set table, the tile table, the cell list, the deferred-point list and
// the status words all live in `scratch` below, which each GPU holds from the
// library's first use on it until the process ends.
//
// Every call on a GPU shares that GPU's copy of `scratch`, which keeps the
// last call's contents: calls that overlap on one GPU (two streams, two host
// threads) race and must be serialised by the caller.  Everything a call reads
// is written by the same call first (the status words and counters are zeroed
// by its first kernels).
#include "gdt_kernels.h"
#include "gpudelaunay.h"

#include <cmath>
#include <mutex>

// 96 MiB: 68 bytes per point, 8 per cell key and the bucket tables of the
// cell-list build -- 2^20 points at density 1 with cell size 1 use 80 MiB.
#define GDT_NOALLOC_SCRATCH_BYTES (96u << 20)
#define GDT_NOALLOC_MAX_DEVICES 64

static __device__ __align__(256) unsigned char scratch[GDT_NOALLOC_SCRATCH_BYTES];

namespace {

// Device address of each GPU's `scratch` (the array's name is a host-side
// stand-in, never a device address) and its compute-unit count, looked up once
// per device with that device current.
std::mutex g_mutex;
unsigned char* g_ptr[GDT_NOALLOC_MAX_DEVICES];
int g_cus[GDT_NOALLOC_MAX_DEVICES];

int scratch_of_current_device(unsigned char** ptr, int* cus)
{
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return GDT_ERR_DEVICE;
    if (dev >= GDT_NOALLOC_MAX_DEVICES) return GDT_ERR_INVALID;
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return GDT_ERR_DEVICE;
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
            return GDT_ERR_DEVICE;
        g_cus[dev] = n;
        g_ptr[dev] = static_cast<unsigned char*>(p);
    }
    *ptr = g_ptr[dev];
    *cus = g_cus[dev];
    return GDT_OK;
}

// Byte offsets of the buffers inside `scratch` (each 256-byte aligned).
struct Layout {
    size_t sets, tiles, status, counters, perm, off64, off32, scell, defer, ccount, cstart, bkt,
        bmat, bscan, total;
};

Layout layout_of(size_t n, size_t keys, size_t ntiles)
{
    Layout L;
    size_t at = 0;
    auto take = [&at](size_t bytes) {
        const size_t here = at;
        at = (at + bytes + 255) & ~(size_t)255;
        return here;
    };
    const size_t bw = gdt_bucket_words(n, keys);
    L.sets = take(sizeof(GdtSet));
    L.tiles = take(ntiles * sizeof(GdtChunk));
    L.status = take(ST_WORDS * sizeof(int));
    L.counters = take(CT_WORDS * sizeof(unsigned long long));
    L.perm = take(n * sizeof(int));
    L.off64 = take(n * sizeof(gdt_d2));
    L.off32 = take(n * sizeof(gdt_f2));
    L.scell = take(n * sizeof(uint32_t));
    L.defer = take(n * sizeof(int));
    L.ccount = take(keys * sizeof(uint32_t));
    L.cstart = take((keys + 1) * sizeof(uint32_t));
    L.bkt = take(n * sizeof(GdtKbRec));
    L.bmat = take(bw * sizeof(uint32_t));
    L.bscan = take(bw * sizeof(uint32_t));
    L.total = at;
    return L;
}

}  // namespace

// The set table and the tile table from kernel arguments (no host buffer
// whose lifetime an asynchronous copy would have to outlast).
__global__ __launch_bounds__(256) void k_noalloc_tables(GdtSet* __restrict__ sets, GdtSet s,
                                                        GdtChunk* __restrict__ tiles, int ntiles)
{
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (t == 0) sets[0] = s;
    if (t < ntiles) tiles[t] = GdtChunk{0, t};
}

extern "C" int gdt_triangulate_noalloc(const double* d_xy, int n, double lx, double ly,
                                       double cell_size, int stride, int* d_nbrs, int* d_counts,
                                       int precision, int* required, void* stream)
{
    if (!d_xy || !d_nbrs || !d_counts || stride < 3) return GDT_ERR_INVALID;
    if (precision != GDT_PREC_FP64 && precision != GDT_PREC_FP32F) return GDT_ERR_INVALID;
    if (n < 4 || !(lx > 0.) || !(ly > 0.) || !(cell_size > 0.) || !std::isfinite(lx) ||
        !std::isfinite(ly) || !std::isfinite(cell_size))
        return GDT_ERR_INVALID;
    const GdtSet S = gdt_make_set(0, n, 0, lx, ly, cell_size, 1, GDT_TILE_SHIFT);
    if (S.xsize > 65535 || S.ysize > 65535 || S.xsize < 3 || S.ysize < 3) return GDT_ERR_INVALID;
    const size_t N = (size_t)n, K = (size_t)S.nkeys;
    const int ntiles = S.nkeys >> (2 * GDT_TILE_SHIFT);
    if (!gdt_cells_two_level(N, K)) return GDT_ERR_SCRATCH;
    const Layout L = layout_of(N, K, (size_t)ntiles);
    if (L.total > (size_t)GDT_NOALLOC_SCRATCH_BYTES) return GDT_ERR_SCRATCH;

    unsigned char* base = nullptr;
    int cus = 0;
    int rc = scratch_of_current_device(&base, &cus);
    if (rc != GDT_OK) return rc;
    hipStream_t s = (hipStream_t)stream;
    GdtSet* d_sets = reinterpret_cast<GdtSet*>(base + L.sets);
    GdtChunk* d_tiles = reinterpret_cast<GdtChunk*>(base + L.tiles);
    int* status = reinterpret_cast<int*>(base + L.status);
    unsigned long long* counters = reinterpret_cast<unsigned long long*>(base + L.counters);

    GdtCells c;
    c.pkey = nullptr;
    c.prank = nullptr;
    c.scan_tmp = nullptr;
    c.slot_of = nullptr;
    c.halt = nullptr;
    c.ccount = reinterpret_cast<uint32_t*>(base + L.ccount);
    c.cstart = reinterpret_cast<uint32_t*>(base + L.cstart);
    c.perm = reinterpret_cast<int*>(base + L.perm);
    c.off64 = reinterpret_cast<gdt_d2*>(base + L.off64);
    c.off32 = reinterpret_cast<gdt_f2*>(base + L.off32);
    c.scell = reinterpret_cast<uint32_t*>(base + L.scell);
    c.bkt = reinterpret_cast<GdtKbRec*>(base + L.bkt);
    c.bmat = reinterpret_cast<uint32_t*>(base + L.bmat);
    c.bscan = reinterpret_cast<uint32_t*>(base + L.bscan);

    hipLaunchKernelGGL(k_noalloc_tables, dim3((unsigned)((ntiles + 255) / 256)), dim3(256), 0, s,
                       d_sets, S, d_tiles, ntiles);
    if (hipGetLastError() != hipSuccess) return GDT_ERR_DEVICE;
    if (gdt_build_cells(d_sets, 1, nullptr, 0, nullptr, 0, n, S.nkeys, d_xy, c, s) != hipSuccess)
        return GDT_ERR_DEVICE;
    if (gdt_status_reset(status, ST_WORDS, counters, CT_WORDS, s) != hipSuccess)
        return GDT_ERR_DEVICE;
    GdtRingIO io;
    io.cstart = c.cstart;
    io.perm = c.perm;
    io.off64 = c.off64;
    io.off32 = c.off32;
    io.scell = c.scell;
    io.nbrs = d_nbrs;
    io.stride = stride;
    io.xy = nullptr;
    // the block size of gdt_triangulate's default configuration
    const int block = ntiles <= cus ? 1024 : 512;
    if (gdt_tiles::one_ring(GDT_TILE_SHIFT, precision, d_sets, 1, d_tiles, ntiles, block, 1, n,
                            nullptr, nullptr, nullptr, nullptr, io, d_counts, status, counters,
                            reinterpret_cast<int*>(base + L.defer), nullptr, s) != hipSuccess)
        return GDT_ERR_DEVICE;
    int h_status[ST_WORDS];
    if (hipMemcpyAsync(h_status, status, sizeof(h_status), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
        return GDT_ERR_DEVICE;
    const int req = h_status[ST_REQUIRED];
    if (required) *required = req;
    const int err = h_status[ST_ERRFLAGS];
    if (err & (ERR_NOCLOSE | ERR_NONEIGHBOUR)) return GDT_ERR_NONGENERIC;
    if (err & ERR_SEARCH) return GDT_ERR_SEARCH;
    return req > stride ? GDT_OVERFLOW : GDT_OK;
}
