// capi.hip -- host side of the C ABI declared in include/gpudelaunay.h:
// context (scratch buffers, set table, chunk tables), argument checking and
// the stream-ordered sequences of kernel launches.
#include "gdt_kernels.h"
#include "gpudelaunay.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <cstring>
#include <new>
#include <vector>

namespace {

template <class T>
struct DBuf {
    T* p = nullptr;
    size_t cap = 0;
    hipError_t reserve(size_t n)
    {
        if (n <= cap && p) return hipSuccess;
        if (p) { hipError_t e = hipFree(p); p = nullptr; cap = 0; if (e != hipSuccess) return e; }
        if (n == 0) n = 1;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&p), n * sizeof(T));
        if (e == hipSuccess) cap = n;
        else p = nullptr;
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct DevGuard {
    int prev = -1;
    explicit DevGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) prev = -1;
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DevGuard()
    {
        int cur = -1;
        if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
    }
};

}  // namespace

struct gdt_context {
    int device = 0;
    int num_cu = 256;                    // compute units of the device
    gdt_config cfg = {512, 1, 1.0, 0, 0, 0, 0};
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};

    std::vector<GdtSet> sets;
    int total_n = 0, total_keys = 0;
    bool cells_valid = false;
    const double* cells_xy = nullptr;            // the coordinates of the current cell list
    DBuf<GdtSet> d_sets;
    DBuf<GdtChunk> pch, kch, rch, tch;
    int npch = 0, nkch = 0, nrch = 0, rch_block = 0, ntch = 0;

    DBuf<uint32_t> pkey, prank, ccount, cstart, scell, scan_tmp, tcount, toff, mark, markscan;
    DBuf<int> perm, worklist, repair, tris, defer;
    DBuf<gdt_d2> off64;
    DBuf<gdt_f2> off32;
    DBuf<GdtKbRec> bkt;
    DBuf<uint32_t> bmat, bscan;
    DBuf<int> status;
    DBuf<unsigned long long> counters, tpart;
    // Voronoi energy: its own set table (only p0, n and the box are used)
    std::vector<GdtSet> vsets;
    DBuf<GdtSet> d_vsets;
    DBuf<double> ecell, partial, vslots;
    // Brownian loop: device state, its pinned mirror, a stream for capture
    DBuf<GdtMdState> md;
    GdtMdState* h_md = nullptr;
    hipStream_t md_stream = nullptr;
    int* h_status = nullptr;
    unsigned long long* h_counters = nullptr;
    int* hd_status = nullptr;                    // device-side pointers of the two
    unsigned long long* hd_counters = nullptr;   // pinned buffers above
    unsigned int* h_seq = nullptr;               // publish sequence number (pinned)
    unsigned int* hd_seq = nullptr;
    unsigned int pub_seq = 0;
    gdt_stats stats;
    // While a Brownian step is enqueued: the device word that halts the step's
    // kernels (&md->halted), else null.
    const int* halt = nullptr;

    GdtCells cells()
    {
        GdtCells c;
        c.pkey = pkey.p; c.prank = prank.p; c.ccount = ccount.p; c.cstart = cstart.p;
        c.perm = perm.p; c.slot_of = nullptr; c.off64 = off64.p; c.off32 = off32.p;
        c.scell = scell.p; c.scan_tmp = scan_tmp.p;
        c.bkt = bkt.p; c.bmat = bmat.p; c.bscan = bscan.p;
        c.halt = halt;
        return c;
    }
};

#define GDT_CHECK(expr)                                   \
    do {                                                  \
        hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return GDT_ERR_DEVICE;      \
    } while (0)

static bool same_sets(const std::vector<GdtSet>& a, const std::vector<GdtSet>& b)
{
    if (a.size() != b.size()) return false;
    return a.empty() || std::memcmp(a.data(), b.data(), a.size() * sizeof(GdtSet)) == 0;
}

static std::vector<GdtChunk> make_chunks(const std::vector<GdtSet>& sets, int per, bool keys)
{
    std::vector<GdtChunk> ch;
    for (int s = 0; s < (int)sets.size(); ++s) {
        const int n = keys ? sets[s].nkeys : sets[s].n;
        for (int st = 0; st < n; st += per) ch.push_back(GdtChunk{s, st});
    }
    return ch;
}

template <class T>
static hipError_t upload(DBuf<T>& buf, const std::vector<T>& v)
{
    hipError_t e = buf.reserve(v.size());
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpy(buf.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

// Install a set table: upload it, build the chunk tables and size the buffers.
static int configure(gdt_context* c, std::vector<GdtSet> sets)
{
    int c0 = 0, p0 = 0;
    for (auto& s : sets) {
        if (s.n <= 0 || s.xsize > 65535 || s.ysize > 65535 || s.xsize < 3 || s.ysize < 3)
            return GDT_ERR_INVALID;
        s.p0 = p0;
        s.c0 = c0;
        p0 += s.n;
        if ((long long)c0 + s.nkeys > 0x7ffffff0LL) return GDT_ERR_INVALID;
        c0 += s.nkeys;
    }
    if (same_sets(sets, c->sets) && c->d_sets.p) return GDT_OK;
    c->cells_valid = false;
    c->sets = sets;
    c->total_n = p0;
    c->total_keys = c0;
    GDT_CHECK(upload(c->d_sets, c->sets));
    auto pch = make_chunks(c->sets, GDT_BLOCK, false);
    auto kch = make_chunks(c->sets, GDT_BLOCK, true);
    GDT_CHECK(upload(c->pch, pch));
    GDT_CHECK(upload(c->kch, kch));
    GDT_CHECK(c->tpart.reserve(pch.size()));
    c->npch = (int)pch.size();
    c->nkch = (int)kch.size();
    // one entry per tile (tiled cell orders only)
    std::vector<GdtChunk> tch;
    for (int si = 0; si < (int)c->sets.size(); ++si)
        if (c->sets[si].tiled)
            for (int t = 0; t < c->sets[si].nkeys >> (2 * c->sets[si].tshift); ++t)
                tch.push_back(GdtChunk{si, t});
    GDT_CHECK(upload(c->tch, tch));
    c->ntch = (int)tch.size();
    GDT_CHECK(c->tpart.reserve(std::max(pch.size(), 2 * tch.size())));   // k_tri_count / k_ring_test_tile
    c->rch_block = 0;
    const size_t N = (size_t)c->total_n, K = (size_t)c->total_keys;
    GDT_CHECK(c->pkey.reserve(N));
    GDT_CHECK(c->prank.reserve(N));
    GDT_CHECK(c->perm.reserve(N));
    GDT_CHECK(c->off64.reserve(N));
    GDT_CHECK(c->off32.reserve(N));
    GDT_CHECK(c->scell.reserve(N));
    GDT_CHECK(c->defer.reserve(N));
    GDT_CHECK(c->ccount.reserve(K));
    GDT_CHECK(c->cstart.reserve(K + 1));
    GDT_CHECK(c->scan_tmp.reserve(gdt_scan_tmp_words(std::max(K, N + 1))));
    GDT_CHECK(c->bkt.reserve(N));
    GDT_CHECK(c->bmat.reserve(gdt_bucket_words(N, K)));
    GDT_CHECK(c->bscan.reserve(gdt_bucket_words(N, K)));
    return GDT_OK;
}

static int ring_chunks(gdt_context* c)
{
    if (c->rch_block == c->cfg.block && c->rch.p) return GDT_OK;
    auto rch = make_chunks(c->sets, c->cfg.block, false);
    GDT_CHECK(upload(c->rch, rch));
    c->nrch = (int)rch.size();
    c->rch_block = c->cfg.block;
    return GDT_OK;
}

// The tile edge of the configuration, or the automatic one: 16x16-cell tiles
// when the 32x32 tiling of the sets would not give every compute unit a
// second tile (small inputs: four times the blocks, each a quarter of the
// work), else 32x32.
static int tile_shift_for(const gdt_context* c, const std::vector<GdtSet>& sets)
{
    if (c->cfg.tile_shift) return c->cfg.tile_shift;
    long long tiles = 0;
    for (const auto& s : sets) tiles += (long long)(s.nkeys >> (2 * s.tshift));
    return tiles <= c->num_cu ? 4 : 5;
}

// Lanes per point of the global triangulation: the configuration's; else a
// wavefront (the star kernel) when 16x16 tiles number at most a quarter of
// the compute units and cells hold at most 1.25 points on average (so that
// the 7x7 blocks of its lanes mostly hold at most 64); two when the tiles do
// not outnumber the compute units (every block then holds one tile's points
// with lanes to spare); else one.
static int lanes_for(const gdt_context* c)
{
    if (c->cfg.lanes) return c->cfg.lanes;
    if (c->sets[0].tshift != 4 || c->ntch > c->num_cu) return 1;
    if (4 * (long long)c->ntch <= c->num_cu && 4 * (long long)c->total_n <= 5LL * c->total_keys)
        return GDT_STAR_LANES;
    return 2;
}

// Threads per block of the tile kernels: the configured count, at most 256
// for 16x16 tiles (about 256 points each).
static int tile_block(const gdt_context* c)
{
    return c->sets[0].tshift == 4 ? std::min(c->cfg.block, 256) : c->cfg.block;
}

// Box s of a caller's array: {lx, ly} pairs (words 2) or box matrices
// {x11, x12, x21, x22} with the lattice vectors as columns (words 4).
static void box_of(const double* boxes, int words, int s, double b[4])
{
    if (words == 2) {
        b[0] = boxes[2 * s];
        b[1] = b[2] = 0.;
        b[3] = boxes[2 * s + 1];
        return;
    }
    for (int k = 0; k < 4; ++k) b[k] = boxes[4 * (size_t)s + k];
}

// A usable box: finite, and a rectangle with positive sides (words 2) or an
// invertible matrix (words 4).
static bool box_ok(const double b[4], int words)
{
    for (int k = 0; k < 4; ++k)
        if (!std::isfinite(b[k])) return false;
    if (words == 2) return b[0] > 0. && b[3] > 0.;
    const double det = b[0] * b[3] - b[1] * b[2];
    return std::isfinite(det) && std::fabs(det) > 0.;
}

// General boxes run the row-major cell order and FP64 predicates (the LDS tile
// kernels and the FP32 filter's bounds are for rectangular cells).
static bool any_general(const gdt_context* c)
{
    for (const auto& s : c->sets)
        if (s.general) return true;
    return false;
}
// (general boxes run FP64; the exact mode reports them as ERR_RANGE)
static int eff_precision(const gdt_context* c, int precision)
{
    return any_general(c) && precision != GDT_PREC_EXACT ? GDT_PREC_FP64 : precision;
}

// Precision codes of the ring entry points.
static bool ring_precision_ok(int p)
{
    return p == GDT_PREC_FP64 || p == GDT_PREC_FP32F || p == GDT_PREC_EXACT;
}
static_assert((int)GDT_PREC_EXACT == (int)GDT_EXACT, "one code for the exact mode");

static int single_set_box(gdt_context* c, int n, const double* box, int words, double cell_size)
{
    if (n < 4 || !box_ok(box, words) || !(cell_size > 0.) || !std::isfinite(cell_size))
        return GDT_ERR_INVALID;
    const double a = cell_size * c->cfg.cell_scale;
    if (!gdt_box_is_rect(box))
        return configure(c, std::vector<GdtSet>(1, gdt_make_set_box(0, n, 0, box, a, 0)));
    const double lx = box[0], ly = box[3];
    std::vector<GdtSet> sets(1, gdt_make_set(0, n, 0, lx, ly, a, c->cfg.tiled, GDT_TSHIFT_MAX));
    const int ts = tile_shift_for(c, sets);
    if (ts != GDT_TSHIFT_MAX) sets[0] = gdt_make_set(0, n, 0, lx, ly, a, c->cfg.tiled, ts);
    return configure(c, sets);
}

static int single_set(gdt_context* c, int n, double lx, double ly, double cell_size)
{
    const double b[4] = {lx, 0., 0., ly};
    return single_set_box(c, n, b, 2, cell_size);
}

// Set s = caller points [offsets[s], offsets[s+1]) in box s of `boxes`
// (words per box: 2 or 4, as box_of).
static int batch_sets(gdt_context* c, int nsets, const int* offsets, const double* boxes,
                      double cell_size, int words = 2)
{
    if (nsets <= 0 || !offsets || !boxes || offsets[0] != 0 || !(cell_size > 0.) ||
        !std::isfinite(cell_size))
        return GDT_ERR_INVALID;
    bool general = false;
    for (int k = 0; k < nsets; ++k) {
        double b[4];
        box_of(boxes, words, k, b);
        if (offsets[k + 1] - offsets[k] < 4 || !box_ok(b, words)) return GDT_ERR_INVALID;
        general = general || !gdt_box_is_rect(b);
    }
    const int tiled = general ? 0 : c->cfg.tiled;
    std::vector<GdtSet> sets;
    sets.reserve((size_t)nsets);
    for (int k = 0; k < nsets; ++k) {
        double b[4];
        box_of(boxes, words, k, b);
        sets.push_back(gdt_make_set_box(offsets[k], offsets[k + 1] - offsets[k], 0, b,
                                        cell_size * c->cfg.cell_scale, tiled, GDT_TSHIFT_MAX));
    }
    const int ts = tile_shift_for(c, sets);
    if (!general && ts != GDT_TSHIFT_MAX)
        for (int k = 0; k < nsets; ++k) {
            double b[4];
            box_of(boxes, words, k, b);
            sets[(size_t)k] = gdt_make_set_box(offsets[k], sets[(size_t)k].n, 0, b,
                                               cell_size * c->cfg.cell_scale, tiled, ts);
        }
    return configure(c, sets);
}

static int reset_status(gdt_context* c, hipStream_t s)
{
    GDT_CHECK(gdt_status_reset(c->status.p, ST_WORDS, c->counters.p, CT_WORDS, s));
    return GDT_OK;
}

// Status words to the host: the publish kernel runs after every kernel of the
// call (stream order) and stores a fresh sequence number after the words; the
// host polls that number instead of sleeping in a stream synchronization, and
// asks the stream every 4096 polls so that a failed stream ends the wait.
static int read_status(gdt_context* c, hipStream_t s)
{
    const unsigned int seq = ++c->pub_seq;
    GDT_CHECK(gdt_status_publish(c->status.p, ST_WORDS, c->counters.p, CT_WORDS, c->hd_status,
                                 c->hd_counters, c->hd_seq, seq, s));
    for (unsigned long it = 1; __atomic_load_n(c->h_seq, __ATOMIC_ACQUIRE) != seq; ++it) {
        if ((it & 4095u) == 0u) {
            const hipError_t e = hipStreamQuery(s);
            if (e == hipSuccess) {
                if (__atomic_load_n(c->h_seq, __ATOMIC_ACQUIRE) != seq) return GDT_ERR_DEVICE;
                break;
            }
            if (e != hipErrorNotReady) return GDT_ERR_DEVICE;
        }
    }
    return GDT_OK;
}

static int status_from_flags(int flags)
{
    if (flags & ERR_RANGE) return GDT_ERR_INVALID;
    if (flags & (ERR_NOCLOSE | ERR_NONEIGHBOUR)) return GDT_ERR_NONGENERIC;
    if (flags & ERR_SEARCH) return GDT_ERR_SEARCH;
    return GDT_OK;
}

static void record_ring_stats(gdt_context* c)
{
    c->stats.predicates = c->h_counters[CT_PRED];
    c->stats.fallbacks = c->h_counters[CT_FALLBACK];
    c->stats.neighbours = c->h_counters[CT_NBRSUM];
    c->stats.max_ring = c->h_status[ST_REQUIRED];
    c->stats.deferred = c->h_status[ST_NDEFER] + c->h_status[ST_NDEFER_TILE];
    if (!c->sets.empty()) {
        c->stats.xsize = c->sets[0].xsize;
        c->stats.ysize = c->sets[0].ysize;
        c->stats.bsx = c->sets[0].bsx;
        c->stats.bsy = c->sets[0].bsy;
    }
}

static GdtRingIO ring_io(gdt_context* c, int* nbrs, int stride)
{
    GdtRingIO io;
    io.cstart = c->cstart.p;
    io.perm = c->perm.p;
    io.off64 = c->off64.p;
    io.off32 = c->off32.p;
    io.scell = c->scell.p;
    io.nbrs = nbrs;
    io.stride = stride;
    io.xy = reinterpret_cast<const gdt_d2*>(c->cells_xy);
    return io;
}

// The cell list of d_xy (no caller index -> slot map: nothing reads one).
// (reset: zero the status words and counters on the way)
static int build_cells(gdt_context* c, const double* d_xy, hipStream_t s, bool reset = false)
{
    if (c->cfg.profile) GDT_CHECK(hipEventRecord(c->ev[0], s));
    GdtCells cl = c->cells();
    GDT_CHECK(gdt_build_cells(c->d_sets.p, (int)c->sets.size(), c->pch.p, c->npch, c->kch.p,
                              c->nkch, c->total_n, c->total_keys, d_xy, cl, s,
                              reset ? c->status.p : nullptr, reset ? c->counters.p : nullptr));
    if (c->cfg.profile) GDT_CHECK(hipEventRecord(c->ev[1], s));
    c->cells_valid = true;
    c->cells_xy = d_xy;
    return GDT_OK;
}

// ---------------------------------------------------------------------------
extern "C" {

const char* gdt_status_string(int st)
{
    switch (st) {
        case GDT_OK: return "ok";
        case GDT_OVERFLOW: return "ring longer than the row stride";
        case GDT_ERR_INVALID: return "invalid argument";
        case GDT_ERR_DEVICE: return "HIP runtime error";
        case GDT_ERR_NONGENERIC: return "non-generic input (a ring did not close)";
        case GDT_ERR_EULER: return "not a triangulation of the torus (Euler check)";
        case GDT_ERR_SEARCH: return "search radius exhausted (box too small)";
        case GDT_ERR_NOMEM: return "out of memory";
        case GDT_ERR_SCRATCH: return "input too large for the fixed scratch of gdt_triangulate_noalloc";
        default: return "unknown status";
    }
}

int gdt_create(gdt_context** out, int device)
{
    if (!out) return GDT_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return GDT_ERR_DEVICE;
    gdt_context* c = new (std::nothrow) gdt_context();
    if (!c) return GDT_ERR_NOMEM;
    c->device = device;
    std::memset(&c->stats, 0, sizeof(c->stats));
    DevGuard g(device);
    if (c->status.reserve(ST_WORDS) != hipSuccess || c->counters.reserve(CT_WORDS) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&c->h_status), ST_WORDS * sizeof(int), 0) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&c->h_counters), CT_WORDS * sizeof(unsigned long long), 0) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void**>(&c->hd_status), c->h_status, 0) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void**>(&c->hd_counters), c->h_counters, 0) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&c->h_seq), sizeof(unsigned int), 0) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void**>(&c->hd_seq), c->h_seq, 0) != hipSuccess ||
        hipEventCreate(&c->ev[0]) != hipSuccess || hipEventCreate(&c->ev[1]) != hipSuccess ||
        hipEventCreate(&c->ev[2]) != hipSuccess || hipEventCreate(&c->ev[3]) != hipSuccess ||
        hipDeviceGetAttribute(&c->num_cu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) {
        gdt_destroy(c);
        return GDT_ERR_DEVICE;
    }
    __atomic_store_n(c->h_seq, 0u, __ATOMIC_RELEASE);
    *out = c;
    return GDT_OK;
}

void gdt_destroy(gdt_context* c)
{
    if (!c) return;
    DevGuard g(c->device);
    (void)hipDeviceSynchronize();
    c->d_sets.release(); c->pch.release(); c->kch.release(); c->rch.release(); c->tch.release();
    c->pkey.release(); c->prank.release(); c->ccount.release(); c->cstart.release();
    c->scell.release(); c->scan_tmp.release(); c->tcount.release(); c->toff.release();
    c->mark.release(); c->markscan.release(); c->perm.release();
    c->worklist.release(); c->repair.release(); c->tris.release(); c->defer.release(); c->off64.release();
    c->off32.release(); c->status.release(); c->counters.release(); c->tpart.release();
    c->d_vsets.release(); c->ecell.release(); c->partial.release(); c->vslots.release();
    c->md.release();
    c->bkt.release(); c->bmat.release(); c->bscan.release();
    if (c->h_md) (void)hipHostFree(c->h_md);
    if (c->md_stream) (void)hipStreamDestroy(c->md_stream);
    for (auto& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->h_status) (void)hipHostFree(c->h_status);
    if (c->h_counters) (void)hipHostFree(c->h_counters);
    if (c->h_seq) (void)hipHostFree(c->h_seq);
    delete c;
}

int gdt_set_config(gdt_context* c, const gdt_config* cfg)
{
    if (!c || !cfg) return GDT_ERR_INVALID;
    if (cfg->block < 64 || cfg->block > 1024 || cfg->block % 64 != 0) return GDT_ERR_INVALID;
    if (!(cfg->cell_scale >= 0.25 && cfg->cell_scale <= 8.0)) return GDT_ERR_INVALID;
    if (cfg->tile_shift != 0 && (cfg->tile_shift < GDT_TSHIFT_MIN || cfg->tile_shift > GDT_TSHIFT_MAX))
        return GDT_ERR_INVALID;
    auto lanes_ok = [](int w) { return w == 0 || w == 1 || w == 2 || w == 4; };
    if (!(lanes_ok(cfg->lanes) || cfg->lanes == GDT_STAR_LANES) || !lanes_ok(cfg->repair_lanes))
        return GDT_ERR_INVALID;
    c->cfg = *cfg;
    c->cfg.tiled = cfg->tiled == 2 ? 2 : (cfg->tiled ? 1 : 0);
    c->sets.clear();          // force a new set table (cell size / order may change)
    c->cells_valid = false;
    c->rch_block = 0;
    return GDT_OK;
}

int gdt_get_config(const gdt_context* c, gdt_config* cfg)
{
    if (!c || !cfg) return GDT_ERR_INVALID;
    *cfg = c->cfg;
    return GDT_OK;
}

int gdt_get_stats(const gdt_context* c, gdt_stats* out)
{
    if (!c || !out) return GDT_ERR_INVALID;
    *out = c->stats;
    return GDT_OK;
}

int gdt_celllist_build(gdt_context* c, const double* d_xy, int n, double lx, double ly,
                       double cell_size, void* stream)
{
    if (!c || !d_xy) return GDT_ERR_INVALID;
    DevGuard g(c->device);
    hipStream_t s = (hipStream_t)stream;
    int rc = single_set(c, n, lx, ly, cell_size);
    if (rc != GDT_OK) return rc;
    return build_cells(c, d_xy, s);
}

int gdt_celllist_build_general(gdt_context* c, const double* d_xy, int n, const double* box,
                               double cell_size, void* stream)
{
    if (!c || !d_xy || !box) return GDT_ERR_INVALID;
    DevGuard g(c->device);
    int rc = single_set_box(c, n, box, 4, cell_size);
    if (rc != GDT_OK) return rc;
    return build_cells(c, d_xy, (hipStream_t)stream);
}

int gdt_celllist_info(gdt_context* c, int* xsize, int* ysize, double* bsx, double* bsy,
                      int* max_occupancy, void* stream)
{
    if (!c || !c->cells_valid || c->sets.size() != 1) return GDT_ERR_INVALID;
    DevGuard g(c->device);
    hipStream_t s = (hipStream_t)stream;
    const GdtSet& S = c->sets[0];
    if (xsize) *xsize = S.xsize;
    if (ysize) *ysize = S.ysize;
    if (bsx) *bsx = S.bsx;
    if (bsy) *bsy = S.bsy;
    if (max_occupancy) {
        GDT_CHECK(gdt_max_occupancy(c->ccount.p, c->total_keys, c->status.p + ST_WORDS - 1, s));
        int h = 0;
        GDT_CHECK(hipMemcpyAsync(&h, c->status.p + ST_WORDS - 1, sizeof(int), hipMemcpyDeviceToHost, s));
        GDT_CHECK(hipStreamSynchronize(s));
        *max_occupancy = h;
    }
    return GDT_OK;
}

int gdt_celllist_export(gdt_context* c, unsigned int* d_cell_sizes, int* d_idxs, int nmax,
                        void* stream)
{
    if (!c || !c->cells_valid || c->sets.size() != 1 || !d_cell_sizes || !d_idxs || nmax < 1)
        return GDT_ERR_INVALID;
    DevGuard g(c->device);
    GdtCells cl = c->cells();
    GDT_CHECK(gdt_export_cells(c->d_sets.p, c->sets[0], cl, d_cell_sizes, d_idxs, nmax,
                               (hipStream_t)stream));
    return GDT_OK;
}

static int finish_rings(gdt_context* c, int stride, int* required, hipStream_t s)
{
    int rc = read_status(c, s);
    if (rc != GDT_OK) return rc;
    record_ring_stats(c);
    const int req = c->h_status[ST_REQUIRED];
    if (required) *required = req;
    rc = status_from_flags(c->h_status[ST_ERRFLAGS]);
    if (rc != GDT_OK) return rc;
    return req > stride ? GDT_OVERFLOW : GDT_OK;
}

static bool tiled_sets(const gdt_context* c)
{
    return !c->sets.empty() && c->sets[0].tiled && c->ntch > 0;
}

// Global triangulation of the installed set table.
static int triangulate_sets(gdt_context* c, const double* d_xy, int stride, int* d_nbrs,
                            int* d_counts, int precision, int* required, hipStream_t s)
{
    int rc;
    precision = eff_precision(c, precision);
    if ((rc = ring_chunks(c)) != GDT_OK) return rc;
    if ((rc = build_cells(c, d_xy, s, true)) != GDT_OK) return rc;
    if (tiled_sets(c) && precision != GDT_PREC_EXACT) {
        // A grid of at most one tile per CU runs one wave of blocks: each lane
        // (or lane group) then takes a single point (1024 threads for about
        // 1024 points of a 32x32 tile), so the block ends after its longest
        // ring instead of after two.
        const int ts = c->sets[0].tshift, lanes = lanes_for(c);
        c->stats.tile_shift = ts;
        if (lanes == GDT_STAR_LANES) {
            // one wavefront per point: about one point per wave, at most four
            // blocks per compute unit (each stages its tile's window)
            c->stats.lanes = lanes;
            const int block = 256, nw = block / 64;
            const long long avg = (c->total_n + c->ntch - 1) / c->ntch;
            int nsplit = (int)std::max(1LL, (avg + nw - 1) / nw);
            nsplit = std::min(nsplit, std::max(1, 4 * c->num_cu / c->ntch));
            GDT_CHECK(gdt_tiles::one_ring_star(ts, c->d_sets.p, (int)c->sets.size(), c->tch.p,
                                               c->ntch, nsplit, block, c->total_n,
                                               ring_io(c, d_nbrs, stride), d_counts, c->status.p,
                                               c->counters.p, c->defer.p,
                                               c->cfg.profile ? c->ev[2] : nullptr, s));
        } else {
            c->stats.lanes = lanes == 4 && precision == GDT_PREC_FP32F ? 2 : lanes;
            const int one_each = std::min(1024, (1 << (2 * ts)) * lanes);
            const int block =
                c->ntch <= c->num_cu ? std::max(tile_block(c), one_each) : tile_block(c);
            GDT_CHECK(gdt_tiles::one_ring(ts, precision, c->d_sets.p, (int)c->sets.size(),
                                          c->tch.p, c->ntch, block, lanes, c->total_n, nullptr,
                                          nullptr, nullptr, nullptr, ring_io(c, d_nbrs, stride),
                                          d_counts, c->status.p, c->counters.p, c->defer.p,
                                          c->cfg.profile ? c->ev[2] : nullptr, s));
        }
    } else {
        c->stats.tile_shift = 0;
        c->stats.lanes = 1;
        const GdtRingIO io = ring_io(c, d_nbrs, stride);
        // (the exact mode hands no point on: one beyond the search radius is an error)
        const bool exact = precision == GDT_PREC_EXACT;
        GDT_CHECK(gdt_launch_one_ring(precision, c->d_sets.p, (int)c->sets.size(), c->rch.p, c->nrch,
                                      c->cfg.block, nullptr, nullptr, 0, 0, io, d_counts,
                                      c->status.p, c->counters.p, exact ? nullptr : c->defer.p, s));
        if (c->cfg.profile) GDT_CHECK(hipEventRecord(c->ev[2], s));
        if (!exact)
            GDT_CHECK(gdt_launch_deferred(precision, c->d_sets.p, (int)c->sets.size(), c->total_n,
                                          io, d_counts, c->status.p, c->counters.p, c->defer.p, s));
    }
    if (c->cfg.profile) GDT_CHECK(hipEventRecord(c->ev[3], s));
    rc = finish_rings(c, stride, required, s);
    if (c->cfg.profile) {
        // The status words reach the host before the runtime need have seen
        // the events complete (an elapsed time read then is "not ready"): wait
        // for the last one, which the stream completes after the others.
        GDT_CHECK(hipEventSynchronize(c->ev[3]));
        float a = 0.f, b = 0.f, d = 0.f;
        GDT_CHECK(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
        GDT_CHECK(hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
        GDT_CHECK(hipEventElapsedTime(&d, c->ev[2], c->ev[3]));
        c->stats.ms_cells = a;
        c->stats.ms_ring = b;
        c->stats.ms_defer = d;
        ++c->stats.profiled_calls;
        c->stats.ms_cells_total += a;
        c->stats.ms_ring_total += b;
        c->stats.ms_defer_total += d;
    }
    return rc;
}

int gdt_triangulate(gdt_context* c, const double* d_xy, int n, double lx, double ly,
                    double cell_size, int stride, int* d_nbrs, int* d_counts, int precision,
                    int* required, void* stream)
{
    if (!c || !d_xy || !d_nbrs || !d_counts || stride < 3) return GDT_ERR_INVALID;
    if (!ring_precision_ok(precision)) return GDT_ERR_INVALID;
    DevGuard g(c->device);
    int rc = single_set(c, n, lx, ly, cell_size);
    if (rc != GDT_OK) return rc;
    return triangulate_sets(c, d_xy, stride, d_nbrs, d_counts, precision, required,
                            (hipStream_t)stream);
}

int gdt_triangulate_batch(gdt_context* c, const double* d_xy, int nsets, const int* offsets,
                          const double* boxes, double cell_size, int stride, int* d_nbrs,
                          int* d_counts, int precision, int* required, void* stream)
{
    if (!c || !d_xy || !d_nbrs || !d_counts || stride < 3) return GDT_ERR_INVALID;
    if (!ring_precision_ok(precision)) return GDT_ERR_INVALID;
    DevGuard g(c->device);
    int rc = batch_sets(c, nsets, offsets, boxes, cell_size);
    if (rc != GDT_OK) return rc;
    return triangulate_sets(c, d_xy, stride, d_nbrs, d_counts, precision, required,
                            (hipStream_t)stream);
}

int gdt_triangulate_general(gdt_context* c, const double* d_xy, int nsets, const int* offsets,
                            const double* boxes, double cell_size, int stride, int* d_nbrs,
                            int* d_counts, int precision, int* required, void* stream)
{
    if (!c || !d_xy || !d_nbrs || !d_counts || stride < 3) return GDT_ERR_INVALID;
    if (!ring_precision_ok(precision)) return GDT_ERR_INVALID;
    DevGuard g(c->device);
    int rc = batch_sets(c, nsets, offsets, boxes, cell_size, 4);
    if (rc != GDT_OK) return rc;
    return triangulate_sets(c, d_xy, stride, d_nbrs, d_counts, precision, required,
                            (hipStream_t)stream);
}

// Owned triangles into `tris` (capacity cap); count stays on the device at
// toff[N].  slot_order: owners in cell-slot order (the current cell list's).
static int enumerate(gdt_context* c, const int* d_nbrs, const int* d_counts, int stride,
                     int* tris, int cap, bool slot_order, hipStream_t s)
{
    GDT_CHECK(c->tcount.reserve((size_t)c->total_n));
    GDT_CHECK(c->toff.reserve((size_t)c->total_n + 1));
    GDT_CHECK(gdt_launch_triangles(c->d_sets.p, c->pch.p, c->npch, c->total_n,
                                   slot_order ? c->perm.p : nullptr, d_nbrs, d_counts, stride,
                                   c->tcount.p, c->toff.p, tris, cap, c->scan_tmp.p, c->tpart.p,
                                   c->counters.p, c->status.p, c->halt, s));
    return GDT_OK;
}

// Empty-circle test of the triangles enumerate() wrote.
static int test_triangles(gdt_context* c, const double* d_xy, int cap, hipStream_t s)
{
    const int n = c->total_n;
    GDT_CHECK(gdt_fill_i32(c->repair.p, -1, (size_t)n, s));
    GDT_CHECK(gdt_launch_test(c->d_sets.p, (int)c->sets.size(), c->tris.p, c->toff.p + n, cap,
                              n, d_xy, c->cells(), c->repair.p, s));
    return GDT_OK;
}

// Triangles of the current rings and their empty-circle test: with the tiled
// cell order in one pass over the rings (no triangle array), else enumerate()
// then test_triangles().  Leaves the triangle count in toff[n] and the sum of
// counts in counters[CT_INSUM] either way.
static int triangles_and_test(gdt_context* c, const double* d_xy, const int* d_nbrs,
                              const int* d_counts, int stride, int cap, hipStream_t s)
{
    const int n = c->total_n;
    if (!tiled_sets(c)) {
        int rc;
        if ((rc = enumerate(c, d_nbrs, d_counts, stride, c->tris.p, cap, false, s)) != GDT_OK) return rc;
        return test_triangles(c, d_xy, cap, s);
    }
    GDT_CHECK(c->toff.reserve((size_t)n + 1));
    GDT_CHECK(gdt_fill_i32(c->repair.p, -1, (size_t)n, s));
    GDT_CHECK(gdt_tiles::ring_test(c->sets[0].tshift, c->d_sets.p, c->tch.p, c->ntch, d_nbrs,
                                   d_counts, stride, n, d_xy, c->cells(), c->repair.p, c->tpart.p,
                                   c->toff.p + n, c->counters.p, c->status.p, s));
    return GDT_OK;
}

int gdt_triangles(gdt_context* c, const int* d_nbrs, const int* d_counts, int n, int stride,
                  int* d_tris, int cap, int* ntri, long long* total_nbrs, void* stream)
{
    if (!c || !d_nbrs || !d_counts || !d_tris || stride < 3 || cap < 0) return GDT_ERR_INVALID;
    DevGuard g(c->device);
    hipStream_t s = (hipStream_t)stream;
    if (c->sets.empty() || c->total_n != n) {
        // geometry is irrelevant for enumeration: use a unit-density placeholder set
        const double L = std::sqrt((double)std::max(n, 16));
        int rc = single_set(c, n, L, L, 1.0);
        if (rc != GDT_OK) return rc;
    }
    int rc = reset_status(c, s);
    if (rc != GDT_OK) return rc;
    if ((rc = enumerate(c, d_nbrs, d_counts, stride, d_tris, cap, false, s)) != GDT_OK) return rc;
    GDT_CHECK(gdt_copy_u32(c->toff.p + c->total_n, reinterpret_cast<uint32_t*>(c->status.p + ST_NTRI), s));
    if ((rc = read_status(c, s)) != GDT_OK) return rc;
    const unsigned int h_nt = (unsigned int)c->h_status[ST_NTRI];
    if (ntri) *ntri = (int)h_nt;
    if (total_nbrs) *total_nbrs = (long long)c->h_counters[CT_INSUM];
    return (int)h_nt > cap ? GDT_ERR_INVALID : GDT_OK;
}

int gdt_test_circumcircles(gdt_context* c, const double* d_xy, int n, const int* d_tris, int ntri,
                           int* d_repair, int* nflagged, void* stream)
{
    if (!c || !d_xy || !d_tris || !d_repair || ntri < 0) return GDT_ERR_INVALID;
    if (!c->cells_valid || c->total_n != n) return GDT_ERR_INVALID;
    DevGuard g(c->device);
    hipStream_t s = (hipStream_t)stream;
    GDT_CHECK(gdt_fill_i32(d_repair, -1, (size_t)n, s));
    GDT_CHECK(gdt_fill_u32(reinterpret_cast<uint32_t*>(c->status.p + ST_NTRI), (uint32_t)ntri, 1, s));
    GdtCells cl = c->cells();
    GDT_CHECK(gdt_launch_test(c->d_sets.p, (int)c->sets.size(), d_tris,
                              reinterpret_cast<const uint32_t*>(c->status.p + ST_NTRI), ntri,
                              c->total_n, d_xy, cl, d_repair, s));
    if (nflagged) {
        GDT_CHECK(c->mark.reserve((size_t)n));
        GDT_CHECK(c->markscan.reserve((size_t)n + 1));
        GDT_CHECK(c->worklist.reserve((size_t)n));
        GDT_CHECK(gdt_launch_compact(d_repair, c->perm.p, n, c->mark.p, c->markscan.p,
                                     c->worklist.p, reinterpret_cast<uint32_t*>(c->status.p + ST_NFLAGGED),
                                     c->scan_tmp.p, nullptr, s));
        int h = 0;
        GDT_CHECK(hipMemcpyAsync(&h, c->status.p + ST_NFLAGGED, sizeof(int), hipMemcpyDeviceToHost, s));
        GDT_CHECK(hipStreamSynchronize(s));
        *nflagged = h;
        c->stats.flagged = h;
    }
    return GDT_OK;
}

// Redo the rings of the flagged points: with the tiled cell order through
// the LDS tile kernel, each tile reading its own slots' flags (tiles without
// any exit before staging), else the flags compacted (slot order) and one
// lane each.  status[ST_NDEFER] and status[ST_NFLAGGED] must be zero on entry
// (reset_status).
static int repair_flagged(gdt_context* c, const int* d_repair, int stride, int* d_nbrs,
                          int* d_counts, int precision, hipStream_t s)
{
    const bool exact = precision == GDT_PREC_EXACT;
    if (tiled_sets(c) && !exact) {
        GDT_CHECK(gdt_tiles::one_ring(c->sets[0].tshift, precision, c->d_sets.p,
                                      (int)c->sets.size(), c->tch.p, c->ntch, tile_block(c),
                                      c->cfg.repair_lanes ? c->cfg.repair_lanes : 2, c->total_n,
                                      nullptr, nullptr, d_repair, c->halt,
                                      ring_io(c, d_nbrs, stride), d_counts, c->status.p,
                                      c->counters.p, c->defer.p, nullptr, s));
        return GDT_OK;
    }
    const size_t n = (size_t)c->total_n;
    GDT_CHECK(c->mark.reserve(n));
    GDT_CHECK(c->markscan.reserve(n + 1));
    GDT_CHECK(c->worklist.reserve(n));
    uint32_t* worklen = reinterpret_cast<uint32_t*>(c->status.p + ST_NFLAGGED);
    GDT_CHECK(gdt_launch_compact(d_repair, c->perm.p, c->total_n, c->mark.p, c->markscan.p,
                                 c->worklist.p, worklen, c->scan_tmp.p, c->halt, s));
    const GdtRingIO io = ring_io(c, d_nbrs, stride);
    GDT_CHECK(gdt_launch_one_ring(precision, c->d_sets.p, (int)c->sets.size(), nullptr, 0,
                                  c->cfg.block, c->worklist.p, worklen, 0, c->total_n, io,
                                  d_counts, c->status.p, c->counters.p,
                                  exact ? nullptr : c->defer.p, s));
    if (!exact)
        GDT_CHECK(gdt_launch_deferred(precision, c->d_sets.p, (int)c->sets.size(), c->total_n, io,
                                      d_counts, c->status.p, c->counters.p, c->defer.p, s));
    return GDT_OK;
}

int gdt_repair(gdt_context* c, const double* d_xy, int n, const int* d_repair, int stride,
               int* d_nbrs, int* d_counts, int precision, int* required, void* stream)
{
    if (!c || !d_xy || !d_repair || !d_nbrs || !d_counts || stride < 3) return GDT_ERR_INVALID;
    if (!ring_precision_ok(precision)) return GDT_ERR_INVALID;
    if (!c->cells_valid || c->total_n != n) return GDT_ERR_INVALID;
    DevGuard g(c->device);
    hipStream_t s = (hipStream_t)stream;
    precision = eff_precision(c, precision);
    int rc = reset_status(c, s);
    if (rc != GDT_OK) return rc;
    if ((rc = repair_flagged(c, d_repair, stride, d_nbrs, d_counts, precision, s)) != GDT_OK) return rc;
    rc = finish_rings(c, stride, required, s);
    c->stats.flagged = c->h_status[ST_NFLAGGED];
    return rc;
}

// Triangles -> cell list -> test -> repair for the installed set table.
static int test_and_repair_sets(gdt_context* c, const double* d_xy, int stride, int* d_nbrs,
                                int* d_counts, int precision, int* nflagged, int* required,
                                hipStream_t s)
{
    const int n = c->total_n;
    const int cap = 2 * n + 64;
    precision = eff_precision(c, precision);
    GDT_CHECK(c->tris.reserve(3 * (size_t)cap));
    GDT_CHECK(c->repair.reserve((size_t)n));
    int rc;
    if ((rc = reset_status(c, s)) != GDT_OK) return rc;
    // 1. cell list of the moved points (DelaunayGPU.cpp:580-586)
    if ((rc = build_cells(c, d_xy, s)) != GDT_OK) return rc;
    // 2. triangles of the current rings (getCircumcirclesGPU, DelaunayGPU.cpp:650-666) and
    // 3. their empty-circle test (testTriangulation, DelaunayGPU.cpp:607-648)
    if ((rc = triangles_and_test(c, d_xy, d_nbrs, d_counts, stride, cap, s)) != GDT_OK) return rc;
    // 4. repair the flagged points (locallyRepairDelaunayTriangulation, DelaunayGPU.cpp:87-115)
    if ((rc = repair_flagged(c, c->repair.p, stride, d_nbrs, d_counts, precision, s)) != GDT_OK) return rc;
    GDT_CHECK(gdt_copy_u32(c->toff.p + n, reinterpret_cast<uint32_t*>(c->status.p + ST_NTRI), s));
    if ((rc = read_status(c, s)) != GDT_OK) return rc;
    const unsigned int h_nt = (unsigned int)c->h_status[ST_NTRI];
    const int nf = c->h_status[ST_NFLAGGED];
    if (nflagged) *nflagged = nf;
    c->stats.flagged = nf;
    c->stats.predicates = c->h_counters[CT_PRED];
    c->stats.fallbacks = c->h_counters[CT_FALLBACK];
    const int req = c->h_status[ST_REQUIRED];
    if (required) *required = req;
    // Euler check on the input triangulation (getCircumcirclesCPU, DelaunayGPU.cpp:701-705)
    if (h_nt != (unsigned)(2 * n) || c->h_counters[CT_INSUM] != 6ULL * (unsigned long long)n)
        return GDT_ERR_EULER;
    rc = status_from_flags(c->h_status[ST_ERRFLAGS]);
    if (rc != GDT_OK) return rc;
    return req > stride ? GDT_OVERFLOW : GDT_OK;
}

int gdt_test_and_repair(gdt_context* c, const double* d_xy, int n, double lx, double ly,
                        double cell_size, int stride, int* d_nbrs, int* d_counts, int precision,
                        int* nflagged, int* required, void* stream)
{
    if (!c || !d_xy || !d_nbrs || !d_counts || stride < 3) return GDT_ERR_INVALID;
    if (!ring_precision_ok(precision)) return GDT_ERR_INVALID;
    DevGuard g(c->device);
    int rc = single_set(c, n, lx, ly, cell_size);
    if (rc != GDT_OK) return rc;
    return test_and_repair_sets(c, d_xy, stride, d_nbrs, d_counts, precision, nflagged, required,
                                (hipStream_t)stream);
}

int gdt_test_and_repair_batch(gdt_context* c, const double* d_xy, int nsets, const int* offsets,
                              const double* boxes, double cell_size, int stride, int* d_nbrs,
                              int* d_counts, int precision, int* nflagged, int* required,
                              void* stream)
{
    if (!c || !d_xy || !d_nbrs || !d_counts || stride < 3) return GDT_ERR_INVALID;
    if (!ring_precision_ok(precision)) return GDT_ERR_INVALID;
    DevGuard g(c->device);
    int rc = batch_sets(c, nsets, offsets, boxes, cell_size);
    if (rc != GDT_OK) return rc;
    return test_and_repair_sets(c, d_xy, stride, d_nbrs, d_counts, precision, nflagged, required,
                                (hipStream_t)stream);
}

int gdt_test_and_repair_general(gdt_context* c, const double* d_xy, int nsets, const int* offsets,
                                const double* boxes, double cell_size, int stride, int* d_nbrs,
                                int* d_counts, int precision, int* nflagged, int* required,
                                void* stream)
{
    if (!c || !d_xy || !d_nbrs || !d_counts || stride < 3) return GDT_ERR_INVALID;
    if (!ring_precision_ok(precision)) return GDT_ERR_INVALID;
    DevGuard g(c->device);
    int rc = batch_sets(c, nsets, offsets, boxes, cell_size, 4);
    if (rc != GDT_OK) return rc;
    return test_and_repair_sets(c, d_xy, stride, d_nbrs, d_counts, precision, nflagged, required,
                                (hipStream_t)stream);
}

// ---- Brownian test-and-repair loop ------------------------------------------
}  // extern "C"

// One step on stream s, device only (capturable): every buffer it touches is
// reserved beforehand by gdt_md_run.  Every kernel of the step that reads or
// writes the positions, the cell list, the rings or the loop state watches
// md->halted (set by the end of a failed step).
static int md_enqueue_step_halting(gdt_context* c, double* d_xy, int stride, int* d_nbrs,
                                   int* d_counts, int precision, double sigma,
                                   unsigned long long seed, hipStream_t s)
{
    const int n = c->total_n, cap = 2 * n + 64;
    int rc;
    GDT_CHECK(gdt_launch_md_displace(c->d_sets.p, c->pch.p, c->npch, d_xy, sigma, seed, c->md.p, s));
    if ((rc = reset_status(c, s)) != GDT_OK) return rc;
    if ((rc = build_cells(c, d_xy, s)) != GDT_OK) return rc;
    if ((rc = triangles_and_test(c, d_xy, d_nbrs, d_counts, stride, cap, s)) != GDT_OK) return rc;
    if ((rc = repair_flagged(c, c->repair.p, stride, d_nbrs, d_counts, precision, s)) != GDT_OK) return rc;
    GDT_CHECK(gdt_launch_md_step_end(c->status.p, c->counters.p, c->toff.p + n, n, stride, c->md.p, s));
    return GDT_OK;
}

static int md_enqueue_step(gdt_context* c, double* d_xy, int stride, int* d_nbrs, int* d_counts,
                           int precision, double sigma, unsigned long long seed, hipStream_t s)
{
    c->halt = &c->md.p->halted;
    const int rc = md_enqueue_step_halting(c, d_xy, stride, d_nbrs, d_counts, precision, sigma,
                                           seed, s);
    c->halt = nullptr;
    return rc;
}

// Restores the profile flag (events are not recorded inside the loop).
struct ProfileOff {
    gdt_context* c;
    int saved;
    explicit ProfileOff(gdt_context* cc) : c(cc), saved(cc->cfg.profile) { c->cfg.profile = 0; }
    ~ProfileOff() { c->cfg.profile = saved; }
};

extern "C" {

}  // extern "C"

static int md_run(gdt_context* c, double* d_xy, int nsets, const int* offsets, const double* boxes,
                  int words, double cell_size, int stride, int* d_nbrs, int* d_counts,
                  int precision, double sigma, unsigned long long seed, int step0, int nsteps,
                  int use_graph, int check_every, gdt_md_report* rep, void* stream)
{
    if (!c || !d_xy || !d_nbrs || !d_counts || !rep || stride < 3 || nsteps < 0 || step0 < 0 ||
        check_every < 1 || !(sigma >= 0.) || !std::isfinite(sigma))
        return GDT_ERR_INVALID;
    if (!ring_precision_ok(precision)) return GDT_ERR_INVALID;
    std::memset(rep, 0, sizeof(*rep));
    rep->first_bad_step = -1;
    DevGuard g(c->device);
    int rc = batch_sets(c, nsets, offsets, boxes, cell_size, words);
    if (rc != GDT_OK) return rc;
    precision = eff_precision(c, precision);
    const size_t n = (size_t)c->total_n, cap = 2 * n + 64;
    GDT_CHECK(c->tris.reserve(3 * cap));
    GDT_CHECK(c->repair.reserve(n));
    GDT_CHECK(c->tcount.reserve(n));
    GDT_CHECK(c->toff.reserve(n + 1));
    GDT_CHECK(c->mark.reserve(n));
    GDT_CHECK(c->markscan.reserve(n + 1));
    GDT_CHECK(c->worklist.reserve(n));
    GDT_CHECK(c->md.reserve(1));
    if (!c->h_md)
        GDT_CHECK(hipHostMalloc(reinterpret_cast<void**>(&c->h_md), sizeof(GdtMdState), 0));
    hipStream_t s = (hipStream_t)stream;
    if (use_graph && !s) {
        // the legacy null stream cannot be captured: run on the context's own
        if (!c->md_stream) GDT_CHECK(hipStreamCreateWithFlags(&c->md_stream, hipStreamNonBlocking));
        GDT_CHECK(hipDeviceSynchronize());
        s = c->md_stream;
    }
    ProfileOff poff(c);
    GDT_CHECK(gdt_launch_md_init(c->md.p, step0, s));
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    if (use_graph) {
        GDT_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        rc = md_enqueue_step(c, d_xy, stride, d_nbrs, d_counts, precision, sigma, seed, s);
        const hipError_t ec = hipStreamEndCapture(s, &graph);
        if (rc == GDT_OK && ec == hipSuccess &&
            hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
            rc = GDT_ERR_DEVICE;
        if (rc == GDT_OK && ec != hipSuccess) rc = GDT_ERR_DEVICE;
        if (rc != GDT_OK) {
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
    }
    int done = 0;
    hipError_t he = hipSuccess;
    while (done < nsteps && rc == GDT_OK && he == hipSuccess) {
        const int chunk = std::min(check_every, nsteps - done);
        for (int t = 0; t < chunk && rc == GDT_OK && he == hipSuccess; ++t) {
            if (use_graph) he = hipGraphLaunch(exec, s);
            else rc = md_enqueue_step(c, d_xy, stride, d_nbrs, d_counts, precision, sigma, seed, s);
        }
        done += chunk;
        if (he == hipSuccess)
            he = hipMemcpyAsync(c->h_md, c->md.p, sizeof(GdtMdState), hipMemcpyDeviceToHost, s);
        if (he == hipSuccess) he = hipStreamSynchronize(s);
        if (he == hipSuccess && rc == GDT_OK && c->h_md->halted) break;
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (rc != GDT_OK) return rc;
    if (he != hipSuccess) return GDT_ERR_DEVICE;
    const GdtMdState& m = *c->h_md;
    rep->steps_done = m.step - step0;
    rep->first_bad_step = m.first_bad;
    rep->required = m.required;
    rep->max_flagged = m.max_flagged;
    rep->skipped_steps = done - rep->steps_done;
    rep->flagged_total = (long long)m.flagged_total;
    c->stats.flagged = m.max_flagged;
    // a halted loop left the cell list of its last built step; the caller
    // re-triangulates, which rebuilds it
    if (m.halted) c->cells_valid = false;
    // a ring longer than the stride truncates its row, which the next step
    // then reports as a malformed triangulation: the overflow is the cause
    if (m.required > stride) return GDT_OVERFLOW;
    if (m.euler) return GDT_ERR_EULER;
    return status_from_flags(m.errflags);
}

extern "C" {

int gdt_md_run(gdt_context* c, double* d_xy, int nsets, const int* offsets, const double* boxes,
               double cell_size, int stride, int* d_nbrs, int* d_counts, int precision,
               double sigma, unsigned long long seed, int step0, int nsteps, int use_graph,
               int check_every, gdt_md_report* rep, void* stream)
{
    return md_run(c, d_xy, nsets, offsets, boxes, 2, cell_size, stride, d_nbrs, d_counts,
                  precision, sigma, seed, step0, nsteps, use_graph, check_every, rep, stream);
}

int gdt_md_run_general(gdt_context* c, double* d_xy, int nsets, const int* offsets,
                       const double* boxes, double cell_size, int stride, int* d_nbrs,
                       int* d_counts, int precision, double sigma, unsigned long long seed,
                       int step0, int nsteps, int use_graph, int check_every,
                       gdt_md_report* rep, void* stream)
{
    return md_run(c, d_xy, nsets, offsets, boxes, 4, cell_size, stride, d_nbrs, d_counts,
                  precision, sigma, seed, step0, nsteps, use_graph, check_every, rep, stream);
}

// ---- Voronoi energy and adjoint -------------------------------------------
}  // extern "C"

static int voronoi_sets(gdt_context* c, int nsets, const int* offsets, const double* boxes,
                        int words = 2)
{
    if (nsets <= 0 || !offsets || !boxes || offsets[0] != 0) return GDT_ERR_INVALID;
    std::vector<GdtSet> v((size_t)nsets);
    for (int s = 0; s < nsets; ++s) {
        double b[4];
        box_of(boxes, words, s, b);
        const int n = offsets[s + 1] - offsets[s];
        if (n < 3 || !box_ok(b, words)) return GDT_ERR_INVALID;
        GdtSet S;
        if (gdt_box_is_rect(b)) {
            std::memset(&S, 0, sizeof(S));
            S.lx = b[0];
            S.ly = b[3];
        } else {
            // minimum images need the reduced basis only (a one-cell grid)
            S = gdt_make_set_box(0, n, 0, b, std::numeric_limits<double>::infinity(), 0);
        }
        S.p0 = offsets[s];
        S.n = n;
        v[(size_t)s] = S;
    }
    if (same_sets(v, c->vsets) && c->d_vsets.p) return GDT_OK;
    c->vsets = v;
    GDT_CHECK(upload(c->d_vsets, c->vsets));
    return GDT_OK;
}

static int voronoi_forward(gdt_context* c, const double* d_xy, int nsets, const int* offsets,
                           const double* boxes, int words, const int* d_nbrs, const int* d_counts, int stride,
                           double a0, double p0, int precision, double* d_area, double* d_perim,
                           double* d_energy, void* stream)
{
    if (!c || !d_xy || !d_nbrs || !d_counts || !d_area || !d_perim || !d_energy || stride < 3)
        return GDT_ERR_INVALID;
    if (precision != GDT_PREC_FP64 && precision != GDT_PREC_FP32F) return GDT_ERR_INVALID;
    DevGuard g(c->device);
    int rc = voronoi_sets(c, nsets, offsets, boxes, words);
    if (rc != GDT_OK) return rc;
    const int total = offsets[nsets];
    GDT_CHECK(c->ecell.reserve((size_t)total));
    GDT_CHECK(c->partial.reserve((size_t)nsets * GDT_SUM_PARTS));
    GDT_CHECK(gdt_launch_voronoi_forward(precision, c->d_vsets.p, nsets, total, d_xy, d_nbrs,
                                         d_counts, stride, a0, p0, d_area, d_perim, c->ecell.p,
                                         c->partial.p, d_energy, (hipStream_t)stream));
    return GDT_OK;
}

static int voronoi_adjoint(gdt_context* c, const double* d_xy, int nsets, const int* offsets,
                           const double* boxes, int words, const int* d_nbrs, const int* d_counts, int stride,
                           double a0, double p0, int precision, const double* d_area,
                           const double* d_perim, const double* d_gscale, double* d_grad,
                           void* stream)
{
    if (!c || !d_xy || !d_nbrs || !d_counts || !d_area || !d_perim || !d_grad || stride < 3)
        return GDT_ERR_INVALID;
    if (precision != GDT_PREC_FP64 && precision != GDT_PREC_FP32F) return GDT_ERR_INVALID;
    DevGuard g(c->device);
    int rc = voronoi_sets(c, nsets, offsets, boxes, words);
    if (rc != GDT_OK) return rc;
    const int total = offsets[nsets];
    GDT_CHECK(c->vslots.reserve(2 * (size_t)total * (size_t)stride));
    GDT_CHECK(gdt_launch_voronoi_adjoint(precision, c->d_vsets.p, nsets, total, d_xy, d_nbrs,
                                         d_counts, stride, a0, p0, d_area, d_perim, d_gscale,
                                         c->vslots.p, d_grad, (hipStream_t)stream));
    return GDT_OK;
}

extern "C" {

int gdt_voronoi_vertices(gdt_context* c, const double* d_xy, int n, double lx, double ly,
                         const int* d_nbrs, const int* d_counts, int stride, double* d_vertices,
                         void* stream)
{
    if (!c || !d_xy || !d_nbrs || !d_counts || !d_vertices || stride < 3) return GDT_ERR_INVALID;
    DevGuard g(c->device);
    const int off[2] = {0, n};
    const double box[2] = {lx, ly};
    int rc = voronoi_sets(c, 1, off, box);
    if (rc != GDT_OK) return rc;
    GDT_CHECK(gdt_launch_voronoi_vertices(c->d_vsets.p, 1, n, d_xy, d_nbrs, d_counts, stride,
                                          d_vertices, (hipStream_t)stream));
    return GDT_OK;
}

int gdt_voronoi_forward(gdt_context* c, const double* d_xy, int n, double lx, double ly,
                        const int* d_nbrs, const int* d_counts, int stride, double a0, double p0,
                        int precision, double* d_area, double* d_perim, double* d_energy,
                        void* stream)
{
    const int off[2] = {0, n};
    const double box[2] = {lx, ly};
    return voronoi_forward(c, d_xy, 1, off, box, 2, d_nbrs, d_counts, stride, a0, p0, precision,
                           d_area, d_perim, d_energy, stream);
}

int gdt_voronoi_adjoint(gdt_context* c, const double* d_xy, int n, double lx, double ly,
                        const int* d_nbrs, const int* d_counts, int stride, double a0, double p0,
                        int precision, const double* d_area, const double* d_perim,
                        const double* d_gscale, double* d_grad, void* stream)
{
    const int off[2] = {0, n};
    const double box[2] = {lx, ly};
    return voronoi_adjoint(c, d_xy, 1, off, box, 2, d_nbrs, d_counts, stride, a0, p0, precision,
                           d_area, d_perim, d_gscale, d_grad, stream);
}

int gdt_voronoi_forward_batch(gdt_context* c, const double* d_xy, int nsets, const int* offsets,
                              const double* boxes, const int* d_nbrs, const int* d_counts,
                              int stride, double a0, double p0, int precision, double* d_area,
                              double* d_perim, double* d_energy, void* stream)
{
    return voronoi_forward(c, d_xy, nsets, offsets, boxes, 2, d_nbrs, d_counts, stride, a0, p0,
                           precision, d_area, d_perim, d_energy, stream);
}

int gdt_voronoi_adjoint_batch(gdt_context* c, const double* d_xy, int nsets, const int* offsets,
                              const double* boxes, const int* d_nbrs, const int* d_counts,
                              int stride, double a0, double p0, int precision,
                              const double* d_area, const double* d_perim, const double* d_gscale,
                              double* d_grad, void* stream)
{
    return voronoi_adjoint(c, d_xy, nsets, offsets, boxes, 2, d_nbrs, d_counts, stride, a0, p0,
                           precision, d_area, d_perim, d_gscale, d_grad, stream);
}

int gdt_voronoi_vertices_general(gdt_context* c, const double* d_xy, int n, const double* box,
                                 const int* d_nbrs, const int* d_counts, int stride,
                                 double* d_vertices, void* stream)
{
    if (!c || !d_xy || !box || !d_nbrs || !d_counts || !d_vertices || stride < 3)
        return GDT_ERR_INVALID;
    DevGuard g(c->device);
    const int off[2] = {0, n};
    int rc = voronoi_sets(c, 1, off, box, 4);
    if (rc != GDT_OK) return rc;
    GDT_CHECK(gdt_launch_voronoi_vertices(c->d_vsets.p, 1, n, d_xy, d_nbrs, d_counts, stride,
                                          d_vertices, (hipStream_t)stream));
    return GDT_OK;
}

int gdt_voronoi_forward_general(gdt_context* c, const double* d_xy, int nsets, const int* offsets,
                                const double* boxes, const int* d_nbrs, const int* d_counts,
                                int stride, double a0, double p0, int precision, double* d_area,
                                double* d_perim, double* d_energy, void* stream)
{
    return voronoi_forward(c, d_xy, nsets, offsets, boxes, 4, d_nbrs, d_counts, stride, a0, p0,
                           precision, d_area, d_perim, d_energy, stream);
}

int gdt_voronoi_adjoint_general(gdt_context* c, const double* d_xy, int nsets, const int* offsets,
                                const double* boxes, const int* d_nbrs, const int* d_counts,
                                int stride, double a0, double p0, int precision,
                                const double* d_area, const double* d_perim,
                                const double* d_gscale, double* d_grad, void* stream)
{
    return voronoi_adjoint(c, d_xy, nsets, offsets, boxes, 4, d_nbrs, d_counts, stride, a0, p0,
                           precision, d_area, d_perim, d_gscale, d_grad, stream);
}

}  // extern "C"
