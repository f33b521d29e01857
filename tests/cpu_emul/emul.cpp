// Host build of the per-point device routines for CPU tests.
// Builds the cell-ordered arrays the GPU cell-list kernels build, then runs
// the direct per-point routine (gdt_point.h), the LDS-staged tile routine
// (gdt_tile.h, one emulated block per tile, deferred points through the
// wavefront routine), the wavefront routine (gdt_wave.h, 64 emulated lanes)
// or the star kernel's steps (tile.hip k_one_ring_star: 64 emulated lanes,
// each running gdt_tile_search_f64 from one candidate of the shared block
// table, the ring walked over their results; FP64 whatever the precision)
// for every point.
// Usage: emul <points.bin> <n> <lx> <ly> <cell_size> <tiled> <precision 64|32> <stride>
//             <out.bin> [mode 0=direct 1=tile 2=wave 3=star] [audit.bin]
// A general box: <lx> = "box:x11,x12,x21,x22" (row-major cells, <ly> ignored).
// out.bin: int32 counts[n], int32 nbrs[n*stride], uint64 npred, uint64 nfb, int32 errflags
// Built with -DGDT_AUDIT_FILTER, every decision the FP32 filter takes from its
// FP32 value alone goes to audit.bin: int32 kind, int32 claimed sign, then the
// FP64 displacements ax, ay, bx, by, cx, cy (GDT_FILTER_AUDIT in gdt_internal.h).
#include <vector>
#ifdef GDT_AUDIT_FILTER
struct AuditRec {
    int kind, sign;
    double ax, ay, bx, by, cx, cy;
};
static std::vector<AuditRec> g_audit;
#define GDT_FILTER_AUDIT(kind, sign, ax, ay, bx, by, cx, cy) \
    g_audit.push_back(AuditRec{(kind), (sign), (ax), (ay), (bx), (by), (cx), (cy)})
#endif
#include "gdt_tile.h"
#include "gdt_wave.h"

#include <cstdio>
#include <cstdlib>
#include <vector>

// LDS entry k as the lane's current neighbour, as the star kernel sets it.
static void gdt_tile_set_a_host(const GdtTileLds& L, GdtTileLane<double>& st, int k)
{
    const gdt_d2 X = L.X64[k];
    st.ak = k;
    st.ax = X.x - st.vx;
    st.ay = X.y - st.vy;
}

// The star kernel's work for the point in slot j, one emulated lane after the
// other (the kernel's ballots and shuffles become searches over the lanes'
// arrays).  ERR_DEFER: the point goes on to the wavefront routine.
static GdtRingResult star_ring(const GdtSet& S, const GdtTileWin& w, const GdtTileLds& L, int j,
                               const GdtRingIO& io)
{
    GdtRingResult res = {0, 0, 0u, 0u};
    uint32_t pre[2 * GDT_SCAN + 2] = {}, own[2 * GDT_SCAN + 2] = {};
    GdtTileLane<double> st;
    st.tbl = own;
    gdt_tile_lane_init<double>(w, L, j, io, S, st);
    const GdtClip whole = {0.f, 0.f, 0};
    int ph = GDT_SCAN;
    int tot = (int)gdt_block_rows(L, st.c0, ph, whole, pre);
    if (tot > 64) {
        ph = GDT_SCAN0;
        tot = (int)gdt_block_rows(L, st.c0, ph, whole, pre);
    }
    int rc = STEP_RUN;
    if (tot > 64) {
        while (rc == STEP_RUN) rc = gdt_tile_step_f64<1>(S, L, io, st, res.npred);
    } else {
        int kg[64], bk[64], lrc[64];
        for (int acc = 0, r = 0; acc < tot; ++r) {
            const int b = (int)(pre[r] & 0xffffu), n = (int)(pre[r] >> 16) - b;
            for (int g = acc; g < acc + n; ++g) kg[g] = b + g - acc;
            acc += n;
        }
        for (int g = 0; g < tot; ++g) {
            GdtTileLane<double> sl = st;
            if (kg[g] != st.iv) gdt_tile_set_a_host(L, sl, kg[g]);
            double bx, by;
            // (no lane extends its search, the point's own included: as the kernel)
            lrc[g] = gdt_tile_search_f64<1>(S, L, sl, res.npred, 0, false, bk[g], bx, by, pre,
                                            (unsigned)tot, ph);
        }
        auto lane_of = [&](int k) {
            for (int g = 0; g < tot; ++g)
                if (kg[g] == k) return g;
            return -1;
        };
        // the full search from entry a (a < 0: the nearest neighbour)
        auto full = [&](int a, int& b) {
            GdtTileLane<double> sl = st;
            if (a >= 0) gdt_tile_set_a_host(L, sl, a);
            double bx, by;
            return gdt_tile_search_f64<1>(S, L, sl, res.npred, 0, true, b, bx, by);
        };
        const int hv = lane_of(st.iv);
        int nn = bk[hv], nrc = lrc[hv];
        if (nrc != STEP_RUN) nrc = full(-1, nn);
        if (nrc != STEP_RUN) {
            rc = STEP_DEFER;
        } else {
            st.row[0] = L.orig[nn];
            st.count = 1;
            for (int cur = nn;;) {
                const int h = lane_of(cur);
                int nx;
                if (h >= 0 && lrc[h] == STEP_RUN) {
                    nx = bk[h];
                } else if (full(cur, nx) != STEP_RUN) {
                    rc = STEP_DEFER;
                    break;
                }
                if (nx == nn) {
                    rc = STEP_DONE;
                    break;
                }
                if (st.count < io.stride) st.row[st.count] = L.orig[nx];
                if (++st.count > GDT_HARD_CAP) {
                    rc = STEP_NOCLOSE;
                    break;
                }
                cur = nx;
            }
        }
    }
    res.count = st.count;
    if (rc == STEP_DEFER) res.err |= ERR_DEFER;
    if (rc == STEP_NOCLOSE) res.err |= ERR_NOCLOSE;
    return res;
}

int main(int argc, char** argv)
{
    if (argc < 10 || argc > 12) { std::fprintf(stderr, "bad args\n"); return 2; }
    const int n = std::atoi(argv[2]);
    double box[4] = {0., 0., 0., 0.};
    const bool general = std::sscanf(argv[3], "box:%lf,%lf,%lf,%lf", &box[0], &box[1], &box[2],
                                     &box[3]) == 4;
    const double lx = general ? 0. : std::atof(argv[3]), ly = general ? 0. : std::atof(argv[4]);
    const double a = std::atof(argv[5]);
    const int tiled = std::atoi(argv[6]), prec = std::atoi(argv[7]), stride = std::atoi(argv[8]);
    const int mode = argc >= 11 ? std::atoi(argv[10]) : 0;
    if ((mode == 1 || mode == 3) && !tiled) {
        std::fprintf(stderr, "tile and star modes need tiled order\n");
        return 2;
    }
    std::vector<double> xy(2 * (size_t)n);
    FILE* f = std::fopen(argv[1], "rb");
    if (!f || std::fread(xy.data(), sizeof(double), xy.size(), f) != xy.size()) return 3;
    std::fclose(f);

    if (general && tiled) { std::fprintf(stderr, "general boxes use row-major cells\n"); return 2; }
    GdtSet S = general ? gdt_make_set_box(0, n, 0, box, a, 0) : gdt_make_set(0, n, 0, lx, ly, a, tiled);
    const int K = S.nkeys;
    std::vector<uint32_t> cnt(K, 0), cstart(K + 1, 0);
    std::vector<int> key(n), perm(n);
    std::vector<gdt_d2> off64(n);
    std::vector<gdt_f2> off32(n);
    std::vector<uint32_t> scell(n);
    for (int i = 0; i < n; ++i) {
        int cx, cy; double ox, oy;
        gdt_bin_point(S, xy[2 * i], xy[2 * i + 1], cx, cy, ox, oy);
        key[i] = gdt_cell_key(S, cx, cy);
        cnt[key[i]]++;
    }
    for (int k = 0; k < K; ++k) cstart[k + 1] = cstart[k] + cnt[k];
    std::vector<uint32_t> fill(K, 0);
    for (int i = 0; i < n; ++i) perm[cstart[key[i]] + fill[key[i]]++] = i;   // index order in cell
    for (int j = 0; j < n; ++j) {
        const int i = perm[j];
        int cx, cy; double ox, oy;
        gdt_bin_point(S, xy[2 * i], xy[2 * i + 1], cx, cy, ox, oy);
        off64[j].x = ox; off64[j].y = oy;
        off32[j].x = (float)ox; off32[j].y = (float)oy;
        scell[j] = (uint32_t)cx | ((uint32_t)cy << 16);
    }
    std::vector<int> nbrs((size_t)n * stride, -1), counts(n, 0);
    GdtRingIO io;
    io.cstart = cstart.data(); io.perm = perm.data(); io.off64 = off64.data();
    io.off32 = off32.data(); io.scell = scell.data(); io.nbrs = nbrs.data(); io.stride = stride;
    io.xy = nullptr;
    unsigned long long npred = 0, nfb = 0;
    int err = 0, ndefer = 0;
    auto acc = [&](int j, const GdtRingResult& r) {
        counts[perm[j]] = r.count;
        npred += r.npred; nfb += r.nfb; err |= r.err;
    };
    GdtWaveHost wave;
    if (mode == 0) {
        for (int j = 0; j < n; ++j)
            acc(j, prec == 32 ? gdt_one_ring_f32(S, j, io) : gdt_one_ring_f64(S, j, io));
    } else if (mode == 2) {
        for (int j = 0; j < n; ++j) acc(j, gdt_wave_ring(S, j, io, wave, true));
    } else {
        // one emulated block per tile: stage counts, scan, copy, then every point
        std::vector<uint32_t> cs(GDT_WIN_CELLS + 1);
        std::vector<int> gstart(GDT_WIN_CELLS), orig(GDT_WIN_MAXPTS);
        std::vector<gdt_d2> X64(GDT_WIN_MAXPTS + 1);
        std::vector<gdt_f2> X32(GDT_WIN_MAXPTS + 1);
        GdtTileLds L = {cs.data(), gstart.data(), X64.data(), X32.data(), orig.data()};
        const int ntiles = S.nkeys / GDT_TILE_CELLS;
        for (int t = 0; t < ntiles; ++t) {
            const GdtTileWin w = gdt_tile_window(S, t);
            std::vector<uint32_t> c_cnt(GDT_WIN_CELLS);
            for (int c = 0; c < GDT_WIN_CELLS; ++c)
                gdt_tile_count(S, w, c, io.cstart, gstart[c], c_cnt[c]);
            uint32_t run = 0;
            for (int c = 0; c < GDT_WIN_CELLS; ++c) { cs[c] = run; run += c_cnt[c]; }
            cs[GDT_WIN_CELLS] = run;
            const bool fits = run <= GDT_WIN_MAXPTS && S.rmax >= GDT_SCAN;
            if (fits)
                for (int c = 0; c < GDT_WIN_CELLS; ++c) {
                    if (prec == 32) gdt_tile_copy<true>(S, c, gstart[c], cs[c], c_cnt[c], io, L);
                    else gdt_tile_copy<false>(S, c, gstart[c], cs[c], c_cnt[c], io, L);
                }
            const int k0 = S.c0 + t * GDT_TILE_CELLS;
            for (int j = (int)io.cstart[k0]; j < (int)io.cstart[k0 + GDT_TILE_CELLS]; ++j) {
                GdtRingResult r;
                if (fits && mode == 3) r = star_ring(S, w, L, j, io);
                else if (fits) r = prec == 32 ? gdt_tile_ring<32>(S, w, L, j, io) : gdt_tile_ring<64>(S, w, L, j, io);
                if (!fits || (r.err & ERR_DEFER)) {
                    const unsigned np = fits ? r.npred : 0u, nf = fits ? r.nfb : 0u;
                    ++ndefer;
                    // the block's own pass has no tie rule; a tie goes on to
                    // the wavefront kernel, which has it
                    r = gdt_wave_ring<false>(S, j, io, wave, true);
                    if (r.err & ERR_TIE) r = gdt_wave_ring<true>(S, j, io, wave, true);
                    r.npred += np;
                    r.nfb += nf;
                }
                acc(j, r);
            }
        }
    }
    if (mode == 1 || mode == 3) std::fprintf(stderr, "deferred %d of %d points\n", ndefer, n);
    f = std::fopen(argv[9], "wb");
    if (!f) return 4;
    std::fwrite(counts.data(), sizeof(int), n, f);
    std::fwrite(nbrs.data(), sizeof(int), nbrs.size(), f);
    std::fwrite(&npred, 8, 1, f);
    std::fwrite(&nfb, 8, 1, f);
    std::fwrite(&err, 4, 1, f);
    std::fclose(f);
#ifdef GDT_AUDIT_FILTER
    if (argc == 12) {
        f = std::fopen(argv[11], "wb");
        if (!f) return 4;
        if (!g_audit.empty()) std::fwrite(g_audit.data(), sizeof(AuditRec), g_audit.size(), f);
        std::fclose(f);
    }
#endif
    return 0;
}
