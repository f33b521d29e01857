// Host build of the per-point device routines (gdt_point.h) for CPU tests.
// Builds the cell-ordered arrays the GPU cell-list kernels build, then runs
// the FP64 or FP32-filtered one-ring routine for every point.
// Usage: emul <points.bin> <n> <lx> <ly> <cell_size> <tiled> <precision 64|32> <stride> <out.bin>
// out.bin: int32 counts[n], int32 nbrs[n*stride], uint64 npred, uint64 nfb, int32 errflags
#include "gdt_point.h"

#include <cstdio>
#include <cstdlib>
#include <vector>

int main(int argc, char** argv)
{
    if (argc != 10) { std::fprintf(stderr, "bad args\n"); return 2; }
    const int n = std::atoi(argv[2]);
    const double lx = std::atof(argv[3]), ly = std::atof(argv[4]), a = std::atof(argv[5]);
    const int tiled = std::atoi(argv[6]), prec = std::atoi(argv[7]), stride = std::atoi(argv[8]);
    std::vector<double> xy(2 * (size_t)n);
    FILE* f = std::fopen(argv[1], "rb");
    if (!f || std::fread(xy.data(), sizeof(double), xy.size(), f) != xy.size()) return 3;
    std::fclose(f);

    GdtSet S = gdt_make_set(0, n, 0, lx, ly, a, tiled);
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
    unsigned long long npred = 0, nfb = 0;
    int err = 0;
    for (int j = 0; j < n; ++j) {
        GdtRingResult r = prec == 32 ? gdt_one_ring_f32(S, j, io) : gdt_one_ring_f64(S, j, io);
        counts[perm[j]] = r.count;
        npred += r.npred; nfb += r.nfb; err |= r.err;
    }
    f = std::fopen(argv[9], "wb");
    if (!f) return 4;
    std::fwrite(counts.data(), sizeof(int), n, f);
    std::fwrite(nbrs.data(), sizeof(int), nbrs.size(), f);
    std::fwrite(&npred, 8, 1, f);
    std::fwrite(&nfb, 8, 1, f);
    std::fwrite(&err, 4, 1, f);
    std::fclose(f);
    return 0;
}
