// Host build of gdt_rng.h: writes md noise of (seed, step, set) for n points
// and one wrapped displacement step of the given points.
// Usage: rng <seed> <step> <set> <n> <sigma> <L> <points.bin|-> <out.bin>
// out.bin: double g[n][2], then (if points given) double xy_new[n][2]
#include "gdt_rng.h"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

int main(int argc, char** argv)
{
    if (argc != 9) { std::fprintf(stderr, "bad args\n"); return 2; }
    const uint64_t seed = std::strtoull(argv[1], nullptr, 10);
    const uint64_t step = std::strtoull(argv[2], nullptr, 10);
    const uint64_t set = std::strtoull(argv[3], nullptr, 10);
    const int n = std::atoi(argv[4]);
    const double sigma = std::atof(argv[5]), L = std::atof(argv[6]);
    std::vector<double> g(2 * (size_t)n), xy;
    const uint64_t key = gdt_md_key(seed, step, set);
    for (int i = 0; i < n; ++i) gdt_gauss_pair(key, (uint64_t)i, g[2 * i], g[2 * i + 1]);
    if (std::strcmp(argv[7], "-") != 0) {
        xy.resize(2 * (size_t)n);
        FILE* f = std::fopen(argv[7], "rb");
        if (!f || std::fread(xy.data(), sizeof(double), xy.size(), f) != xy.size()) return 3;
        std::fclose(f);
        for (int i = 0; i < 2 * n; ++i) xy[i] = gdt_wrap_coord(xy[i] + sigma * g[i], L);
    }
    FILE* f = std::fopen(argv[8], "wb");
    if (!f) return 4;
    std::fwrite(g.data(), sizeof(double), g.size(), f);
    if (!xy.empty()) std::fwrite(xy.data(), sizeof(double), xy.size(), f);
    std::fclose(f);
    return 0;
}
