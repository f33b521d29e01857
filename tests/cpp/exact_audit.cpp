// Audit build of the exact mode's host routine (hostpath.cpp compiled with
// GDT_AUDIT_EXACT by tests/test_exact_mode.py): triangulates the points of a
// binary file (int n, double lx, ly, cell, then n (x, y) pairs) in exact
// precision on one thread and prints, per predicate kind (orientation,
// in-circle, distance), the decisions taken from the FP64 value alone and how
// many of them disagree with the exact sign (an exact zero included: the tie
// rule decides those).
#include "gpudelaunay.h"

#include <cstdio>
#include <vector>

static unsigned long long g_taken[3], g_wrong[3];

void gdt_exact_audit(int kind, int claimed, int exact)
{
    ++g_taken[kind];
    if (claimed != exact) ++g_wrong[kind];
}

int main(int argc, char** argv)
{
    if (argc < 2) return 2;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int n = 0;
    double box[3];
    if (std::fread(&n, sizeof n, 1, f) != 1 || std::fread(box, sizeof(double), 3, f) != 3) return 2;
    std::vector<double> xy(2 * (size_t)n);
    if (std::fread(xy.data(), sizeof(double), xy.size(), f) != xy.size()) return 2;
    std::fclose(f);
    const int stride = 32;
    std::vector<int> nbrs((size_t)n * stride), counts(n);
    int req = 0;
    const int st = gdt_host_triangulate(xy.data(), n, box[0], box[1], box[2], stride, nbrs.data(),
                                        counts.data(), GDT_PREC_EXACT, 1, &req);
    std::printf("%d", st);
    for (int k = 0; k < 3; ++k) std::printf(" %llu %llu", g_taken[k], g_wrong[k]);
    std::printf("\n");
    return 0;
}
