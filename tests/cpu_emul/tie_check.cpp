// Host build of the exact-tie rule (gdt_internal.h) for tests/test_cpu_emulation.py:
// reads "ic pv pa pb pc ax ay bx by cx cy" lines (caller indices, then the
// displacements of a, b, c from v) and prints gdt_incircle_tie's answer (1:
// c replaces b) for each.
#include "gdt_internal.h"

#include <cstdio>
#include <cstring>

int main()
{
    char tag[8];
    while (std::scanf("%7s", tag) == 1) {
        if (std::strcmp(tag, "ic") != 0) return 2;
        int pv, pa, pb, pc;
        double ax, ay, bx, by, cx, cy;
        if (std::scanf("%d %d %d %d %lf %lf %lf %lf %lf %lf", &pv, &pa, &pb, &pc, &ax, &ay, &bx,
                       &by, &cx, &cy) != 10)
            return 2;
        std::printf("%d\n", gdt_incircle_tie(pv, pa, pb, pc, ax, ay, bx, by, cx, cy) ? 1 : 0);
    }
    return 0;
}
