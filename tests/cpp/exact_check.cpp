// Host harness of the exact-mode expansion predicates (gdt_exact.h), built
// by tests/test_exact_mode.py with g++: each input line holds the caller
// coordinates of v, a, b, c and the box sides as C99 hex floats; the output
// line holds the exact signs of orient(a, c), in-circle(a, b, c) and
// |c|^2 - |b|^2 for the displacements from v (minimum images), the overflow
// flag, and the components of a's exact x displacement as hex floats.
#include "gdt_exact.h"

#include <cstdio>
#include <cstdlib>

int main()
{
    char buf[1024];
    while (std::fgets(buf, sizeof buf, stdin)) {
        double q[10];
        char* p = buf;
        for (int i = 0; i < 10; ++i) q[i] = std::strtod(p, &p);
        const gdt_d2 v = {q[0], q[1]}, a = {q[2], q[3]}, b = {q[4], q[5]}, c = {q[6], q[7]};
        const double lx = q[8], ly = q[9];
        GdtXCtx cx;
        cx.overflow = false;
        auto disp = [&](const gdt_d2& u, GdtXVec& X) {
            double fx = u.x - v.x, fy = u.y - v.y;
            fx -= lx * std::rint(fx / lx);
            fy -= ly * std::rint(fy / ly);
            gdt_x_disp(v, u, lx, ly, fx, fy, X.x, X.y, cx);
        };
        GdtXVec A, B, C;
        disp(a, A);
        disp(b, B);
        disp(c, C);
        const int so = gdt_x_orient(A, C, cx);
        const int si = gdt_x_incircle(A, B, C, cx);
        const int sd = gdt_x_dist_cmp(C, B, cx);
        std::printf("%d %d %d %d", so, si, sd, cx.overflow ? 1 : 0);
        for (int i = 0; i < A.x.n; ++i) std::printf(" %a", A.x.t[i]);
        std::printf("\n");
    }
    return 0;
}
