// gdt_vor.h -- Voronoi vertices of a ring table (host/device): minimum-image
// displacements and circumcentres, shared by the Voronoi kernels and the CPU
// mode's GPUVoroCur (DelaunayGPU.cu:921-922, DelaunayGPU.h:57-58).
#pragma once

#include <type_traits>

#include "gdt_internal.h"

namespace gdt_vor {

struct Disp {
    double x, y;
};

// Minimum-image displacement x_j - x_i (points inside the box; general boxes:
// neighbours much closer than half a box thickness).
GDT_HD inline Disp vor_disp(const GdtSet& S, const double* __restrict__ xy, int i, int j)
{
    Disp d;
    d.x = xy[2 * (size_t)j] - xy[2 * (size_t)i];
    d.y = xy[2 * (size_t)j + 1] - xy[2 * (size_t)i + 1];
    if (S.general) {
        gdt_min_image(S, d.x, d.y);
        return d;
    }
    const double hx = 0.5 * S.lx, hy = 0.5 * S.ly;
    if (d.x > hx) d.x -= S.lx;
    else if (d.x < -hx) d.x += S.lx;
    if (d.y > hy) d.y -= S.ly;
    else if (d.y < -hy) d.y += S.ly;
    return d;
}

template <class R>
struct V2 {
    R x, y;
};

template <class R>
GDT_HD inline V2<R> vor_circ(V2<R> a, V2<R> b)
{
    const R an = a.x * a.x + a.y * a.y, bn = b.x * b.x + b.y * b.y;
    const R s = R(0.5) / (a.x * b.y - a.y * b.x);
    V2<R> c;
    c.x = s * (an * b.y - a.y * bn);
    c.y = s * (a.x * bn - an * b.x);
    return c;
}

template <class R>
GDT_HD inline V2<R> vor_round(const Disp& d)
{
    V2<R> v;
    v.x = (R)d.x;
    v.y = (R)d.y;
    return v;
}

template <class R>
GDT_HD inline R vor_len(R x, R y)
{
    return sqrt(x * x + y * y);
}

// Slot gradient dE_i/dh_m from the vertices m-1, m, m+1.
template <class R>
GDT_HD inline V2<R> vor_slot_grad(V2<R> hp, V2<R> hm, V2<R> hn, R ca, R cp)
{
    const R lp = vor_len<R>(hm.x - hp.x, hm.y - hp.y);
    const R ln = vor_len<R>(hn.x - hm.x, hn.y - hm.y);
    V2<R> g;
    g.x = ca * R(0.5) * (hn.y - hp.y) + cp * ((hm.x - hp.x) / lp - (hn.x - hm.x) / ln);
    g.y = ca * R(0.5) * (hp.x - hn.x) + cp * ((hm.y - hp.y) / lp - (hn.y - hm.y) / ln);
    return g;
}

template <int PREC>
using VReal = typename std::conditional<PREC == 32, float, double>::type;

// A row usable as a ring: 3 <= count <= stride, every entry a point index.
GDT_HD inline bool vor_row_ok(const int* __restrict__ row, int k, int stride, int total_n)
{
    if (k < 3 || k > stride) return false;
    for (int m = 0; m < k; ++m)
        if (row[m] < 0 || row[m] >= total_n) return false;
    return true;
}

}  // namespace gdt_vor
