// voronoi.hip -- Voronoi cell areas and perimeters of the tessellation dual to
// a ring table, the energy E = sum_i (A_i - a0)^2 + (P_i - p0)^2, and its
// adjoint dE/dx (hand-written; no automatic differentiation).
//
// The reference keeps the Voronoi vertices of every point relative to it
// (GPUVoroCur, DelaunayGPU.cu:921-922, DelaunayGPU.h:57-58) and has unused
// polygon-area helpers (functions.h:219-230); the energy and its gradient have
// no reference counterpart.  Vertex m of point i is the circumcentre of
// (x_i, x_{n_m}, x_{n_{m+1}}), computed relative to x_i from minimum-image
// displacements: h = c(a, b) with a = d(n_m), b = d(n_{m+1}),
//     c(a,b) = (|a|^2 b_y - |b|^2 a_y, |b|^2 a_x - |a|^2 b_x) / (2 a x b).
//
// FP64 path: everything in FP64.  FP32 path: displacements in FP64, rounded
// to FP32, vertices / areas / perimeters / slot gradients in FP32, sums of
// cell energies and the gradient output in FP64.
//
// Adjoint.  dE_i/dh_m = 2(A_i-a0) dA/dh_m + 2(P_i-p0) dP/dh_m with
//     dA/dh_m = ((h_{m+1}-h_{m-1})_y, (h_{m-1}-h_{m+1})_x) / 2,
//     dP/dh_m = (h_m-h_{m-1})/|h_m-h_{m-1}| - (h_{m+1}-h_m)/|h_{m+1}-h_m|
// is written per ring slot (k_voronoi_slots).  A vertex is shared by the three
// cells of its triangle (p, j, l): slot m of p, slot pos_j(p)-1 of j and slot
// pos_l(p) of l.  Point p gathers, for each of its vertices, the sum of the
// three slot gradients times dh/dx_p = I - dc/da - dc/db (k_voronoi_gather),
// so every gradient entry is written once, without atomics.
//
// Rows that are not a ring of a triangulation (count < 3 or > stride, or a
// vertex whose triangle is missing from a neighbour's ring) give NaN area,
// perimeter, energy and gradient; the kernels never synchronise with the host.
#include "gdt_kernels.h"

#include <type_traits>

#include "gdt_vor.h"

using namespace gdt_vor;


// Per cell: area, perimeter and energy term (vertices streamed, O(1) state).
template <int PREC>
__global__ __launch_bounds__(256) void k_voronoi_cells(const GdtSet* __restrict__ sets, int nsets,
                                                       int total_n, const double* __restrict__ xy,
                                                       const int* __restrict__ nbrs,
                                                       const int* __restrict__ counts, int stride,
                                                       double a0, double p0,
                                                       double* __restrict__ area,
                                                       double* __restrict__ perim,
                                                       double* __restrict__ ecell)
{
    using R = VReal<PREC>;
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= total_n) return;
    const GdtSet S = sets[gdt_set_of(sets, nsets, i)];
    const int k = counts[i];
    const int* row = nbrs + (size_t)i * stride;
    if (!vor_row_ok(row, k, stride, total_n)) {
        const double nan = __builtin_nan("");
        area[i] = nan;
        perim[i] = nan;
        ecell[i] = nan;
        return;
    }
    const V2<R> d0 = vor_round<R>(vor_disp(S, xy, i, row[0]));
    V2<R> a = d0, h0 = {R(0), R(0)}, hp = {R(0), R(0)};
    R A = R(0), P = R(0);
    for (int m = 0; m < k; ++m) {
        const V2<R> b = m + 1 < k ? vor_round<R>(vor_disp(S, xy, i, row[m + 1])) : d0;
        const V2<R> h = vor_circ<R>(a, b);
        if (m == 0) {
            h0 = h;
        } else {
            A += hp.x * h.y - h.x * hp.y;
            P += vor_len<R>(h.x - hp.x, h.y - hp.y);
        }
        hp = h;
        a = b;
    }
    A += hp.x * h0.y - h0.x * hp.y;
    P += vor_len<R>(h0.x - hp.x, h0.y - hp.y);
    const double Ad = 0.5 * (double)A, Pd = (double)P;
    area[i] = Ad;
    perim[i] = Pd;
    ecell[i] = (Ad - a0) * (Ad - a0) + (Pd - p0) * (Pd - p0);
}

// Deterministic per-set sums of v (fixed slices and tree order, no atomics).
__device__ inline double block_sum_f64(double acc, double* s)
{
    s[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    return s[0];
}

// Block (b, s): slice b of GDT_SUM_PARTS of set s -> partial[s * GDT_SUM_PARTS + b].
__global__ __launch_bounds__(256) void k_sum_slices(const GdtSet* __restrict__ sets,
                                                    const double* __restrict__ v,
                                                    double* __restrict__ partial)
{
    __shared__ double s[256];
    const GdtSet S = sets[blockIdx.y];
    const int per = (S.n + GDT_SUM_PARTS - 1) / GDT_SUM_PARTS;
    const int b = (int)blockIdx.x * per, e = min(b + per, S.n);
    double acc = 0.;
    for (int i = b + (int)threadIdx.x; i < e; i += (int)blockDim.x) acc += v[S.p0 + i];
    const double t = block_sum_f64(acc, s);
    if (threadIdx.x == 0) partial[(size_t)blockIdx.y * GDT_SUM_PARTS + blockIdx.x] = t;
}

// Block s: sum of the GDT_SUM_PARTS partials of set s.
__global__ __launch_bounds__(256) void k_sum_parts(const double* __restrict__ partial,
                                                   double* __restrict__ out)
{
    __shared__ double s[256];
    double acc = 0.;
    for (int i = (int)threadIdx.x; i < GDT_SUM_PARTS; i += (int)blockDim.x)
        acc += partial[(size_t)blockIdx.x * GDT_SUM_PARTS + i];
    const double t = block_sum_f64(acc, s);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// Per cell: slot gradients dE_i/dh_m, scaled by gscale[set] (1 if null).
template <int PREC>
__global__ __launch_bounds__(256) void k_voronoi_slots(const GdtSet* __restrict__ sets, int nsets,
                                                       int total_n, const double* __restrict__ xy,
                                                       const int* __restrict__ nbrs,
                                                       const int* __restrict__ counts, int stride,
                                                       double a0, double p0,
                                                       const double* __restrict__ area,
                                                       const double* __restrict__ perim,
                                                       const double* __restrict__ gscale,
                                                       double* __restrict__ G)
{
    using R = VReal<PREC>;
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= total_n) return;
    const int k = counts[i];
    const int* row = nbrs + (size_t)i * stride;
    double* Gi = G + 2 * (size_t)i * stride;
    if (!vor_row_ok(row, k, stride, total_n)) {
        for (int m = 0; m < 2 * stride; ++m) Gi[m] = __builtin_nan("");
        return;
    }
    const int si = gdt_set_of(sets, nsets, i);
    const GdtSet S = sets[si];
    const double g = gscale ? gscale[si] : 1.0;
    const R ca = (R)(2.0 * (area[i] - a0) * g), cp = (R)(2.0 * (perim[i] - p0) * g);
    const V2<R> d0 = vor_round<R>(vor_disp(S, xy, i, row[0]));
    const V2<R> d1 = vor_round<R>(vor_disp(S, xy, i, row[1]));
    const V2<R> dl = vor_round<R>(vor_disp(S, xy, i, row[k - 1]));
    V2<R> hprev = vor_circ<R>(dl, d0);           // vertex k-1
    const V2<R> h0 = vor_circ<R>(d0, d1);         // vertex 0
    V2<R> hcur = h0, b = d1;
    for (int m = 0; m < k; ++m) {
        V2<R> hnext;
        if (m + 1 < k) {
            const V2<R> c = m + 2 < k ? vor_round<R>(vor_disp(S, xy, i, row[m + 2])) : d0;
            hnext = vor_circ<R>(b, c);
            b = c;
        } else {
            hnext = h0;
        }
        const V2<R> gm = vor_slot_grad<R>(hprev, hcur, hnext, ca, cp);
        Gi[2 * m] = (double)gm.x;
        Gi[2 * m + 1] = (double)gm.y;
        hprev = hcur;
        hcur = hnext;
    }
}

// Position of p in the ring of j, or -1 (also for a row longer than stride).
__device__ inline int vor_pos(const int* __restrict__ nbrs, const int* __restrict__ counts,
                              int stride, int j, int p)
{
    const int k = counts[j];
    if (k < 1 || k > stride) return -1;
    const int* row = nbrs + (size_t)j * stride;
    for (int q = 0; q < k; ++q)
        if (row[q] == p) return q;
    return -1;
}

// Per point: dE/dx_p = sum over p's vertices of (sum of the three slot
// gradients) times (I - dc/da - dc/db).
template <int PREC>
__global__ __launch_bounds__(256) void k_voronoi_gather(const GdtSet* __restrict__ sets, int nsets,
                                                        int total_n, const double* __restrict__ xy,
                                                        const int* __restrict__ nbrs,
                                                        const int* __restrict__ counts, int stride,
                                                        const double* __restrict__ G,
                                                        double* __restrict__ grad)
{
    using R = VReal<PREC>;
    const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= total_n) return;
    const int k = counts[p];
    const int* row = nbrs + (size_t)p * stride;
    if (!vor_row_ok(row, k, stride, total_n)) {
        grad[2 * (size_t)p] = grad[2 * (size_t)p + 1] = __builtin_nan("");
        return;
    }
    const GdtSet S = sets[gdt_set_of(sets, nsets, p)];
    double gx = 0., gy = 0.;
    int bad = 0;
    for (int m = 0; m < k; ++m) {
        const int j = row[m], l = row[m + 1 < k ? m + 1 : 0];
        double sx = G[2 * ((size_t)p * stride + m)], sy = G[2 * ((size_t)p * stride + m) + 1];
        const int pj = vor_pos(nbrs, counts, stride, j, p);
        const int pl = vor_pos(nbrs, counts, stride, l, p);
        if (pj < 0 || pl < 0) {
            bad = 1;
            continue;
        }
        const int kj = counts[j];
        const size_t sj = (size_t)j * stride + (pj + kj - 1) % kj;
        const size_t sl = (size_t)l * stride + pl;
        sx += G[2 * sj] + G[2 * sl];
        sy += G[2 * sj + 1] + G[2 * sl + 1];
        const V2<R> a = vor_round<R>(vor_disp(S, xy, p, j));
        const V2<R> b = vor_round<R>(vor_disp(S, xy, p, l));
        const V2<R> c = vor_circ<R>(a, b);
        const R an = a.x * a.x + a.y * a.y, bn = b.x * b.x + b.y * b.y;
        const R i2D = R(0.5) / (a.x * b.y - a.y * b.x);
        // M[r][s] = dc_r/da_s + dc_r/db_s
        const R Mxx = i2D * (R(2) * a.x * b.y - R(2) * c.x * b.y - R(2) * b.x * a.y + R(2) * c.x * a.y);
        const R Myx = i2D * (bn - R(2) * a.x * b.x - R(2) * c.y * b.y + R(2) * b.x * a.x - an +
                             R(2) * c.y * a.y);
        const R Mxy = i2D * (R(2) * a.y * b.y - bn + R(2) * c.x * b.x + an - R(2) * b.y * a.y -
                             R(2) * c.x * a.x);
        const R Myy = i2D * (-R(2) * a.y * b.x + R(2) * c.y * b.x + R(2) * b.y * a.x - R(2) * c.y * a.x);
        gx += sx * (1.0 - (double)Mxx) - sy * (double)Myx;
        gy += -sx * (double)Mxy + sy * (1.0 - (double)Myy);
    }
    if (bad) gx = gy = __builtin_nan("");
    grad[2 * (size_t)p] = gx;
    grad[2 * (size_t)p + 1] = gy;
}

// Voronoi vertices relative to each point, one per ring slot:
// out[i*stride + m] = circumcentre of (x_i, x_{n_m}, x_{n_{m+1}}) - x_i
// (the layout of the reference's GPUVoroCur); invalid rows are NaN.
__global__ __launch_bounds__(256) void k_voronoi_vertices(const GdtSet* __restrict__ sets,
                                                          int nsets, int total_n,
                                                          const double* __restrict__ xy,
                                                          const int* __restrict__ nbrs,
                                                          const int* __restrict__ counts,
                                                          int stride, double* __restrict__ out)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= total_n) return;
    const int k = counts[i];
    const int* row = nbrs + (size_t)i * stride;
    double* o = out + 2 * (size_t)i * stride;
    if (!vor_row_ok(row, k, stride, total_n)) {
        for (int m = 0; m < 2 * stride; ++m) o[m] = __builtin_nan("");
        return;
    }
    const GdtSet S = sets[gdt_set_of(sets, nsets, i)];
    const V2<double> d0 = vor_round<double>(vor_disp(S, xy, i, row[0]));
    V2<double> a = d0;
    for (int m = 0; m < k; ++m) {
        const V2<double> b = m + 1 < k ? vor_round<double>(vor_disp(S, xy, i, row[m + 1])) : d0;
        const V2<double> h = vor_circ<double>(a, b);
        o[2 * m] = h.x;
        o[2 * m + 1] = h.y;
        a = b;
    }
}

hipError_t gdt_launch_voronoi_vertices(const GdtSet* d_sets, int nsets, int total_n,
                                       const double* xy, const int* nbrs, const int* counts,
                                       int stride, double* out, hipStream_t s)
{
    hipLaunchKernelGGL(k_voronoi_vertices, dim3((total_n + 255) / 256), dim3(256), 0, s, d_sets,
                       nsets, total_n, xy, nbrs, counts, stride, out);
    return hipGetLastError();
}

hipError_t gdt_launch_voronoi_forward(int precision, const GdtSet* d_sets, int nsets, int total_n,
                                      const double* xy, const int* nbrs, const int* counts,
                                      int stride, double a0, double p0, double* area, double* perim,
                                      double* ecell, double* partial, double* energy,
                                      hipStream_t s)
{
    const int grid = (total_n + 255) / 256;
    if (precision == 32)
        hipLaunchKernelGGL(k_voronoi_cells<32>, dim3(grid), dim3(256), 0, s, d_sets, nsets, total_n,
                           xy, nbrs, counts, stride, a0, p0, area, perim, ecell);
    else
        hipLaunchKernelGGL(k_voronoi_cells<64>, dim3(grid), dim3(256), 0, s, d_sets, nsets, total_n,
                           xy, nbrs, counts, stride, a0, p0, area, perim, ecell);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_sum_slices, dim3(GDT_SUM_PARTS, nsets), dim3(256), 0, s, d_sets, ecell,
                       partial);
    hipLaunchKernelGGL(k_sum_parts, dim3(nsets), dim3(256), 0, s, partial, energy);
    return hipGetLastError();
}

hipError_t gdt_launch_voronoi_adjoint(int precision, const GdtSet* d_sets, int nsets, int total_n,
                                      const double* xy, const int* nbrs, const int* counts,
                                      int stride, double a0, double p0, const double* area,
                                      const double* perim, const double* gscale, double* G,
                                      double* grad, hipStream_t s)
{
    const int grid = (total_n + 255) / 256;
    if (precision == 32) {
        hipLaunchKernelGGL(k_voronoi_slots<32>, dim3(grid), dim3(256), 0, s, d_sets, nsets, total_n,
                           xy, nbrs, counts, stride, a0, p0, area, perim, gscale, G);
        hipLaunchKernelGGL(k_voronoi_gather<32>, dim3(grid), dim3(256), 0, s, d_sets, nsets,
                           total_n, xy, nbrs, counts, stride, G, grad);
    } else {
        hipLaunchKernelGGL(k_voronoi_slots<64>, dim3(grid), dim3(256), 0, s, d_sets, nsets, total_n,
                           xy, nbrs, counts, stride, a0, p0, area, perim, gscale, G);
        hipLaunchKernelGGL(k_voronoi_gather<64>, dim3(grid), dim3(256), 0, s, d_sets, nsets,
                           total_n, xy, nbrs, counts, stride, G, grad);
    }
    return hipGetLastError();
}
