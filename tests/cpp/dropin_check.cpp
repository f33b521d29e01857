// Drives the C++ drop-in classes (include/gpudelaunay.hpp) the way a caller of
// the reference does (triangulation.cpp:224-299, 349-415): DelaunayGPU over
// GPUArray<double2>, triangulate, read the rings through ArrayHandle, move
// the points on the host, test-and-repair; cellListGPU binning; GPUVoroCur.
//
// Usage: dropin_check <points.bin> <n> <lx> <ly> <moved.bin> <precision> <out.bin> [cpu]
// A general box: <lx> = "box" and <ly> = "x11,x12,x21,x22" (setGeneral).
// With `cpu`, DelaunayGPU runs in CPU mode (setGPUcompute(false), 4 OpenMP
// threads) and the cell list is binned by cellListGPU::compute on the host.
// Either way the profiler (del.prof) must have timed both calls.
// out.bin: int32 MaxSize0, int32 rings0[n][MaxSize0], int32 counts0[n],
//          int32 MaxSize1, int32 rings1[n][MaxSize1], int32 counts1[n],
//          int32 lastRepaired, int32 xsize, int32 ysize, int32 Nmax,
//          uint32 cell_sizes[xsize*ysize], int32 idxs[xsize*ysize*Nmax],
//          double2 GPUVoroCur[n][MaxSize1]
#include "gpudelaunay.hpp"

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

static bool read_points(const char* path, int n, std::vector<double>& v)
{
    v.resize(2 * (size_t)n);
    FILE* f = std::fopen(path, "rb");
    if (!f) return false;
    const bool ok = std::fread(v.data(), sizeof(double), v.size(), f) == v.size();
    std::fclose(f);
    return ok;
}

template <class T>
static void dump(FILE* f, GPUArray<T>& a)
{
    ArrayHandle<T> h(a, access_location::host, access_mode::read);
    std::fwrite(h.data, sizeof(T), a.getNumElements(), f);
}

int main(int argc, char** argv)
{
    if (argc != 8 && argc != 9) { std::fprintf(stderr, "bad args\n"); return 2; }
    const bool cpu = argc == 9 && std::string(argv[8]) == "cpu";
    const int n = std::atoi(argv[2]);
    const bool general = std::string(argv[3]) == "box";
    double bm[4] = {0., 0., 0., 0.};
    if (general && std::sscanf(argv[4], "%lf,%lf,%lf,%lf", &bm[0], &bm[1], &bm[2], &bm[3]) != 4)
        return 2;
    const double lx = general ? 0. : std::atof(argv[3]), ly = general ? 0. : std::atof(argv[4]);
    const int prec = std::atoi(argv[6]);
    std::vector<double> xy, moved;
    if (!read_points(argv[1], n, xy) || !read_points(argv[5], n, moved)) return 3;
    try {
        GPUArray<double2> points((unsigned int)n);
        {
            ArrayHandle<double2> h(points, access_location::host, access_mode::overwrite);
            for (int i = 0; i < n; ++i) h.data[i] = make_double2(xy[2 * i], xy[2 * i + 1]);
        }
        PeriodicBoxPtr box = general
                                 ? std::make_shared<periodicBoundaries>(bm[0], bm[1], bm[2], bm[3])
                                 : std::make_shared<periodicBoundaries>(lx, ly);
        // a small guess on purpose: the call grows MaxSize and the arrays
        DelaunayGPU del(n, 4, 1.0, box);
        del.setPrecision(prec);
        if (cpu) {
            del.setGPUcompute(false);
            del.setOMPthreads(4);
        }
        GPUArray<int> tri, num;
        del.globalDelaunayTriangulation(points, tri, num);
        FILE* f = std::fopen(argv[7], "wb");
        if (!f) return 4;
        int ms = del.MaxSize;
        std::fwrite(&ms, 4, 1, f);
        dump(f, tri);
        dump(f, num);

        // move on the host (the reference driver moves the points through a
        // host handle as well), then test and repair
        {
            ArrayHandle<double2> h(points, access_location::host, access_mode::readwrite);
            for (int i = 0; i < n; ++i) h.data[i] = make_double2(moved[2 * i], moved[2 * i + 1]);
        }
        del.testAndRepairDelaunayTriangulation(points, tri, num);
        ms = del.MaxSize;
        std::fwrite(&ms, 4, 1, f);
        dump(f, tri);
        dump(f, num);
        std::fwrite(&del.lastRepaired, 4, 1, f);

        cellListGPU cl;
        cl.setBox(box);
        cl.setNp(n);
        cl.setGridSize(1.0);
        if (cpu) cl.compute(points);
        else cl.computeGPU(points);
        int xs = cl.getXsize(), ys = cl.getYsize(), nmax = cl.getNmax();
        std::fwrite(&xs, 4, 1, f);
        std::fwrite(&ys, 4, 1, f);
        std::fwrite(&nmax, 4, 1, f);
        dump(f, cl.cell_sizes);
        dump(f, cl.idxs);
        dump(f, del.GPUVoroCur);
        std::fclose(f);
        if (!(del.prof.timing("globalTriangulation") > 0.) || !(del.prof.timing("testAndRepair") > 0.)) {
            std::fprintf(stderr, "error: the profiler did not time the calls\n");
            return 5;
        }
        del.prof.print();
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
