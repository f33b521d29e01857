// gpudelaunay.hpp -- C++ drop-in for the reference's host classes, over the
// C ABI of gpudelaunay.h (header only; link libgpudelaunay.so and amdhip64).
//
// Class surface kept from the reference so that callers compile unchanged:
//   Index2D                       indexer.h:32-62
//   periodicBoundaries, PeriodicBoxPtr   periodicBoundaries.h:25-73, 191
//   GPUArray<T>, ArrayHandle<T>,  gpuarray.h:26-62, 305-380, 429-590,
//   ArrayHandleAsync<T>           983-1127 (host/device mirror), 1316-1373 (1-D and
//                                 2-D resize), mapped arrays
//   cellListGPU                   cellListGPU.h:17-121
//   DelaunayGPU                   DelaunayGPU.h:18-58
//   multiProfiler                 multiProfiler.h:9-75
// Errors: a failing HIP call or library status throws std::runtime_error (the
// reference prints and throws std::exception, std_include.h:61-72).
// setGPUcompute(false) selects the library's CPU mode (gdt_host_*: the same
// per-point routines on the host over setOMPthreads OpenMP threads), which
// reads and writes the arrays through host handles, as the reference's CPU
// branches do.  Only rectangular boxes.
#pragma once

#include <hip/hip_runtime_api.h>
#include <hip/hip_vector_types.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <iostream>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "gpudelaunay.h"

namespace gdt_detail {

inline void hip_check(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

inline void status_check(int st, const char* what)
{
    if (st != GDT_OK) throw std::runtime_error(std::string(what) + ": " + gdt_status_string(st));
}

}  // namespace gdt_detail

// ---------------------------------------------------------------------------
//! Named wall-clock timers with the reference's interface (multiProfiler.h:9-75):
//! start/end bracket a region, timing() is its mean duration in seconds over
//! the completed calls, print() lists every region.
class multiProfiler {
public:
    using clock = std::chrono::high_resolution_clock;
    multiProfiler() = default;

    void addName(const std::string& name)
    {
        if (processMap.count(name)) return;
        processMap.emplace(name, numberOfProfiles);
        startTimes.push_back(clock::now());
        endTimes.push_back(clock::now());
        functionCalls.push_back(0);
        timeTaken.push_back(0.);
        ++numberOfProfiles;
    }
    void start(const std::string& name)
    {
        addName(name);
        startTimes[(size_t)processMap[name]] = clock::now();
    }
    void end(const std::string& name)
    {
        auto it = processMap.find(name);
        if (it == processMap.end()) return;
        const size_t k = (size_t)it->second;
        endTimes[k] = clock::now();
        timeTaken[k] += std::chrono::duration<double>(endTimes[k] - startTimes[k]).count();
        functionCalls[k] += 1;
    }
    double timing(const std::string& name) const
    {
        auto it = processMap.find(name);
        if (it == processMap.end() || functionCalls[(size_t)it->second] == 0) return 0.;
        return timeTaken[(size_t)it->second] / functionCalls[(size_t)it->second];
    }
    void print() const
    {
        for (const auto& kv : processMap)
            std::cout << "process " << kv.first << " averaged\t " << timing(kv.first) << " \t over "
                      << functionCalls[(size_t)kv.second] << " function calls" << std::endl;
    }

    int numberOfProfiles = 0;
    std::vector<clock::time_point> startTimes;
    std::vector<clock::time_point> endTimes;
    std::vector<int> functionCalls;
    std::vector<double> timeTaken;
    std::map<std::string, int> processMap;
};

// ---------------------------------------------------------------------------
//! Row-major 2D indexer: (i, j) -> j*w + i (indexer.h:32-62).
class Index2D {
public:
    explicit Index2D(unsigned int w = 0) : m_w(w), m_h(w) {}
    Index2D(unsigned int w, unsigned int h) : m_w(w), m_h(h) {}
    unsigned int operator()(unsigned int i, unsigned int j) const { return j * m_w + i; }
    unsigned int getNumElements() const { return m_w * m_h; }
    unsigned int getW() const { return m_w; }
    unsigned int getH() const { return m_h; }

private:
    unsigned int m_w, m_h;
};

// ---------------------------------------------------------------------------
//! Periodic box (periodicBoundaries.h:25-73): a rectangle (setSquare, 75-84)
//! or a general box (setGeneral, 86-101) whose matrix [[x11, x12], [x21, x22]]
//! has the lattice vectors as columns.
class periodicBoundaries {
public:
    periodicBoundaries() { setSquare(1.0, 1.0); }
    periodicBoundaries(double x, double y) { setSquare(x, y); }
    periodicBoundaries(double a, double b, double c, double d) { setGeneral(a, b, c, d); }
    void setSquare(double x, double y)
    {
        if (!(x > 0.) || !(y > 0.)) throw std::invalid_argument("box sides must be positive");
        x11 = x;
        x12 = x21 = 0.;
        x22 = y;
        invert();
        isSquare = true;
    }
    void setGeneral(double a, double b, double c, double d)
    {
        const double det = a * d - b * c;
        if (!std::isfinite(a) || !std::isfinite(b) || !std::isfinite(c) || !std::isfinite(d) ||
            !std::isfinite(det) || det == 0.)
            throw std::invalid_argument("the box matrix must be finite and invertible");
        x11 = a;
        x12 = b;
        x21 = c;
        x22 = d;
        invert();
        isSquare = false;
    }
    void getBoxDims(double& xx, double& xy, double& yx, double& yy) const
    {
        xx = x11;
        xy = x12;
        yx = x21;
        yy = x22;
    }
    void getBoxInvDims(double& xx, double& xy, double& yx, double& yy) const
    {
        xx = xi11;
        xy = xi12;
        yx = xi21;
        yy = xi22;
    }
    bool isBoxSquare() const { return isSquare; }
    //! whether the library runs this box on its rectangular path
    bool isRectangle() const { return x12 == 0. && x21 == 0. && x11 > 0. && x22 > 0.; }
    //! {x11, x12, x21, x22}, as the general-box calls take it
    void boxMatrix(double* m) const
    {
        m[0] = x11;
        m[1] = x12;
        m[2] = x21;
        m[3] = x22;
    }
    //! unit-cell coordinates to the box and back (periodicBoundaries.h:103-129)
    void Trans(const double2& p, double2& out) const
    {
        const double x = x11 * p.x + x12 * p.y, y = x21 * p.x + x22 * p.y;
        out.x = x;
        out.y = y;
    }
    void invTrans(const double2& p, double2& out) const
    {
        const double x = xi11 * p.x + xi12 * p.y, y = xi21 * p.x + xi22 * p.y;
        out.x = x;
        out.y = y;
    }
    //! nearest-image displacement p1 - p2 (periodicBoundaries.h:154-179): per
    //! axis for rectangles, on unit-cell coordinates wrapped into [-1/2, 1/2]
    //! for general boxes
    void minDist(const double2& p1, const double2& p2, double2& out) const
    {
        if (isSquare) {
            double dx = p1.x - p2.x, dy = p1.y - p2.y;
            dx -= x11 * std::nearbyint(dx / x11);
            dy -= x22 * std::nearbyint(dy / x22);
            out.x = dx;
            out.y = dy;
            return;
        }
        double2 a, b;
        invTrans(p1, a);
        invTrans(p2, b);
        double2 u;
        u.x = a.x - b.x;
        u.y = a.y - b.y;
        u.x -= std::nearbyint(u.x);
        u.y -= std::nearbyint(u.y);
        Trans(u, out);
    }
    //! wrap into the unit cell (periodicBoundaries.h:131-152)
    void putInBoxReal(double2& p) const
    {
        if (isSquare) {
            p.x -= std::floor(p.x / x11) * x11;
            p.y -= std::floor(p.y / x22) * x22;
            if (p.x >= x11) p.x -= x11;
            if (p.y >= x22) p.y -= x22;
            return;
        }
        double2 u;
        invTrans(p, u);
        u.x -= std::floor(u.x);
        u.y -= std::floor(u.y);
        if (u.x >= 1.) u.x -= 1.;
        if (u.y >= 1.) u.y -= 1.;
        Trans(u, p);
    }
    //! p + disp, put back in the box (periodicBoundaries.h:181-189)
    void move(double2& p, const double2& disp) const
    {
        p.x += disp.x;
        p.y += disp.y;
        putInBoxReal(p);
    }
    double x11 = 1., x12 = 0., x21 = 0., x22 = 1.;
    double xi11 = 1., xi12 = 0., xi21 = 0., xi22 = 1.;
    bool isSquare = true;

private:
    void invert()
    {
        const double det = x11 * x22 - x12 * x21;
        xi11 = x22 / det;
        xi12 = -x12 / det;
        xi21 = -x21 / det;
        xi22 = x11 / det;
    }
};
typedef std::shared_ptr<periodicBoundaries> PeriodicBoxPtr;

// ---------------------------------------------------------------------------
struct access_location { enum Enum { host, device }; };
struct access_mode { enum Enum { read, readwrite, overwrite }; };
struct data_location { enum Enum { host, device, hostdevice }; };

//! Host/device mirrored array (gpuarray.h:429-590, 983-1127).  Data live where
//! they were last written; an acquire copies only when the requested side is
//! stale and the mode is not overwrite.  One handle at a time (gpuarray.h:991).
//! 2-D arrays pad each row to the reference's pitch, width + 16 - width % 16
//! (gpuarray.h:673-677).  Mapped arrays live in pinned host memory that the
//! device addresses directly (hipHostMalloc with hipHostMallocMapped, not the
//! reference's registration of existing memory): acquires copy nothing, and a
//! synchronous host acquire of device data synchronizes the device
//! (gpuarray.h:936-940).  Asynchronous acquires (ArrayHandleAsync) copy on the
//! null stream.
template <class T>
class GPUArray {
public:
    GPUArray() = default;
    explicit GPUArray(unsigned int num_elements) { allocate(num_elements, num_elements, num_elements ? 1u : 0u); }
    GPUArray(unsigned int width, unsigned int height)
    {
        const unsigned int p = pitch_of(width);
        allocate(p * height, p, height);
    }
    GPUArray(unsigned int num_elements, bool mapped) : m_mapped(mapped)
    {
        allocate(num_elements, num_elements, num_elements ? 1u : 0u);
    }
    explicit GPUArray(bool mapped) : m_mapped(mapped) {}
    GPUArray(unsigned int width, unsigned int height, bool mapped) : m_mapped(mapped)
    {
        const unsigned int p = pitch_of(width);
        allocate(p * height, p, height);
    }
    GPUArray(const GPUArray& o) { copy_from(o); }
    GPUArray& operator=(const GPUArray& o)
    {
        if (this != &o) {
            deallocate();
            copy_from(o);
        }
        return *this;
    }
    GPUArray(GPUArray&& o) noexcept { swap(o); }
    GPUArray& operator=(GPUArray&& o) noexcept
    {
        swap(o);
        return *this;
    }
    ~GPUArray() { deallocate(); }

    unsigned int getNumElements() const { return m_n; }
    //! Row width in elements including padding (1-D: the element count).
    unsigned int getPitch() const { return m_pitch; }
    //! Rows (1-D: 1).
    unsigned int getHeight() const { return m_height; }
    bool isNull() const { return m_h == nullptr; }
    bool isAcquired() const { return m_acquired; }

    void swap(GPUArray& o) noexcept
    {
        std::swap(m_n, o.m_n);
        std::swap(m_pitch, o.m_pitch);
        std::swap(m_height, o.m_height);
        std::swap(m_h, o.m_h);
        std::swap(m_d, o.m_d);
        std::swap(m_loc, o.m_loc);
        std::swap(m_acquired, o.m_acquired);
        std::swap(m_mapped, o.m_mapped);
    }

    //! Resize, keeping the first min(old, new) elements of both copies; new
    //! elements are zero (gpuarray.h:1316-1335).
    void resize(unsigned int n)
    {
        if (m_acquired) throw std::logic_error("GPUArray::resize while acquired");
        if (n == m_n && m_height <= 1) return;
        GPUArray<T> t(n, m_mapped);
        copy_rows(t, std::min(n, m_n), 1, 0, 0);
        t.m_loc = m_loc;
        swap(t);
    }
    //! Resize a 2-D array, keeping the overlapping rows and columns of both
    //! copies; new elements are zero (gpuarray.h:1344-1373).
    void resize(unsigned int width, unsigned int height)
    {
        if (m_acquired) throw std::logic_error("GPUArray::resize while acquired");
        GPUArray<T> t(width, height, m_mapped);
        if (!isNull())
            copy_rows(t, std::min(m_pitch, t.m_pitch), std::min(m_height, height), m_pitch,
                      t.m_pitch);
        t.m_loc = isNull() ? t.m_loc : m_loc;
        swap(t);
    }

    //! Zero both copies from element `first` on (gpuarray.h:908-924).
    void memclear(unsigned int first = 0)
    {
        if (isNull() || first >= m_n) return;
        const size_t bytes = (size_t)(m_n - first) * sizeof(T);
        std::memset(m_h + first, 0, bytes);
        if (!m_mapped)
            gdt_detail::hip_check(hipMemset(m_d + first, 0, bytes), "GPUArray::memclear");
    }

    T* acquire(access_location::Enum where, access_mode::Enum mode, bool async = false)
    {
        if (m_acquired) throw std::logic_error("GPUArray already acquired");
        if (isNull()) return nullptr;
        m_acquired = true;
        if (where == access_location::host) {
            if (m_loc == data_location::device && mode != access_mode::overwrite) to_host(async);
            const bool both = mode == access_mode::read && m_loc != data_location::host;
            m_loc = both ? data_location::hostdevice : data_location::host;
            return m_h;
        }
        if (m_loc == data_location::host && mode != access_mode::overwrite) to_device(async);
        const bool both = mode == access_mode::read && m_loc != data_location::device;
        m_loc = both ? data_location::hostdevice : data_location::device;
        return m_d;
    }
    void release() { m_acquired = false; }

private:
    static unsigned int pitch_of(unsigned int width) { return width + (16u - (width & 15u)); }

    void allocate(unsigned int n, unsigned int pitch, unsigned int height)
    {
        m_n = n;
        m_pitch = pitch;
        m_height = height;
        m_loc = data_location::hostdevice;
        if (!n) return;
        const size_t bytes = (size_t)n * sizeof(T);
        if (m_mapped) {
            gdt_detail::hip_check(hipHostMalloc(reinterpret_cast<void**>(&m_h), bytes,
                                                hipHostMallocMapped),
                                  "GPUArray mapped allocation");
            gdt_detail::hip_check(hipHostGetDevicePointer(reinterpret_cast<void**>(&m_d), m_h, 0),
                                  "GPUArray mapped device pointer");
            std::memset(m_h, 0, bytes);
            return;
        }
        gdt_detail::hip_check(hipHostMalloc(reinterpret_cast<void**>(&m_h), bytes, 0),
                              "GPUArray host allocation");
        gdt_detail::hip_check(hipMalloc(reinterpret_cast<void**>(&m_d), bytes),
                              "GPUArray device allocation");
        std::memset(m_h, 0, bytes);
        gdt_detail::hip_check(hipMemset(m_d, 0, bytes), "GPUArray device clear");
    }
    void deallocate()
    {
        if (m_h) (void)hipHostFree(m_h);
        if (m_d && !m_mapped) (void)hipFree(m_d);
        m_h = nullptr;
        m_d = nullptr;
        m_n = m_pitch = m_height = 0;
        m_acquired = false;
        m_loc = data_location::hostdevice;
    }
    void to_host(bool async)
    {
        const size_t bytes = (size_t)m_n * sizeof(T);
        if (m_mapped) {
            if (!async) gdt_detail::hip_check(hipDeviceSynchronize(), "GPUArray mapped sync");
            return;
        }
        gdt_detail::hip_check(async ? hipMemcpyAsync(m_h, m_d, bytes, hipMemcpyDeviceToHost, nullptr)
                                    : hipMemcpy(m_h, m_d, bytes, hipMemcpyDeviceToHost),
                              "GPUArray device->host");
    }
    void to_device(bool async)
    {
        if (m_mapped) return;
        const size_t bytes = (size_t)m_n * sizeof(T);
        gdt_detail::hip_check(async ? hipMemcpyAsync(m_d, m_h, bytes, hipMemcpyHostToDevice, nullptr)
                                    : hipMemcpy(m_d, m_h, bytes, hipMemcpyHostToDevice),
                              "GPUArray host->device");
    }
    //! rows x cols elements of both copies into t (row pitches sp, tp; 0: one row)
    void copy_rows(GPUArray& t, unsigned int cols, unsigned int rows, unsigned int sp,
                   unsigned int tp) const
    {
        if (!cols || !rows || isNull() || t.isNull()) return;
        for (unsigned int r = 0; r < rows; ++r)
            std::memcpy(t.m_h + (size_t)r * tp, m_h + (size_t)r * sp, (size_t)cols * sizeof(T));
        if (m_mapped || t.m_mapped) return;
        gdt_detail::hip_check(hipMemcpy2D(t.m_d, (size_t)std::max(tp, cols) * sizeof(T), m_d,
                                          (size_t)std::max(sp, cols) * sizeof(T),
                                          (size_t)cols * sizeof(T), rows, hipMemcpyDeviceToDevice),
                              "GPUArray resize");
    }
    void copy_from(const GPUArray& o)
    {
        m_mapped = o.m_mapped;
        allocate(o.m_n, o.m_pitch, o.m_height);
        if (!m_n) return;
        const size_t bytes = (size_t)m_n * sizeof(T);
        if (o.m_loc != data_location::device || m_mapped) std::memcpy(m_h, o.m_h, bytes);
        if (o.m_loc != data_location::host && !m_mapped)
            gdt_detail::hip_check(hipMemcpy(m_d, o.m_d, bytes, hipMemcpyDeviceToDevice),
                                  "GPUArray copy");
        m_loc = o.m_loc;
    }

    unsigned int m_n = 0, m_pitch = 0, m_height = 0;
    T* m_h = nullptr;
    T* m_d = nullptr;
    data_location::Enum m_loc = data_location::hostdevice;
    bool m_acquired = false;
    bool m_mapped = false;
};

//! Scoped access to a GPUArray (gpuarray.h:305-328).
template <class T>
class ArrayHandle {
public:
    ArrayHandle(GPUArray<T>& a, access_location::Enum where = access_location::host,
                access_mode::Enum mode = access_mode::readwrite)
        : data(a.acquire(where, mode)), m_arr(a)
    {
    }
    ArrayHandle(const GPUArray<T>& a, access_location::Enum where = access_location::host,
                access_mode::Enum mode = access_mode::read)
        : ArrayHandle(const_cast<GPUArray<T>&>(a), where, mode)
    {
    }
    ~ArrayHandle() { m_arr.release(); }
    ArrayHandle(const ArrayHandle&) = delete;
    ArrayHandle& operator=(const ArrayHandle&) = delete;
    T* const data;

private:
    GPUArray<T>& m_arr;
};

//! The same with asynchronous copies on the null stream (gpuarray.h:356-380):
//! host data acquired this way are valid after the caller synchronizes.
template <class T>
class ArrayHandleAsync {
public:
    ArrayHandleAsync(GPUArray<T>& a, access_location::Enum where = access_location::host,
                     access_mode::Enum mode = access_mode::readwrite)
        : data(a.acquire(where, mode, true)), m_arr(a)
    {
    }
    ArrayHandleAsync(const GPUArray<T>& a, access_location::Enum where = access_location::host,
                     access_mode::Enum mode = access_mode::read)
        : ArrayHandleAsync(const_cast<GPUArray<T>&>(a), where, mode)
    {
    }
    ~ArrayHandleAsync() { m_arr.release(); }
    ArrayHandleAsync(const ArrayHandleAsync&) = delete;
    ArrayHandleAsync& operator=(const ArrayHandleAsync&) = delete;
    T* const data;

private:
    GPUArray<T>& m_arr;
};

// ---------------------------------------------------------------------------
//! Owns one library context on the current HIP device.
class gdtContext {
public:
    gdtContext()
    {
        int dev = 0;
        gdt_detail::hip_check(hipGetDevice(&dev), "hipGetDevice");
        gdt_detail::status_check(gdt_create(&m_c, dev), "gdt_create");
    }
    ~gdtContext() { gdt_destroy(m_c); }
    gdtContext(const gdtContext&) = delete;
    gdtContext& operator=(const gdtContext&) = delete;
    gdt_context* get() const { return m_c; }

private:
    gdt_context* m_c = nullptr;
};

// ---------------------------------------------------------------------------
//! Cell list (cellListGPU.h:17-121): a counting sort inside the library,
//! presented as the reference's cell_sizes / padded idxs on request.
class cellListGPU {
public:
    cellListGPU() : Box(std::make_shared<periodicBoundaries>()), m_ctx(std::make_shared<gdtContext>()) {}
    explicit cellListGPU(std::shared_ptr<gdtContext> ctx)
        : Box(std::make_shared<periodicBoundaries>()), m_ctx(std::move(ctx))
    {
    }

    void setNp(int nn) { Np = nn; }
    void setBox(periodicBoundaries& bx) { Box = std::make_shared<periodicBoundaries>(bx); }
    void setBox(PeriodicBoxPtr bx) { Box = std::move(bx); }

    //! cellListGPU::setGridSize (cellListGPU.cpp:97-117): cells per axis
    //! floor(L/a) rounded up to even; the y side uses Ly/ysize.  A general
    //! box's grid (cells of its reduced lattice basis) is known after compute.
    void setGridSize(double a)
    {
        m_a = a;
        if (!Box->isRectangle()) {
            xsize = ysize = totalCells = 0;
            boxsize = 0.;
            return;
        }
        xsize = (int)std::floor(Box->x11 / a);
        if (xsize % 2) ++xsize;
        ysize = (int)std::floor(Box->x22 / a);
        if (ysize % 2) ++ysize;
        xsize = std::max(xsize, 2);
        ysize = std::max(ysize, 2);
        boxsize = Box->x11 / xsize;
        totalCells = xsize * ysize;
        cell_indexer = Index2D(xsize, ysize);
    }
    int getNmax() { return Nmax; }
    int getXsize() { return xsize; }
    int getYsize() { return ysize; }
    double getBoxsize() { return boxsize; }

    //! rectangular boxes (the reference's cell list)
    int positionToCellIndex(double x, double y)
    {
        if (!Box->isRectangle())
            throw std::invalid_argument("positionToCellIndex: a general box's cells follow its "
                                        "reduced lattice basis");
        const double bsy = Box->x22 / ysize;
        const int cx = std::min(xsize - 1, std::max(0, (int)std::floor(x / boxsize)));
        const int cy = std::min(ysize - 1, std::max(0, (int)std::floor(y / bsy)));
        return cell_indexer(cx, cy);
    }
    //! all cells within `width` (the square around the cell)
    void getCellNeighbors(int cellIndex, int width, std::vector<int>& out)
    {
        const int w = std::min(width, xsize / 2);
        const int cx = cellIndex % xsize, cy = cellIndex / xsize;
        out.clear();
        for (int i = -w; i <= w; ++i)
            for (int j = -w; j <= w; ++j)
                out.push_back(cell_indexer(wrap(cx + j, xsize), wrap(cy + i, ysize)));
    }
    //! cells on the square shell at distance `width`
    void getCellShellNeighbors(int cellIndex, int width, std::vector<int>& out)
    {
        const int w = std::min(width, xsize);
        const int cx = cellIndex % xsize, cy = cellIndex / xsize;
        out.clear();
        for (int i = -w; i <= w; ++i)
            for (int j = -w; j <= w; ++j)
                if (i == -w || i == w || j == -w || j == w)
                    out.push_back(cell_indexer(wrap(cx + j, xsize), wrap(cy + i, ysize)));
    }

    //! Bin the points on the device; fills cell_sizes and idxs (reference layout).
    void computeGPU(GPUArray<double2>& points)
    {
        const int n = (int)points.getNumElements();
        if (xsize == 0) setGridSize(m_a);
        {
            ArrayHandle<double2> p(points, access_location::device, access_mode::read);
            const double* xy = reinterpret_cast<const double*>(p.data);
            double bm[4];
            Box->boxMatrix(bm);
            gdt_detail::status_check(
                Box->isRectangle()
                    ? gdt_celllist_build(m_ctx->get(), xy, n, Box->x11, Box->x22, m_a, nullptr)
                    : gdt_celllist_build_general(m_ctx->get(), xy, n, bm, m_a, nullptr),
                "cellListGPU::computeGPU");
        }
        Np = n;
        int xs = 0, ys = 0, mo = 0;
        double bx = 0., by = 0.;
        gdt_detail::status_check(gdt_celllist_info(m_ctx->get(), &xs, &ys, &bx, &by, &mo, nullptr),
                                 "gdt_celllist_info");
        xsize = xs;
        ysize = ys;
        boxsize = bx;
        totalCells = xs * ys;
        Nmax = std::max(Nmax, mo);
        cell_indexer = Index2D(xsize, ysize);
        cell_list_indexer = Index2D(Nmax, totalCells);
        if ((int)cell_sizes.getNumElements() != totalCells) cell_sizes.resize(totalCells);
        if ((int)idxs.getNumElements() != totalCells * Nmax) idxs.resize(totalCells * Nmax);
        ArrayHandle<unsigned int> cs(cell_sizes, access_location::device, access_mode::overwrite);
        ArrayHandle<int> ix(idxs, access_location::device, access_mode::overwrite);
        gdt_detail::status_check(gdt_celllist_export(m_ctx->get(), cs.data, ix.data, Nmax, nullptr),
                                 "gdt_celllist_export");
        gdt_detail::hip_check(hipDeviceSynchronize(), "cellListGPU::computeGPU");
    }
    //! Bin the points on the host (cellListGPU::compute, cellListGPU.cpp:289-330):
    //! the same buckets in the same order as computeGPU.
    void compute(GPUArray<double2>& points)
    {
        const int n = (int)points.getNumElements();
        if (xsize == 0) setGridSize(m_a);
        ArrayHandle<double2> p(points, access_location::host, access_mode::read);
        const double* xy = reinterpret_cast<const double*>(p.data);
        int xs = 0, ys = 0, mo = 0;
        double bm[4];
        Box->boxMatrix(bm);
        gdt_detail::status_check(gdt_host_celllist_general(xy, n, bm, m_a, nullptr, nullptr, 0, &xs,
                                                           &ys, &mo),
                                 "cellListGPU::compute");
        Np = n;
        xsize = xs;
        ysize = ys;
        boxsize = Box->isRectangle() ? Box->x11 / xs : 0.;
        totalCells = xs * ys;
        Nmax = std::max(Nmax, mo);
        cell_indexer = Index2D(xsize, ysize);
        cell_list_indexer = Index2D(Nmax, totalCells);
        if ((int)cell_sizes.getNumElements() != totalCells) cell_sizes.resize(totalCells);
        if ((int)idxs.getNumElements() != totalCells * Nmax) idxs.resize(totalCells * Nmax);
        ArrayHandle<unsigned int> cs(cell_sizes, access_location::host, access_mode::overwrite);
        ArrayHandle<int> ix(idxs, access_location::host, access_mode::overwrite);
        gdt_detail::status_check(gdt_host_celllist_general(xy, n, bm, m_a, cs.data, ix.data, Nmax,
                                                           &xs, &ys, &mo),
                                 "cellListGPU::compute");
    }
    void computeGPU() { computeGPU(particles); }
    void compute() { compute(particles); }

    const GPUArray<unsigned int>& getCellSizeArray() const { return cell_sizes; }
    const GPUArray<int>& getIdxArray() const { return idxs; }

    Index2D cell_indexer;
    Index2D cell_list_indexer;
    GPUArray<double2> particles;
    GPUArray<unsigned int> cell_sizes;
    GPUArray<int> idxs;
    PeriodicBoxPtr Box;

protected:
    static int wrap(int v, int n) { return ((v % n) + n) % n; }
    std::shared_ptr<gdtContext> m_ctx;
    int Np = 0, xsize = 0, ysize = 0, totalCells = 0, Nmax = 0;
    double boxsize = 0., m_a = 1.;
};

// ---------------------------------------------------------------------------
//! Periodic Delaunay triangulation (DelaunayGPU.h:18-58).
class DelaunayGPU {
public:
    DelaunayGPU() : m_ctx(std::make_shared<gdtContext>()), cList(m_ctx) {}
    DelaunayGPU(int N, int maximumNeighborsGuess, double cellSize, PeriodicBoxPtr bx)
        : DelaunayGPU()
    {
        initialize(N, maximumNeighborsGuess, cellSize, std::move(bx));
    }

    void initialize(int N, int maximumNeighborsGuess, double cellSize, PeriodicBoxPtr bx)
    {
        prof.start("initialization");
        Ncells = N;
        MaxSize = std::max(4, maximumNeighborsGuess);
        cellsize = cellSize;
        setBox(std::move(bx));
        cList.setNp(N);
        cList.setGridSize(cellsize);
        prof.end("initialization");
    }
    void resize(const int nmax) { MaxSize = nmax; }
    void setCellListSize(double csize)
    {
        cellsize = csize;
        cList.setGridSize(cellsize);
    }
    void updateList(GPUArray<double2>& points)
    {
        if (GPUcompute) cList.computeGPU(points);
        else cList.compute(points);
    }
    void setBox(PeriodicBoxPtr bx)
    {
        Box = bx ? std::move(bx) : std::make_shared<periodicBoundaries>();
        cList.setBox(Box);
    }
    void setSafetyMode(bool sm) { safetyMode = sm; }
    //! false: CPU mode (the library's per-point routines on the host, OpenMP)
    void setGPUcompute(bool flag) { GPUcompute = flag; }
    void setOMPthreads(unsigned int num) { OMPThreadsNum = num; }
    //! 64: all-FP64 predicates; 32: FP32 with a certified bound + FP64 fallback
    void setPrecision(int p)
    {
        if (p != GDT_PREC_FP64 && p != GDT_PREC_FP32F && p != GDT_PREC_EXACT)
            throw std::invalid_argument("precision");
        precision = p;
    }
    //! GPUVoroCur is filled after every call while this is on (default).
    void setVoronoiVertices(bool on) { voronoiVertices = on; }

    //! DelaunayGPU::globalDelaunayTriangulation (DelaunayGPU.cpp:118-157):
    //! grows MaxSize and GPUTriangulation until every ring fits.
    void globalDelaunayTriangulation(GPUArray<double2>& points, GPUArray<int>& GPUTriangulation,
                                     GPUArray<int>& cellNeighborNum)
    {
        const int n = (int)points.getNumElements();
        if (n == 0) return;
        Ncells = n;
        if ((int)cellNeighborNum.getNumElements() != n) cellNeighborNum.resize(n);
        prof.start("globalTriangulation");
        const access_location::Enum where = GPUcompute ? access_location::device : access_location::host;
        for (;;) {
            if ((int)GPUTriangulation.getNumElements() != n * MaxSize)
                GPUTriangulation.resize(n * MaxSize);
            int req = 0, st;
            {
                ArrayHandle<double2> p(points, where, access_mode::read);
                ArrayHandle<int> t(GPUTriangulation, where, access_mode::overwrite);
                ArrayHandle<int> c(cellNeighborNum, where, access_mode::overwrite);
                const double* xy = reinterpret_cast<const double*>(p.data);
                double bm[4];
                Box->boxMatrix(bm);
                const int off[2] = {0, n};
                if (!GPUcompute)
                    st = gdt_host_triangulate_general(xy, n, bm, cellsize, MaxSize, t.data, c.data,
                                                      precision, (int)OMPThreadsNum, &req);
                else if (Box->isRectangle())
                    st = gdt_triangulate(m_ctx->get(), xy, n, Box->x11, Box->x22, cellsize, MaxSize,
                                         t.data, c.data, precision, &req, nullptr);
                else
                    st = gdt_triangulate_general(m_ctx->get(), xy, 1, off, bm, cellsize, MaxSize,
                                                 t.data, c.data, precision, &req, nullptr);
            }
            if (st == GDT_OVERFLOW) {
                MaxSize = req;
                continue;
            }
            prof.end("globalTriangulation");
            gdt_detail::status_check(st, "DelaunayGPU::globalDelaunayTriangulation");
            break;
        }
        fill_voronoi(points, GPUTriangulation, cellNeighborNum);
    }

    //! DelaunayGPU::testAndRepairDelaunayTriangulation (DelaunayGPU.cpp:567-602).
    void testAndRepairDelaunayTriangulation(GPUArray<double2>& points,
                                            GPUArray<int>& GPUTriangulation,
                                            GPUArray<int>& cellNeighborNum)
    {
        const int n = (int)points.getNumElements();
        if ((int)GPUTriangulation.getNumElements() != n * MaxSize)
            throw std::invalid_argument("GPUTriangulation must hold N*MaxSize entries");
        int nf = 0, req = 0, st;
        prof.start("testAndRepair");
        {
            const access_location::Enum where =
                GPUcompute ? access_location::device : access_location::host;
            ArrayHandle<double2> p(points, where, access_mode::read);
            ArrayHandle<int> t(GPUTriangulation, where, access_mode::readwrite);
            ArrayHandle<int> c(cellNeighborNum, where, access_mode::readwrite);
            const double* xy = reinterpret_cast<const double*>(p.data);
            double bm[4];
            Box->boxMatrix(bm);
            const int off[2] = {0, n};
            if (!GPUcompute)
                st = gdt_host_test_and_repair_general(xy, n, bm, cellsize, MaxSize, t.data, c.data,
                                                      precision, (int)OMPThreadsNum, &nf, &req);
            else if (Box->isRectangle())
                st = gdt_test_and_repair(m_ctx->get(), xy, n, Box->x11, Box->x22, cellsize,
                                         MaxSize, t.data, c.data, precision, &nf, &req, nullptr);
            else
                st = gdt_test_and_repair_general(m_ctx->get(), xy, 1, off, bm, cellsize, MaxSize,
                                                 t.data, c.data, precision, &nf, &req, nullptr);
        }
        prof.end("testAndRepair");
        lastRepaired = nf;
        if (st == GDT_OVERFLOW) {
            MaxSize = req;
            globalDelaunayTriangulation(points, GPUTriangulation, cellNeighborNum);
            return;
        }
        gdt_detail::status_check(st, "DelaunayGPU::testAndRepairDelaunayTriangulation");
        fill_voronoi(points, GPUTriangulation, cellNeighborNum);
    }

    gdt_stats stats() const
    {
        gdt_stats s;
        gdt_detail::status_check(gdt_get_stats(m_ctx->get(), &s), "gdt_get_stats");
        return s;
    }

    //! wall-clock timers of the calls (initialization, globalTriangulation, testAndRepair)
    multiProfiler prof;
    PeriodicBoxPtr Box;
    int MaxSize = 4;
    //! Voronoi vertex m of point i relative to it at [i*MaxSize + m]
    GPUArray<double2> GPUVoroCur;
    int lastRepaired = 0;

protected:
    void fill_voronoi(GPUArray<double2>& points, GPUArray<int>& tri, GPUArray<int>& num)
    {
        if (!voronoiVertices) return;
        const int n = (int)points.getNumElements();
        if ((int)GPUVoroCur.getNumElements() != n * MaxSize) GPUVoroCur.resize(n * MaxSize);
        const access_location::Enum where = GPUcompute ? access_location::device : access_location::host;
        ArrayHandle<double2> p(points, where, access_mode::read);
        ArrayHandle<int> t(tri, where, access_mode::read);
        ArrayHandle<int> c(num, where, access_mode::read);
        ArrayHandle<double2> v(GPUVoroCur, where, access_mode::overwrite);
        const double* xy = reinterpret_cast<const double*>(p.data);
        double* out = reinterpret_cast<double*>(v.data);
        double bm[4];
        Box->boxMatrix(bm);
        int st;
        if (!GPUcompute)
            st = gdt_host_voronoi_vertices_general(xy, n, bm, t.data, c.data, MaxSize, out,
                                                   (int)OMPThreadsNum);
        else if (Box->isRectangle())
            st = gdt_voronoi_vertices(m_ctx->get(), xy, n, Box->x11, Box->x22, t.data, c.data,
                                      MaxSize, out, nullptr);
        else
            st = gdt_voronoi_vertices_general(m_ctx->get(), xy, n, bm, t.data, c.data, MaxSize,
                                              out, nullptr);
        gdt_detail::status_check(st, "GPUVoroCur");
    }

    std::shared_ptr<gdtContext> m_ctx;
    cellListGPU cList;
    int Ncells = 0;
    double cellsize = 1.10;     // reference default (DelaunayGPU.cpp:7)
    bool safetyMode = false;
    bool GPUcompute = true;
    unsigned int OMPThreadsNum = 1;
    int precision = GDT_PREC_FP64;
    bool voronoiVertices = true;
};
