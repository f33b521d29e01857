// Checks the host/device mirror of GPUArray (include/gpudelaunay.hpp) the way
// the reference's callers use it (gpuarray.h:429-590, 983-1127, 1316-1373):
// the location state machine, 1-D and 2-D resizes of both copies (pitch
// width + 16 - width % 16), copies, memclear, mapped arrays and asynchronous
// handles.  Device data are read and written with hipMemcpy/hipMemset through
// the device pointers the handles return.  Prints one line per failed check;
// exit status 0 when all pass.
#include "gpudelaunay.hpp"

#include <cstdio>
#include <vector>

static int g_fail = 0;
#define CHECK(cond)                                                          \
    do {                                                                     \
        if (!(cond)) {                                                       \
            std::fprintf(stderr, "FAILED line %d: %s\n", __LINE__, #cond);   \
            ++g_fail;                                                        \
        }                                                                    \
    } while (0)

static void hip_ok(hipError_t e) { CHECK(e == hipSuccess); }

static std::vector<int> device_read(const int* d, size_t n)
{
    std::vector<int> v(n);
    hip_ok(hipMemcpy(v.data(), d, n * sizeof(int), hipMemcpyDeviceToHost));
    return v;
}

int main()
{
    try {
        // 1-D: host writes reach the device; device writes reach the host
        {
            GPUArray<int> a((unsigned int)1000);
            CHECK(a.getNumElements() == 1000u && a.getPitch() == 1000u && a.getHeight() == 1u);
            {
                ArrayHandle<int> h(a, access_location::host, access_mode::overwrite);
                for (int i = 0; i < 1000; ++i) h.data[i] = 3 * i;
                CHECK(a.isAcquired());
            }
            CHECK(!a.isAcquired());
            {
                ArrayHandle<int> d(a, access_location::device, access_mode::readwrite);
                const std::vector<int> v = device_read(d.data, 1000);
                CHECK(v[999] == 2997 && v[0] == 0);
                hip_ok(hipMemset(d.data, 0, 10 * sizeof(int)));   // the device copy is now the valid one
            }
            {
                ArrayHandle<int> h(a, access_location::host, access_mode::read);
                CHECK(h.data[5] == 0 && h.data[10] == 30);
            }
            // a copy keeps the data; resize keeps the prefix and zeroes the rest
            GPUArray<int> b(a);
            b.resize(1500);
            {
                ArrayHandle<int> h(b, access_location::host, access_mode::read);
                CHECK(h.data[10] == 30 && h.data[999] == 2997 && h.data[1000] == 0 && h.data[1499] == 0);
            }
            // resize while the device copy is the valid one
            {
                ArrayHandle<int> d(b, access_location::device, access_mode::readwrite);
                hip_ok(hipMemset(d.data + 20, 0, sizeof(int)));
            }
            b.resize(500);
            {
                ArrayHandle<int> h(b, access_location::host, access_mode::read);
                CHECK(b.getNumElements() == 500u && h.data[20] == 0 && h.data[21] == 63 &&
                      h.data[499] == 1497);
            }
            b.memclear(100);
            {
                ArrayHandle<int> h(b, access_location::host, access_mode::read);
                CHECK(h.data[99] == 297 && h.data[100] == 0 && h.data[499] == 0);
            }
        }
        // 2-D: the reference's pitch, rows and columns kept by resize on both sides
        {
            GPUArray<int> m((unsigned int)10, (unsigned int)3);
            CHECK(m.getPitch() == 16u && m.getHeight() == 3u && m.getNumElements() == 48u);
            GPUArray<int> m16((unsigned int)16, (unsigned int)2);
            CHECK(m16.getPitch() == 32u);     // a multiple of 16 still gains 16
            {
                ArrayHandle<int> h(m, access_location::host, access_mode::overwrite);
                for (int r = 0; r < 3; ++r)
                    for (int c = 0; c < 16; ++c) h.data[r * 16 + c] = 100 * r + c;
            }
            {
                ArrayHandle<int> d(m, access_location::device, access_mode::readwrite);
                (void)d;    // the device copy becomes the valid one
            }
            m.resize(20, 4);
            CHECK(m.getPitch() == 32u && m.getHeight() == 4u && m.getNumElements() == 128u);
            {
                ArrayHandle<int> h(m, access_location::host, access_mode::read);
                bool ok = true;
                for (int r = 0; r < 4; ++r)
                    for (int c = 0; c < 32; ++c) {
                        const int want = (r < 3 && c < 16) ? 100 * r + c : 0;
                        ok = ok && h.data[r * 32 + c] == want;
                    }
                CHECK(ok);
            }
            m.resize(5, 2);     // pitch 16: the first 16 columns of 2 rows
            {
                ArrayHandle<int> h(m, access_location::host, access_mode::read);
                CHECK(m.getPitch() == 16u && h.data[15] == 15 && h.data[16 + 7] == 107);
            }
        }
        // mapped: one allocation, no copies; the device pointer reads the host data
        {
            GPUArray<int> p((unsigned int)256, true);
            {
                ArrayHandle<int> h(p, access_location::host, access_mode::overwrite);
                for (int i = 0; i < 256; ++i) h.data[i] = i + 7;
            }
            {
                ArrayHandle<int> d(p, access_location::device, access_mode::read);
                const std::vector<int> v = device_read(d.data, 256);
                CHECK(v[0] == 7 && v[255] == 262);
            }
            p.resize(300);
            {
                ArrayHandle<int> h(p, access_location::host, access_mode::read);
                CHECK(h.data[255] == 262 && h.data[299] == 0);
            }
        }
        // asynchronous handle: valid after the caller synchronizes
        {
            GPUArray<int> q((unsigned int)4096);
            {
                ArrayHandle<int> d(q, access_location::device, access_mode::overwrite);
                hip_ok(hipMemset(d.data, 1, 4096 * sizeof(int)));
            }
            {
                ArrayHandleAsync<int> h(q, access_location::host, access_mode::read);
                hip_ok(hipDeviceSynchronize());
                CHECK(h.data[0] == 0x01010101 && h.data[4095] == 0x01010101);
            }
        }
        // one handle at a time
        {
            GPUArray<int> r((unsigned int)8);
            ArrayHandle<int> h(r, access_location::host, access_mode::read);
            bool threw = false;
            try {
                ArrayHandle<int> h2(r, access_location::device, access_mode::read);
            } catch (const std::logic_error&) {
                threw = true;
            }
            CHECK(threw);
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "exception: %s\n", e.what());
        return 2;
    }
    if (g_fail) return 1;
    std::printf("gpuarray_check: all passed\n");
    return 0;
}
