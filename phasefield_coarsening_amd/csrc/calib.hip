// calib.hip -- libppfcal.so: plain streaming kernels of known traffic, used
// to calibrate measurements of the step kernels.
//
//   ppfcal_copy    : read n doubles, write n doubles (16 B per element)
//   ppfcal_stream3 : read three FP64 fields, write three (48 B per site), in
//                    flat memory order or in the single-step kernel's
//                    z-marching access pattern (a workgroup walks a column
//                    tile through a chunk of z planes), no stencil: the
//                    ceiling of that access pattern
//
// bench.py reports the copy rate as the achievable HBM ceiling next to the
// 8 TB/s peak, and the rocprofv3 --pmc traffic runs divide the step
// kernel's FETCH_SIZE / WRITE_SIZE by what the same counters read for this
// copy (whose true traffic is known).  No counterpart in the reference.
#include <hip/hip_runtime.h>

#include <stddef.h>

typedef double dv2 __attribute__((ext_vector_type(2)));

// One 16-byte element per thread, streaming (non-temporal) stores: the
// fastest plain copy form measured on MI355X.
__global__ __launch_bounds__(256) void k_copy_f64(const dv2 *__restrict__ in, dv2 *__restrict__ out,
                                                  size_t n2)
{
    const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (q < n2)
        __builtin_nontemporal_store(in[q], out + q);
}

// Copy `bytes` (rounded down to 16) from one device buffer to another,
// `iters` times; *ms gets the average kernel time from HIP events.
// Returns 0, or a negative value if a HIP call failed.
extern "C" int ppfcal_copy(int device, size_t bytes, int iters, double *ms)
{
    if (!ms || iters < 1 || bytes < 16)
        return -1;
    if (hipSetDevice(device) != hipSuccess)
        return -3;
    const size_t n2 = bytes / 16;
    dv2 *a = nullptr, *b = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipMalloc(&a, n2 * 16) != hipSuccess || hipMalloc(&b, n2 * 16) != hipSuccess ||
        hipMemset(a, 0, n2 * 16) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        rc = -2;
    } else {
        const unsigned blocks = (unsigned)((n2 + 255) / 256);
        hipLaunchKernelGGL(k_copy_f64, dim3(blocks), dim3(256), 0, 0, a, b, n2);
        hipEventRecord(e0, 0);
        for (int i = 0; i < iters; i++)
            hipLaunchKernelGGL(k_copy_f64, dim3(blocks), dim3(256), 0, 0, a, b, n2);
        hipEventRecord(e1, 0);
        if (hipEventSynchronize(e1) != hipSuccess) {
            rc = -3;
        } else {
            float f = 0.0f;
            hipEventElapsedTime(&f, e0, e1);
            *ms = (double)f / iters;
        }
    }
    if (e0)
        hipEventDestroy(e0);
    if (e1)
        hipEventDestroy(e1);
    if (a)
        hipFree(a);
    if (b)
        hipFree(b);
    return rc;
}

// 3-in/3-out FP64 stream, one site per thread in flat order (8 B lanes).
__global__ __launch_bounds__(256) void k_stream3_flat(const double *__restrict__ a,
                                                      const double *__restrict__ b,
                                                      const double *__restrict__ c,
                                                      double *__restrict__ o1,
                                                      double *__restrict__ o2,
                                                      double *__restrict__ o3, size_t n)
{
    const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (q < n) {
        const double x = a[q], y = b[q], z = c[q];
        __builtin_nontemporal_store(x + y, o1 + q);
        __builtin_nontemporal_store(y + z, o2 + q);
        __builtin_nontemporal_store(z + x, o3 + q);
    }
}

// The same stream in the z-marching pattern: a 512-thread workgroup owns a
// TX x TY column tile and walks tz planes, the next plane's three values
// loaded one plane ahead.  Tiles fastest in the workgroup order, as the
// single-step kernel's default.  nx % TX == 0, ny % TY == 0, nz % tz == 0.
template <int TX, int TY>
__global__ __launch_bounds__(512) void k_stream3_march(const double *__restrict__ a,
                                                       const double *__restrict__ b,
                                                       const double *__restrict__ c,
                                                       double *__restrict__ o1,
                                                       double *__restrict__ o2,
                                                       double *__restrict__ o3, int nx, int ny,
                                                       int tz)
{
    const int ntx = nx / TX, ntile = ntx * (ny / TY);
    const int tile = blockIdx.x % ntile, zc = blockIdx.x / ntile;
    const size_t plane = (size_t)nx * ny;
    const size_t col = (size_t)((tile / ntx) * TY + threadIdx.x / TX) * nx +
                       (tile % ntx) * TX + threadIdx.x % TX;
    const int z0 = zc * tz;
    double x = a[z0 * plane + col], y = b[z0 * plane + col], w = c[z0 * plane + col];
    for (int z = z0; z < z0 + tz; z++) {
        double xn = x, yn = y, wn = w;
        if (z + 1 < z0 + tz) {
            xn = a[(z + 1) * plane + col];
            yn = b[(z + 1) * plane + col];
            wn = c[(z + 1) * plane + col];
        }
        __builtin_nontemporal_store(x + y, o1 + z * plane + col);
        __builtin_nontemporal_store(y + w, o2 + z * plane + col);
        __builtin_nontemporal_store(w + x, o3 + z * plane + col);
        x = xn;
        y = yn;
        w = wn;
    }
}

// pattern 0: flat; 1: z march, 64x8 tiles; 2: 128x4; 3: 512x1.  tz planes
// per chunk (patterns 1-3).  *ms gets the average kernel time over `iters`
// launches from HIP events.  Returns 0, -1 for bad arguments (sizes that
// the pattern's tiles or tz do not divide), or -2/-3 if a HIP call failed.
extern "C" int ppfcal_stream3(int device, int nx, int ny, int nz, int pattern, int tz, int iters,
                              double *ms)
{
    static const int tx_of[4] = {0, 64, 128, 512}, ty_of[4] = {0, 8, 4, 1};
    if (!ms || iters < 1 || nx < 1 || ny < 1 || nz < 1 || pattern < 0 || pattern > 3)
        return -1;
    if (pattern > 0 && (tz < 1 || nx % tx_of[pattern] || ny % ty_of[pattern] || nz % tz))
        return -1;
    if (hipSetDevice(device) != hipSuccess)
        return -3;
    const size_t n = (size_t)nx * ny * nz;
    double *buf[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    for (int i = 0; i < 6 && rc == 0; i++)
        if (hipMalloc(&buf[i], n * sizeof(double)) != hipSuccess ||
            hipMemset(buf[i], 0, n * sizeof(double)) != hipSuccess)
            rc = -2;
    if (rc == 0 && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = -2;
    if (rc == 0) {
        const unsigned ntiles = pattern ? (unsigned)((nx / tx_of[pattern]) * (ny / ty_of[pattern])) : 0;
        const unsigned grid = pattern ? ntiles * (unsigned)(nz / tz) : (unsigned)((n + 255) / 256);
        auto launch = [&]() {
            if (pattern == 0)
                hipLaunchKernelGGL(k_stream3_flat, dim3(grid), dim3(256), 0, 0, buf[0], buf[1],
                                   buf[2], buf[3], buf[4], buf[5], n);
            else if (pattern == 1)
                hipLaunchKernelGGL((k_stream3_march<64, 8>), dim3(grid), dim3(512), 0, 0, buf[0],
                                   buf[1], buf[2], buf[3], buf[4], buf[5], nx, ny, tz);
            else if (pattern == 2)
                hipLaunchKernelGGL((k_stream3_march<128, 4>), dim3(grid), dim3(512), 0, 0, buf[0],
                                   buf[1], buf[2], buf[3], buf[4], buf[5], nx, ny, tz);
            else
                hipLaunchKernelGGL((k_stream3_march<512, 1>), dim3(grid), dim3(512), 0, 0, buf[0],
                                   buf[1], buf[2], buf[3], buf[4], buf[5], nx, ny, tz);
        };
        launch();
        hipEventRecord(e0, 0);
        for (int i = 0; i < iters; i++)
            launch();
        hipEventRecord(e1, 0);
        if (hipEventSynchronize(e1) != hipSuccess) {
            rc = -3;
        } else {
            float f = 0.0f;
            hipEventElapsedTime(&f, e0, e1);
            *ms = (double)f / iters;
        }
    }
    if (e0)
        hipEventDestroy(e0);
    if (e1)
        hipEventDestroy(e1);
    for (double *p : buf)
        if (p)
            hipFree(p);
    return rc;
}
