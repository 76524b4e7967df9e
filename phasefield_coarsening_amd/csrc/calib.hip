// calib.hip -- libppfcal.so: plain streaming kernels of known traffic, used
// to calibrate measurements of the step kernels.
//
//   ppfcal_copy : read n doubles, write n doubles (16 B per element)
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
