// ppf_reduce.h -- deterministic workgroup reductions (wave64 shuffles, then
// LDS across waves).  Every sum is taken in a fixed tree order, so a result
// does not change from run to run.
#pragma once

#include <hip/hip_runtime.h>

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v = fmin(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1)
        v = fmax(v, __shfl_xor(v, off, 64));
    return v;
}

// Sum three values over a workgroup of NT threads (NT a multiple of 64);
// thread 0 writes them to out[0..2].  Every thread must call it.
template <int NT>
__device__ __forceinline__ void block_sum3(const double v[3], double *out)
{
    static_assert(NT % 64 == 0 && NT <= 1024, "workgroup must be whole waves");
    constexpr int NW = NT / 64;
    __shared__ double part[3][NW];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double a[3];
    for (int i = 0; i < 3; i++)
        a[i] = wave_sum(v[i]);
    if (lane == 0)
        for (int i = 0; i < 3; i++)
            part[i][w] = a[i];
    __syncthreads();
    if (threadIdx.x == 0)
        for (int i = 0; i < 3; i++) {
            double t = 0.0;
            for (int j = 0; j < NW; j++)
                t += part[i][j];
            out[i] = t;
        }
}

// Error-free sum of two doubles (Knuth TwoSum): s + e == a + b exactly.
__device__ __forceinline__ void two_sum(double a, double b, double &s, double &e)
{
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Compensated (double-double) accumulator.
struct DD {
    double hi, lo;
};

__device__ __forceinline__ DD dd_add(DD a, DD b)
{
    double s, e;
    two_sum(a.hi, b.hi, s, e);
    e += a.lo + b.lo;
    DD r;
    two_sum(s, e, r.hi, r.lo);
    return r;
}

__device__ __forceinline__ DD dd_add1(DD a, double x)
{
    double s, e;
    two_sum(a.hi, x, s, e);
    e += a.lo;
    DD r;
    two_sum(s, e, r.hi, r.lo);
    return r;
}

__device__ __forceinline__ DD wave_dd_sum(DD v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        DD o;
        o.hi = __shfl_xor(v.hi, off, 64);
        o.lo = __shfl_xor(v.lo, off, 64);
        v = dd_add(v, o);
    }
    return v;
}
