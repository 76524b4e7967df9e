// k_ccl.h -- connected-component labelling of the solid phase {etaS > thr}
// with 6-connectivity (periodic or bounded), for particle counts, sizes and
// the mean particle radius.
//
// Lock-free union-find on one int32 label per site: every solid site starts
// as its own root (L[q] = q), each site unites with its +x, +y, +z solid
// neighbours (atomicMin links the larger root under the smaller, so labels
// only ever decrease and every parent chain ends), then every site is
// pointed at its root and root sizes are counted.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

__device__ __forceinline__ int ccl_load(const int *L, int x)
{
    return __hip_atomic_load(L + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving: each visited node is re-pointed at its
// grandparent (atomicMin, so a label only ever decreases and still names an
// ancestor), which keeps the chains short in large merged components.
__device__ __forceinline__ int ccl_find(int *L, int x)
{
    int y = ccl_load(L, x);
    while (y != x) {
        const int z = ccl_load(L, y);
        if (z < y)
            atomicMin(L + x, z);
        x = y;
        y = z;
    }
    return x;
}

__device__ __forceinline__ void ccl_union(int *L, int a, int b)
{
    for (;;) {
        a = ccl_find(L, a);
        b = ccl_find(L, b);
        if (a == b)
            return;
        if (a < b) {
            const int old = atomicMin(L + b, a);
            if (old == b)
                return;
            b = old;
        } else {
            const int old = atomicMin(L + a, b);
            if (old == a)
                return;
            a = old;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_ccl_init(const T *__restrict__ s, int *__restrict__ L,
                                                  int *__restrict__ cnt, size_t n, double thr)
{
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
        L[q] = (double)s[q] > thr ? (int)q : -1;
        cnt[q] = 0;
    }
}

template <int BC>
__global__ __launch_bounds__(256) void k_ccl_merge(int *__restrict__ L, int nx, int ny, int nz)
{
    const size_t n = (size_t)nx * ny * nz;
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
        if (ccl_load(L, (int)q) < 0)
            continue;
        const int x = (int)(q % nx), y = (int)((q / nx) % ny), z = (int)(q / ((size_t)nx * ny));
        int nb[3] = {-1, -1, -1};
        if (x + 1 < nx)
            nb[0] = (int)q + 1;
        else if (BC == 0 && nx > 1)
            nb[0] = (int)q - (nx - 1);
        if (y + 1 < ny)
            nb[1] = (int)q + nx;
        else if (BC == 0 && ny > 1)
            nb[1] = (int)q - (ny - 1) * nx;
        if (z + 1 < nz)
            nb[2] = (int)(q + (size_t)nx * ny);
        else if (BC == 0 && nz > 1)
            nb[2] = (int)(q - (size_t)(nz - 1) * nx * ny);
        for (int a = 0; a < 3; a++)
            if (nb[a] >= 0 && ccl_load(L, nb[a]) >= 0)
                ccl_union(L, (int)q, nb[a]);
    }
}

// Point every solid site at its root and count the sites of each root.
// Lanes of a wave with the same root add their count with one atomic (one
// large component would otherwise serialise millions of atomics on one
// counter).
static __global__ __launch_bounds__(256) void k_ccl_flatten(int *__restrict__ L, int *__restrict__ cnt,
                                                     size_t n)
{
    const size_t stride = (size_t)gridDim.x * 256;
    // every lane runs the same number of iterations so the ballots see the
    // whole wave
    const size_t iters = (n + stride - 1) / stride;
    for (size_t it = 0; it < iters; it++) {
        const size_t q = it * stride + (size_t)blockIdx.x * 256 + threadIdx.x;
        int r = -1;
        if (q < n && L[q] >= 0)
            r = ccl_find(L, (int)q);
        unsigned long long pending = __ballot(r >= 0);
        while (pending) {
            const int leader = __ffsll((unsigned long long)pending) - 1;
            const int rl = __shfl(r, leader, 64);
            const unsigned long long same = __ballot(r == rl);
            if ((int)(threadIdx.x & 63) == leader)
                atomicAdd(cnt + rl, (int)__popcll(same));
            pending &= ~same;
        }
    }
}

// Number of roots (first pass, list == nullptr), or append each root's size
// to list (second pass; *nroots restarts from zero, cap bounds the writes).
static __global__ __launch_bounds__(256) void k_ccl_collect(const int *__restrict__ L,
                                                     const int *__restrict__ cnt, size_t n,
                                                     unsigned long long *__restrict__ nroots,
                                                     int *__restrict__ list, unsigned long long cap)
{
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256) {
        if (L[q] != (int)q)
            continue;
        const unsigned long long i = atomicAdd(nroots, 1ull);
        if (list && i < cap)
            list[i] = cnt[q];
    }
}
