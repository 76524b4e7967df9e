// k_ens.h -- batched ensembles: many independent grids of uneven size, one
// persistent launch per step.
//
// The work list holds one item per (domain, column tile, z chunk).  A fixed
// grid of workgroups (a few per CU) pulls items from an atomic counter until
// the list is exhausted, so small and large domains share the device without
// a launch per domain.  Each item carries its domain's tile shape (chosen by
// the ensemble autotuner per grid size); all shapes have 256 threads and run
// the same per-site arithmetic as the single-grid kernel (tile_march), so an
// ensemble member evolves bit for bit as it would alone.
#pragma once

#include "k_misc.h"
#include "k_step.h"

// Per-domain table entry on the device.
struct EnsDom {
    void *f[2][3];  // [copy][field]: etaS, etaL, c
    int nx, ny, nz, pad;
};

// Tile shapes of the ensemble kernel (all 256 threads).
enum { ENS_64x4 = 0, ENS_32x8 = 1, ENS_16x16 = 2, ENS_NSHAPES = 3 };

struct EnsItem {
    int dom;
    int shape;
    int x0, y0, z0, zend;
};

// LDS for the largest shape: 2 buffers x 4 fields x (TY+2) x (TX+2), max over
// 64x4 (6 x 66 = 396), 32x8 (10 x 34 = 340), 16x16 (18 x 18 = 324).
constexpr int ENS_LDS_SITES = 2 * 4 * 396;

template <class P, int BC, bool CONLY>
__global__ __launch_bounds__(256) void k_ensemble(const EnsDom *__restrict__ doms,
                                                  const EnsItem *__restrict__ items, int nitems,
                                                  int *__restrict__ counter, int parity,
                                                  typename P::K k)
{
    using T = typename P::T;
    __shared__ T sh[ENS_LDS_SITES];
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0)
            next = atomicAdd(counter, 1);
        __syncthreads();
        const int it = next;
        __syncthreads();  // every thread has read `next` (and left the last item)
        if (it >= nitems)
            return;
        const EnsItem w = items[it];
        // the member's pointers read straight from the descriptor (a
        // private copy indexed by the runtime parity would live in scratch)
        const EnsDom *d = doms + w.dom;
        const T *si = (const T *)d->f[parity][0];
        const T *li = (const T *)d->f[parity][1];
        const T *ci = (const T *)d->f[parity][2];
        T *so = (T *)d->f[1 - parity][0];
        T *lo = (T *)d->f[1 - parity][1];
        T *co = (T *)d->f[1 - parity][2];
        const int dnx = d->nx, dny = d->ny, dnz = d->nz;
        double acc[3];
        if (w.shape == ENS_64x4)
            tile_march<P, 64, 4, BC, CONLY, false, false>(
                si, li, ci, so, lo, co, dnx, dny, dnz, w.x0, w.y0, w.z0, w.zend, k,
                reinterpret_cast<T(*)[4][6][66]>(sh), acc);
        else if (w.shape == ENS_32x8)
            tile_march<P, 32, 8, BC, CONLY, false, false>(
                si, li, ci, so, lo, co, dnx, dny, dnz, w.x0, w.y0, w.z0, w.zend, k,
                reinterpret_cast<T(*)[4][10][34]>(sh), acc);
        else
            tile_march<P, 16, 16, BC, CONLY, false, false>(
                si, li, ci, so, lo, co, dnx, dny, dnz, w.x0, w.y0, w.z0, w.zend, k,
                reinterpret_cast<T(*)[4][18][18]>(sh), acc);
    }
}

// Per-domain double-double sums of c: blockIdx.y is the domain, each of the
// gridDim.x workgroups of a domain writes one (hi, lo) pair.
template <typename T>
__global__ __launch_bounds__(256) void k_ens_sum(const EnsDom *__restrict__ doms, int parity,
                                                 double *__restrict__ partials)
{
    const EnsDom d = doms[blockIdx.y];
    const T *c = (const T *)d.f[parity][2];
    const size_t n = (size_t)d.nx * d.ny * d.nz;
    DD acc = {0.0, 0.0};
    for (size_t q = (size_t)blockIdx.x * 256 + threadIdx.x; q < n; q += (size_t)gridDim.x * 256)
        acc = dd_add1(acc, (double)c[q]);
    acc = wave_dd_sum(acc);
    __shared__ double part[2][4];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) {
        part[0][w] = acc.hi;
        part[1][w] = acc.lo;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        DD t = {part[0][0], part[1][0]};
        for (int j = 1; j < 4; j++)
            t = dd_add(t, DD{part[0][j], part[1][j]});
        double *o = partials + 2 * ((size_t)blockIdx.y * gridDim.x + blockIdx.x);
        o[0] = t.hi;
        o[1] = t.lo;
    }
}

// One workgroup per domain: out[2*dom] = (target - sum)/n, out[2*dom+1] = sum.
static __global__ __launch_bounds__(256) void k_ens_delta(const EnsDom *__restrict__ doms,
                                                   const double *__restrict__ partials, int nb,
                                                   const double *__restrict__ targets,
                                                   double *__restrict__ out)
{
    const int dom = blockIdx.x;
    const EnsDom d = doms[dom];
    k_mass_delta_body(partials + 2 * (size_t)dom * nb, nb, targets[dom],
                      (double)d.nx * d.ny * d.nz, out + 2 * (size_t)dom);
}

// Stochastically rounded shift of each domain's c by its delta
// (blockIdx.y = domain); the per-site rule is k_mass_shift's.
static __global__ __launch_bounds__(256) void k_ens_shift(const EnsDom *__restrict__ doms, int parity,
                                                   const double *__restrict__ deltas,
                                                   uint64_t seed)
{
    const EnsDom d = doms[blockIdx.y];
    float *c = (float *)d.f[parity][2];
    const size_t n = (size_t)d.nx * d.ny * d.nz;
    mass_shift_body(c, n, deltas[2 * (size_t)blockIdx.y], seed, (size_t)blockIdx.x * 256,
                    (size_t)gridDim.x * 256);
}

// ---------------------------------------------------------------------------
// Two steps per launch (periodic or, with BC = 1, no-flux members): the same persistent work list
// walk with tile_march2 as the item body, so each member advances two steps
// bit for bit as two single steps.  tile_march2 needs its two rings within
// the workgroup (4 TX + 4 TY + 4 <= 256), so the shapes are 32x8 and 16x16.
// ---------------------------------------------------------------------------
enum { ENS2_32x8 = 0, ENS2_16x16 = 1, ENS2_NSHAPES = 2 };

// LDS in values: A1 (4 x (TY+4) x (TX+4)) then A2 (4 x (TY+2) x (TX+2)), the
// larger of the two shapes: 32x8 needs 1728 + 1360, 16x16 1600 + 1296.  FP32
// members double-buffer both (one barrier per plane; 24.7 KB per workgroup);
// FP64 keeps single buffers so that four workgroups fit a CU.
constexpr int ENS2_LDS = 4 * 12 * 36 + 4 * 10 * 34;

// Register target: FP32 members five waves per SIMD (the single-grid
// kernel's budget, no spills); FP64 two (~200 VGPRs with both member shapes
// in one kernel; three waves per SIMD spill).
template <class P, int BC = 0>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(sizeof(typename P::T) == 4 ? 5 : 2))) void k_ensemble2(
    const EnsDom *__restrict__ doms,
                                                   const EnsItem *__restrict__ items, int nitems,
                                                   int *__restrict__ counter, int parity,
                                                   typename P::K k)
{
    using T = typename P::T;
    constexpr bool DB = sizeof(T) == 4;
    constexpr int NB = DB ? 2 : 1;
    __shared__ __align__(16) T sh[NB * ENS2_LDS];
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0)
            next = atomicAdd(counter, 1);
        __syncthreads();
        const int it = next;
        __syncthreads();  // every thread has read `next` (and left the last item)
        if (it >= nitems)
            return;
        const EnsItem w = items[it];
        // the member's pointers read straight from the descriptor (a
        // private copy indexed by the runtime parity would live in scratch)
        const EnsDom *d = doms + w.dom;
        const T *si = (const T *)d->f[parity][0];
        const T *li = (const T *)d->f[parity][1];
        const T *ci = (const T *)d->f[parity][2];
        T *so = (T *)d->f[1 - parity][0];
        T *lo = (T *)d->f[1 - parity][1];
        T *co = (T *)d->f[1 - parity][2];
        const int dnx = d->nx, dny = d->ny, dnz = d->nz;
        // no register rotation in the march: with both tile shapes in one
        // kernel the FP32 form would spill with it at five waves per SIMD
        constexpr bool ROT = false;
        if (w.shape == ENS2_32x8)
            tile_march2<P, 32, 8, false, DB, BC, ROT>(
                si, li, ci, so, lo, co, dnx, dny, dnz, w.x0, w.y0, w.z0, w.zend, k,
                reinterpret_cast<T(*)[12][36]>(sh),
                reinterpret_cast<T(*)[10][34]>(sh + NB * 4 * 12 * 36));
        else
            tile_march2<P, 16, 16, false, DB, BC, ROT>(
                si, li, ci, so, lo, co, dnx, dny, dnz, w.x0, w.y0, w.z0, w.zend, k,
                reinterpret_cast<T(*)[20][20]>(sh),
                reinterpret_cast<T(*)[18][18]>(sh + NB * 4 * 20 * 20));
    }
}
