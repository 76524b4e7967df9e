// k_step.h -- step kernels: the naive FP64 form and the tiled form for any
// physics policy (PhysF64: bitwise equal to the reference; PhysF32).
//
// One launch = one sweep of phasefield.c:301-403 over the whole grid, with
// the ghost fill of phasefield.c:1112-1189 replaced by the index map
// ghost<BC>.  Input and output are separate buffers (ping-pong); the
// reference updates in place after computing every increment from the old
// state (pass C, phasefield.c:393-401), which is the same thing.
#pragma once

#include <type_traits>

#include "ppf_physics.h"
#include "ppf_reduce.h"

// Any coordinate (also beyond one cell outside the box: threads of a
// partial tile) to a valid source index.
template <int BC>
__device__ __forceinline__ int ghost_any(int i, int n)
{
    if (BC == 0) {
        int r = i % n;
        return r < 0 ? r + n : r;
    }
    return i < 0 ? 0 : (i >= n ? n - 1 : i);
}

// A pointer every lane holds the same value of, moved to scalar registers
// (readfirstlane) so that loads can use the scalar-base + 32-bit lane
// offset addressing form.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *uniform_ptr(T *p)
{
    const unsigned long long v = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return (__attribute__((address_space(1))) T *)(((unsigned long long)hi << 32) | lo);
}

// Element i of a plane whose base is uniform: the byte offset i*sizeof(T) is
// formed in 32 bits (contexts keep a plane under 4 GiB), so the access is
// one global instruction with the base in scalar registers and a 32-bit
// lane offset, with no 64-bit address arithmetic per lane.
template <typename T>
__device__ __forceinline__ T ld_plane(const __attribute__((address_space(1))) T *base, unsigned i)
{
    using C = const __attribute__((address_space(1))) char;
    return *(const __attribute__((address_space(1))) T *)((C *)base + i * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *at_plane(
    __attribute__((address_space(1))) T *base, unsigned i)
{
    using C = __attribute__((address_space(1))) char;
    return (__attribute__((address_space(1))) T *)((C *)base + i * (unsigned)sizeof(T));
}
// The same with the lane's byte offset b = i*sizeof(T) already formed (a
// march that keeps b across planes saves the shift per access group).
template <typename T>
__device__ __forceinline__ T ld_plane_b(const __attribute__((address_space(1))) T *base,
                                        unsigned b)
{
    using C = const __attribute__((address_space(1))) char;
    return *(const __attribute__((address_space(1))) T *)((C *)base + b);
}
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *at_plane_b(
    __attribute__((address_space(1))) T *base, unsigned b)
{
    using C = __attribute__((address_space(1))) char;
    return (__attribute__((address_space(1))) T *)((C *)base + b);
}

// The step-t (or t+1) state of one site column at one plane, as the z march
// keeps it in registers: the raw fields and pass A's per-site values.
template <typename T>
struct Col {
    T s, l, c, h, mu, M;
};

// The LDS planes of the rotating two-step march hold the four fields of a
// site (s, l, mu, M) side by side in 16-byte vectors: one per site in FP32,
// two in FP64 ((s, l) and (mu, M), the second set NS sites after the
// first).  A neighbour's four fields then come in one or two wide reads at
// an immediate offset from the thread's own site, from one address register
// per buffer, where separate field planes (a few KB apart, beyond the reach
// of the paired reads' offsets) need an address register per field.
template <typename T>
struct LdsVec;
template <>
struct LdsVec<float> {
    using V = float4;
    static constexpr int NG = 1;
};
template <>
struct LdsVec<double> {
    using V = double2;
    static constexpr int NG = 2;
};

template <int NS, typename T, typename V>
__device__ __forceinline__ void lds_put4(V *b, int i, T s, T l, T mu, T M)
{
    if constexpr (sizeof(T) == 4) {
        b[i] = make_float4(s, l, mu, M);
    } else {
        b[i] = make_double2(s, l);
        b[NS + i] = make_double2(mu, M);
    }
}

template <int NS, typename T, typename V>
__device__ __forceinline__ void lds_get4(const V *b, int i, T &s, T &l, T &mu, T &M)
{
    if constexpr (sizeof(T) == 4) {
        const V v = b[i];
        s = v.x;
        l = v.y;
        mu = v.z;
        M = v.w;
    } else {
        const V a = b[i], c = b[NS + i];
        s = a.x;
        l = a.y;
        mu = c.x;
        M = c.y;
    }
}

// The four in-plane neighbours of site i (row width RW) into slots 2..5 of
// the update's operand arrays: +y, -y, +x, -x.
template <int NS, int RW, typename T, typename V>
__device__ __forceinline__ void lds_nbrs(const V *b, int i, T sn[6], T ln[6], T mun[6], T Mn[6])
{
    lds_get4<NS>(b, i + RW, sn[2], ln[2], mun[2], Mn[2]);
    lds_get4<NS>(b, i - RW, sn[3], ln[3], mun[3], Mn[3]);
    lds_get4<NS>(b, i + 1, sn[4], ln[4], mun[4], Mn[4]);
    lds_get4<NS>(b, i - 1, sn[5], ln[5], mun[5], Mn[5]);
}

// Marks lane offsets as redefined at the top of each plane of a z march (an
// empty asm, no instruction): the compiler then forms each plane's
// addresses from the 32-bit offsets inside the loop body, as scalar base +
// 32-bit vector offset, instead of hoisting their 64-bit zero extensions
// out of the loop and adding 64-bit register pairs per access.
#define PPF_FRESH(...) asm volatile("" : __VA_ARGS__)

// ---------------------------------------------------------------------------
// Naive form: one thread per site, every neighbour read from global memory and
// its h, mu, M recomputed (7x the pointwise work).  Kept as an independent
// second implementation for cross-checks.
// ---------------------------------------------------------------------------
template <int BC, bool EQA, bool CONLY, bool REPORT>
__global__ __launch_bounds__(256) void k_naive_f64(
    const double *__restrict__ s_in, const double *__restrict__ l_in,
    const double *__restrict__ c_in, double *__restrict__ s_out, double *__restrict__ l_out,
    double *__restrict__ c_out, int nx, int ny, int nz, KConst k, double *__restrict__ partials)
{
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    double acc[3] = {0.0, 0.0, 0.0};
    if (x < nx && y < ny) {
        const size_t plane = (size_t)nx * ny;
        const size_t q = (size_t)z * plane + (size_t)y * nx + x;
        // neighbour flat indices in reference order z+1, z-1, y+1, y-1, x+1, x-1
        size_t nb[6];
        nb[0] = (size_t)ghost<BC>(z + 1, nz) * plane + (size_t)y * nx + x;
        nb[1] = (size_t)ghost<BC>(z - 1, nz) * plane + (size_t)y * nx + x;
        nb[2] = (size_t)z * plane + (size_t)ghost<BC>(y + 1, ny) * nx + x;
        nb[3] = (size_t)z * plane + (size_t)ghost<BC>(y - 1, ny) * nx + x;
        nb[4] = (size_t)z * plane + (size_t)y * nx + ghost<BC>(x + 1, nx);
        nb[5] = (size_t)z * plane + (size_t)y * nx + ghost<BC>(x - 1, nx);
        const double s = s_in[q], l = l_in[q], c = c_in[q];
        double h, mu, M;
        pointwise<EQA>(s, l, c, k, h, mu, M);
        double sn[6], ln[6], mun[6], Mn[6];
        for (int a = 0; a < 6; a++) {
            sn[a] = s_in[nb[a]];
            ln[a] = l_in[nb[a]];
            double hh;
            pointwise<EQA>(sn[a], ln[a], c_in[nb[a]], k, hh, mun[a], Mn[a]);
        }
        if (CONLY) {
            const double dc = dc_only(mu, M, mun, Mn, k);
            c_out[q] = c + dc;
            acc[0] = dc * dc;
        } else {
            double so, lo, co, ds, dl, dc;
            site_update<EQA>(s, l, c, h, mu, M, sn, ln, mun, Mn, k, so, lo, co, ds, dl, dc);
            s_out[q] = so;
            l_out[q] = lo;
            c_out[q] = co;
            acc[0] = dc * dc;
            acc[1] = ds * ds;
            acc[2] = dl * dl;
        }
    }
    if (REPORT) {
        const int bid = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
        block_sum3<256>(acc, partials + 3 * (size_t)bid);
    }
}

// ---------------------------------------------------------------------------
// Tiled form.  A workgroup owns a TX x TY column tile of the (x, y) plane and
// marches it through tz planes of z.  Per plane:
//   * each thread holds its column's s, l, mu, M for planes z-1, z and z+1
//     in registers (z neighbours), computing h, mu, M once per site;
//   * the current plane's s, l, mu, M of the tile and of its one-cell ring
//     (faces only; corners are never read) go to LDS, double buffered so one
//     barrier per plane suffices;
//   * every thread updates its site from LDS (x, y neighbours) and registers
//     (z neighbours) and writes the new (s, l, c).
// With PIPE the march is software pipelined: the raw (s, l, c) of the
// column two planes ahead and of the ring one plane ahead are loaded an
// iteration before they are used, so each wave keeps two planes of loads in
// flight while it computes (more registers, fewer waves per SIMD).  With
// PIPE, CD > 2 loads the column CD planes ahead instead of two (CD - 2 more
// planes in flight, 6 VGPRs each in FP64).  Compulsory HBM traffic is one read and one write of each field
// per site; ring reads and the extra planes of each tz chunk come mostly
// from L2.
// ---------------------------------------------------------------------------
// One workgroup's march over the column tile [x0, x0+TX) x [y0, y0+TY) and
// the planes [z0, zend) of one grid.  sh is the workgroup's LDS,
// 2 x 4 x (TY+2) x (TX+2) values; acc gets the REPORT sums of squared
// increments.  Every thread of the workgroup must call it.
template <class P, int TX, int TY, int BC, bool CONLY, bool REPORT, bool PIPE, int CD = 2>
__device__ __forceinline__ void tile_march(
    const typename P::T *__restrict__ s_in, const typename P::T *__restrict__ l_in,
    const typename P::T *__restrict__ c_in, typename P::T *__restrict__ s_out,
    typename P::T *__restrict__ l_out, typename P::T *__restrict__ c_out, int nx, int ny, int nz,
    int x0, int y0, int z0, int zend, const typename P::K &k,
    typename P::T (*__restrict__ sh)[4][TY + 2][TX + 2], double acc[3])
{
    constexpr int NRING = 2 * TX + 2 * TY;
    static_assert(NRING <= TX * TY, "ring larger than the workgroup");
    using T = typename P::T;
    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    const int x = x0 + tx, y = y0 + ty;
    const bool active = x < nx && y < ny;
    const size_t plane = (size_t)nx * ny;
    // Addressing: the plane base (z is uniform over the workgroup) lives in
    // scalar registers; each lane keeps only 32-bit offsets within a plane
    // (nx*ny < 2^31 is checked at context creation).
#define PLANE(p, zz) uniform_ptr((p) + (size_t)(zz) * plane)
    unsigned col = (unsigned)ghost_any<BC>(y, ny) * (unsigned)nx + (unsigned)ghost_any<BC>(x, nx);
    unsigned ocol = (unsigned)y * (unsigned)nx + (unsigned)x;

    // This thread's ring site, if any: left, right, bottom, top faces.
    int rx = 0, ry = 0;
    if (tid < TY) {
        rx = 0;
        ry = tid + 1;
    } else if (tid < 2 * TY) {
        rx = TX + 1;
        ry = tid - TY + 1;
    } else if (tid < 2 * TY + TX) {
        rx = tid - 2 * TY + 1;
        ry = 0;
    } else if (tid < NRING) {
        rx = tid - 2 * TY - TX + 1;
        ry = TY + 1;
    }
    const bool ring = tid < NRING;
    unsigned rcol = (unsigned)ghost_any<BC>(y0 + ry - 1, ny) * (unsigned)nx +
                          (unsigned)ghost_any<BC>(x0 + rx - 1, nx);

    // Prologue: planes z0-1 and z0 computed, plane z0+1 and the ring of z0
    // loaded raw.
    T s_p, l_p, mu_p, M_p;
    T s_c, l_c, c_c, h_c, mu_c, M_c;
    T rs_n, rl_n, rc_n;                // raw column values one plane ahead
    T rs_r = 0, rl_r = 0, rc_r = 0;    // raw ring values of the current plane
    constexpr int NQ = CD > 2 ? CD - 2 : 1;
    T qs[NQ], ql[NQ], qc[NQ];          // raw column values 2 .. CD-1 planes ahead
    {
        const int zp = ghost<BC>(z0 - 1, nz), zn = ghost_any<BC>(z0 + 1, nz);
        s_p = ld_plane(PLANE(s_in, zp), col);
        l_p = ld_plane(PLANE(l_in, zp), col);
        const T c_p = ld_plane(PLANE(c_in, zp), col);
        s_c = ld_plane(PLANE(s_in, z0), col);
        l_c = ld_plane(PLANE(l_in, z0), col);
        c_c = ld_plane(PLANE(c_in, z0), col);
        rs_n = ld_plane(PLANE(s_in, zn), col);
        rl_n = ld_plane(PLANE(l_in, zn), col);
        rc_n = ld_plane(PLANE(c_in, zn), col);
        if (ring) {
            rs_r = ld_plane(PLANE(s_in, z0), rcol);
            rl_r = ld_plane(PLANE(l_in, z0), rcol);
            rc_r = ld_plane(PLANE(c_in, z0), rcol);
        }
#pragma unroll
        for (int k = 0; k < NQ; k++) {
            qs[k] = ql[k] = qc[k] = 0;
            if (PIPE && CD > 2 && z0 + 1 + k < zend) {
                const int z2 = ghost_any<BC>(z0 + 2 + k, nz);
                qs[k] = ld_plane(PLANE(s_in, z2), col);
                ql[k] = ld_plane(PLANE(l_in, z2), col);
                qc[k] = ld_plane(PLANE(c_in, z2), col);
            }
        }
        T hp;
        P::point(s_p, l_p, c_p, k, hp, mu_p, M_p);
        P::point(s_c, l_c, c_c, k, h_c, mu_c, M_c);
    }

    for (int z = z0; z < zend; ++z) {
        PPF_FRESH("+v"(col), "+v"(ocol), "+v"(rcol));
        const int buf = (z - z0) & 1;
        const bool more = z + 1 < zend;
        if (!PIPE && z > z0) {
            // unpipelined form: this plane's ring and the next plane now
            const int zn = ghost_any<BC>(z + 1, nz);
            rs_n = ld_plane(PLANE(s_in, zn), col);
            rl_n = ld_plane(PLANE(l_in, zn), col);
            rc_n = ld_plane(PLANE(c_in, zn), col);
            if (ring) {
                rs_r = ld_plane(PLANE(s_in, z), rcol);
                rl_r = ld_plane(PLANE(l_in, z), rcol);
                rc_r = ld_plane(PLANE(c_in, z), rcol);
            }
        }
        // prefetch: column CD planes ahead, ring one plane ahead
        T ps = 0, pl = 0, pc = 0, prs = 0, prl = 0, prc = 0;
        if (PIPE && z + CD - 1 < zend) {
            const int zf = ghost_any<BC>(z + CD, nz);
            ps = ld_plane(PLANE(s_in, zf), col);
            pl = ld_plane(PLANE(l_in, zf), col);
            pc = ld_plane(PLANE(c_in, zf), col);
        }
        if (PIPE && more) {
            if (ring) {
                prs = ld_plane(PLANE(s_in, z + 1), rcol);
                prl = ld_plane(PLANE(l_in, z + 1), rcol);
                prc = ld_plane(PLANE(c_in, z + 1), rcol);
            }
        }

        if (!CONLY) {
            sh[buf][0][ty + 1][tx + 1] = s_c;
            sh[buf][1][ty + 1][tx + 1] = l_c;
        }
        sh[buf][2][ty + 1][tx + 1] = mu_c;
        sh[buf][3][ty + 1][tx + 1] = M_c;
        if (ring) {
            T hr, mur, Mr;
            P::point(rs_r, rl_r, rc_r, k, hr, mur, Mr);
            if (!CONLY) {
                sh[buf][0][ry][rx] = rs_r;
                sh[buf][1][ry][rx] = rl_r;
            }
            sh[buf][2][ry][rx] = mur;
            sh[buf][3][ry][rx] = Mr;
        }
        T h_n, mu_n, M_n;
        P::point(rs_n, rl_n, rc_n, k, h_n, mu_n, M_n);
        __syncthreads();

        if (active) {
            const T mun[6] = {mu_n, mu_p, sh[buf][2][ty + 2][tx + 1], sh[buf][2][ty][tx + 1],
                                   sh[buf][2][ty + 1][tx + 2], sh[buf][2][ty + 1][tx]};
            const T Mn[6] = {M_n, M_p, sh[buf][3][ty + 2][tx + 1], sh[buf][3][ty][tx + 1],
                                  sh[buf][3][ty + 1][tx + 2], sh[buf][3][ty + 1][tx]};
            if (CONLY) {
                const T dc = P::conly(mu_c, M_c, mun, Mn, k);
                __builtin_nontemporal_store((T)(c_c + dc), at_plane(PLANE(c_out, z), ocol));
                if (REPORT)
                    acc[0] += (double)dc * (double)dc;
            } else {
                const T sn[6] = {rs_n, s_p, sh[buf][0][ty + 2][tx + 1], sh[buf][0][ty][tx + 1],
                                      sh[buf][0][ty + 1][tx + 2], sh[buf][0][ty + 1][tx]};
                const T ln[6] = {rl_n, l_p, sh[buf][1][ty + 2][tx + 1], sh[buf][1][ty][tx + 1],
                                      sh[buf][1][ty + 1][tx + 2], sh[buf][1][ty + 1][tx]};
                T so, lo, co;
                double ds, dl, dc;
                P::update(s_c, l_c, c_c, h_c, mu_c, M_c, sn, ln, mun, Mn, k, so, lo, co, ds, dl,
                          dc);
                // streaming stores: the new state is not read again this step
                __builtin_nontemporal_store(so, at_plane(PLANE(s_out, z), ocol));
                __builtin_nontemporal_store(lo, at_plane(PLANE(l_out, z), ocol));
                __builtin_nontemporal_store(co, at_plane(PLANE(c_out, z), ocol));
                if (REPORT) {
                    acc[0] += dc * dc;
                    acc[1] += ds * ds;
                    acc[2] += dl * dl;
                }
            }
        }
        s_p = s_c;
        l_p = l_c;
        mu_p = mu_c;
        M_p = M_c;
        s_c = rs_n;
        l_c = rl_n;
        c_c = rc_n;
        h_c = h_n;
        mu_c = mu_n;
        M_c = M_n;
        if (PIPE) {
            if (CD > 2) {
                rs_n = qs[0];
                rl_n = ql[0];
                rc_n = qc[0];
#pragma unroll
                for (int k = 0; k + 1 < NQ; k++) {
                    qs[k] = qs[k + 1];
                    ql[k] = ql[k + 1];
                    qc[k] = qc[k + 1];
                }
                qs[NQ - 1] = ps;
                ql[NQ - 1] = pl;
                qc[NQ - 1] = pc;
            } else {
                rs_n = ps;
                rl_n = pl;
                rc_n = pc;
            }
            rs_r = prs;
            rl_r = prl;
            rc_r = prc;
        }
    }
#undef PLANE
}

template <class P, int TX, int TY, int BC, bool CONLY, bool REPORT, bool PIPE, int CD = 2>
__global__ __launch_bounds__(TX *TY) void k_tiled(
    const typename P::T *__restrict__ s_in, const typename P::T *__restrict__ l_in,
    const typename P::T *__restrict__ c_in, typename P::T *__restrict__ s_out,
    typename P::T *__restrict__ l_out, typename P::T *__restrict__ c_out, int nx, int ny, int nz,
    int tz, int ntx, int nty, int order, typename P::K k, double *__restrict__ partials)
{
    constexpr int NT = TX * TY;
    using T = typename P::T;
    __shared__ T sh[2][4][TY + 2][TX + 2];
    const int ntile = ntx * nty;
    const int nblk = gridDim.x;
    int b = blockIdx.x;
    // Workgroups are dealt round-robin to the 8 XCDs (b % 8); with the XCD
    // bit of `order` set, each XCD instead gets one contiguous run of the
    // logical order, so tiles that share ring and halo planes share an L2.
    if ((order & 2) && (nblk & 7) == 0)
        b = (b & 7) * (nblk >> 3) + (b >> 3);
    // order bit 0: z chunk fastest (all chunks of one tile adjacent);
    // otherwise tiles fastest (one z chunk of every tile adjacent).
    const int tile = (order & 1) ? b / ((nblk + ntile - 1) / ntile) : b % ntile;
    const int zc = (order & 1) ? b % ((nblk + ntile - 1) / ntile) : b / ntile;
    const int x0 = (tile % ntx) * TX, y0 = (tile / ntx) * TY, z0 = zc * tz;
    const int zend = min(z0 + tz, nz);
    double acc[3] = {0.0, 0.0, 0.0};
    tile_march<P, TX, TY, BC, CONLY, REPORT, PIPE, CD>(s_in, l_in, c_in, s_out, l_out, c_out, nx, ny,
                                                   nz, x0, y0, z0, zend, k, sh, acc);
    if (REPORT)
        block_sum3<NT>(acc, partials + 3 * (size_t)blockIdx.x);
}

// ---------------------------------------------------------------------------
// Two steps per launch (periodic grids; BC = 1: no-flux).  The new state of step t+2 is
// computed from step t in one z march, with step t+1 kept on chip:
//   stage 1, plane z:   step t -> t+1 on the tile and its one-cell face ring
//                        (the ring sites are the in-plane neighbours stage 2
//                        needs); step t of the tile, of that ring, of the
//                        two-cell face ring and of the four one-ring corners
//                        sits in LDS (A1);
//   stage 2, plane z-1: step t+1 -> t+2 on the tile, from A2 (step t+1 of
//                        the tile and its ring) and registers (z neighbours).
// Every site, ring sites included, is updated at its periodic image with the
// same operands in the same order as a single step, so two steps here equal
// two single steps bit for bit.  With mirror (no-flux) boundaries a ring site
// or plane outside the grid is not updated at all (computing it at its image
// would reorder the image site's two x or y neighbours): it takes the step
// t+1 of the site it mirrors, which that site's thread writes.  Each output
// plane is read and written once for two steps: 24 B per site update in
// FP64 instead of 48.  A z chunk [z0, zend) runs stage 1 over
// [z0-1, zend] and so reads step t over [z0-2, zend+1].
// Per plane: A1 written, barrier, stage 1 (tile and ring) and stage 2,
// barrier, A2 written.  The barriers separate every write of A1 or A2 from
// the reads of the previous plane, so both are single buffered.
// Threads: tid < NR1 also own a ring site (its own z column of step t);
// NR1 <= tid < NR1 + NO2 load one outer site (two-cell ring or corner) of
// the current plane.
// PF: the raw step-t values a plane needs (tile and ring columns at zs+1,
// outer site at zs) are loaded one iteration earlier, so the loads stay in
// flight over a whole plane of arithmetic and both barriers (which wait for
// LDS only) instead of stalling the plane's first h, mu, M.  Nine more
// values in registers per thread.
// DB: A1 and A2 double buffered (a1, a2 hold two sets of 4 arrays each), one
// barrier per plane.  A1 of plane zs goes to buffer zs&1, last read by stage
// 1 two planes earlier (the previous plane's barrier has passed); A2 of plane
// zs is written right after stage 1 into the buffer stage 2 does not read
// this plane, last read by stage 2 of the previous plane (before this
// plane's barrier).
// ---------------------------------------------------------------------------
// ROT: the march without early loads rotates three register sets per
// column instead of copying (see below); callers with a tighter register
// budget (the no-flux FP32 ensemble kernel) switch it off.
// FULL (ROT, periodic): every tile site lies in the grid (nx % TX == 0 and
// ny % TY == 0), so the output stores are unconditional and are deferred:
// a plane's step t+2 values stay in registers and are stored after the
// next plane's loads have been issued.  Loads and stores share one in-order
// counter of outstanding vector memory operations, so a store issued before
// a plane's loads would have to complete before those loads could be used;
// issued after them, it completes while the plane is computed.
template <class P, int TX, int TY, bool PF = false, bool DB = false, int BC = 0,
          bool ROT = P::kRotate, bool FULL = false>
__device__ __forceinline__ void tile_march2(
    const typename P::T *__restrict__ s_in, const typename P::T *__restrict__ l_in,
    const typename P::T *__restrict__ c_in, typename P::T *__restrict__ s_out,
    typename P::T *__restrict__ l_out, typename P::T *__restrict__ c_out, int nx, int ny, int nz,
    int x0, int y0, int z0, int zend, const typename P::K &k,
    typename P::T (*__restrict__ a1)[TY + 4][TX + 4], typename P::T (*__restrict__ a2)[TY + 2][TX + 2])
{
    constexpr int NR1 = 2 * TX + 2 * TY, NO2 = 2 * TX + 2 * TY + 4;
    static_assert(NR1 + NO2 <= TX * TY, "ring sites exceed the workgroup");
    using T = typename P::T;
    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    const bool active = x0 + tx < nx && y0 + ty < ny;
    const size_t plane = (size_t)nx * ny;
#define PLANE(p, zz) uniform_ptr((p) + (size_t)(zz) * plane)
    // site (x0 + ax, y0 + ay) of the tile frame -> offset within a plane
    auto colof = [&](int ax, int ay) {
        return (unsigned)ghost_any<BC>(y0 + ay, ny) * (unsigned)nx + (unsigned)ghost_any<BC>(x0 + ax, nx);
    };
    unsigned col = colof(tx, ty);
    unsigned ocol = (unsigned)(y0 + ty) * (unsigned)nx + (unsigned)(x0 + tx);

    // one-cell ring site (A2 frame rx, ry; tile frame rx-1, ry-1)
    int rx = 0, ry = 0;
    if (tid < TY) {
        rx = 0;
        ry = tid + 1;
    } else if (tid < 2 * TY) {
        rx = TX + 1;
        ry = tid - TY + 1;
    } else if (tid < 2 * TY + TX) {
        rx = tid - 2 * TY + 1;
        ry = 0;
    } else if (tid < NR1) {
        rx = tid - 2 * TY - TX + 1;
        ry = TY + 1;
    }
    const bool ring = tid < NR1;
    unsigned rcol = colof(rx - 1, ry - 1);
    // mirror boundaries: a ring site outside the grid is a ghost, whose step
    // t+1 is its mirror site's; the tile thread of that site writes it
    const bool ring_in = BC == 0 || (x0 + rx - 1 >= 0 && x0 + rx - 1 < nx && y0 + ry - 1 >= 0 &&
                                     y0 + ry - 1 < ny);
    // outer site (A1 frame ox, oy; tile frame ox-2, oy-2)
    int ox = 0, oy = 0;
    {
        const int j = tid - NR1;
        if (j < 0) {
        } else if (j < TY) {
            ox = 0;
            oy = j + 2;
        } else if (j < 2 * TY) {
            ox = TX + 3;
            oy = j - TY + 2;
        } else if (j < 2 * TY + TX) {
            ox = j - 2 * TY + 2;
            oy = 0;
        } else if (j < 2 * TY + 2 * TX) {
            ox = j - 2 * TY - TX + 2;
            oy = TY + 3;
        } else if (j < NO2) {
            const int c = j - 2 * TY - 2 * TX;
            ox = (c & 1) ? TX + 2 : 1;
            oy = (c & 2) ? TY + 2 : 1;
        }
    }
    const bool outer = tid >= NR1 && tid < NR1 + NO2;
    unsigned o2col = colof(ox - 2, oy - 2);

    // step t columns of the tile site (p: z-1, c: z, n: z+1 raw) and of the
    // ring site (q*)
    T s_p, l_p, mu_p, M_p, s_c, l_c, c_c, h_c, mu_c, M_c;
    T qs_p = 0, ql_p = 0, qmu_p = 0, qM_p = 0, qs_c = 0, ql_c = 0, qc_c = 0, qh_c = 0, qmu_c = 0,
      qM_c = 0;
    {
        const int zp = ghost_any<BC>(z0 - 2, nz), zc = ghost_any<BC>(z0 - 1, nz);
        T hp;
        s_p = ld_plane(PLANE(s_in, zp), col);
        l_p = ld_plane(PLANE(l_in, zp), col);
        P::point(s_p, l_p, ld_plane(PLANE(c_in, zp), col), k, hp, mu_p, M_p);
        s_c = ld_plane(PLANE(s_in, zc), col);
        l_c = ld_plane(PLANE(l_in, zc), col);
        c_c = ld_plane(PLANE(c_in, zc), col);
        P::point(s_c, l_c, c_c, k, h_c, mu_c, M_c);
        if (ring) {
            qs_p = ld_plane(PLANE(s_in, zp), rcol);
            ql_p = ld_plane(PLANE(l_in, zp), rcol);
            T qh;
            P::point(qs_p, ql_p, ld_plane(PLANE(c_in, zp), rcol), k, qh, qmu_p, qM_p);
            qs_c = ld_plane(PLANE(s_in, zc), rcol);
            ql_c = ld_plane(PLANE(l_in, zc), rcol);
            qc_c = ld_plane(PLANE(c_in, zc), rcol);
            P::point(qs_c, ql_c, qc_c, k, qh_c, qmu_c, qM_c);
        }
    }
    // step t+1 column of the tile site: u*_p plane zs-2, u*_c plane zs-1
    T us_p = 0, ul_p = 0, umu_p = 0, uM_p = 0, us_c = 0, ul_c = 0, uc_c = 0, uh_c = 0, umu_c = 0,
      uM_c = 0;
    // step t raw values of plane zs+1 (tile and ring columns) and of plane
    // zs (outer site); with PF the first plane's are loaded here
    T rs_n = 0, rl_n = 0, rc_n = 0, qs_n = 0, ql_n = 0, qc_n = 0, os = 0, ol = 0, oc = 0;
    if (PF) {
        const int zn = ghost_any<BC>(z0, nz), zc = ghost_any<BC>(z0 - 1, nz);
        rs_n = ld_plane(PLANE(s_in, zn), col);
        rl_n = ld_plane(PLANE(l_in, zn), col);
        rc_n = ld_plane(PLANE(c_in, zn), col);
        if (ring) {
            qs_n = ld_plane(PLANE(s_in, zn), rcol);
            ql_n = ld_plane(PLANE(l_in, zn), rcol);
            qc_n = ld_plane(PLANE(c_in, zn), rcol);
        }
        if (outer) {
            os = ld_plane(PLANE(s_in, zc), o2col);
            ol = ld_plane(PLANE(l_in, zc), o2col);
            oc = ld_plane(PLANE(c_in, zc), o2col);
        }
    }

    if constexpr (!PF && ROT) {
        // Without the early loads the state of three consecutive planes
        // rotates through three register sets (tile column, ring column and
        // step t+1 column: z-1, z, z+1), and the march is unrolled by three
        // with the sets passed in rotated order, so no plane's values are
        // copied from one register to another.
        Col<T> ta{s_p, l_p, 0, 0, mu_p, M_p}, tb{s_c, l_c, c_c, h_c, mu_c, M_c}, tn{};
        Col<T> qa{qs_p, ql_p, 0, 0, qmu_p, qM_p}, qb{qs_c, ql_c, qc_c, qh_c, qmu_c, qM_c}, qn{};
        Col<T> ua{}, ub{}, un{};
        // LDS as 16-byte site vectors (LdsVec): A1 buffer b at v1 + b*NG*S1,
        // A2 buffer b at v2 + b*NG*S2; site indices of this thread's tile,
        // ring and outer sites
        using V = typename LdsVec<T>::V;
        constexpr int NG = LdsVec<T>::NG;
        constexpr int W1 = TX + 4, W2 = TX + 2, S1 = (TY + 4) * W1, S2 = (TY + 2) * W2;
        static_assert(sizeof(V) * NG == 4 * sizeof(T), "four fields per site");
        V *const v1 = reinterpret_cast<V *>(&a1[0][0][0]);
        V *const v2 = reinterpret_cast<V *>(&a2[0][0][0]);
        const int i1 = (ty + 2) * W1 + tx + 2, ir1 = (ry + 1) * W1 + rx + 1, io1 = oy * W1 + ox;
        const int i2 = (ty + 1) * W2 + tx + 1, ir2 = ry * W2 + rx;
        // the lane offsets as byte offsets
        unsigned bcol = col * (unsigned)sizeof(T), bocol = ocol * (unsigned)sizeof(T),
                 brcol = rcol * (unsigned)sizeof(T), bo2col = o2col * (unsigned)sizeof(T);
        // FULL: step t+2 values of plane zs-1, stored by the next plane
        constexpr bool DEF = FULL && BC == 0;
        T ps = 0, pl = 0, pc = 0;
        // OEARLY: the outer site's loads are issued a plane ahead (eo*), and
        // the point values of the z+1 columns are computed after the plane's
        // barrier
        constexpr bool OEARLY = true;
        T eos = 0, eol = 0, eoc = 0;
        // PEND (an integral constant): a deferred store of plane zs-2 is
        // pending, which also means zs >= z0 + 2
        auto plane_step = [&](auto PEND, const int zs, const Col<T> &tp, const Col<T> &tc,
                              Col<T> &tnx, const Col<T> &qp, const Col<T> &qc, Col<T> &qnx,
                              const Col<T> &up, Col<T> &uc, Col<T> &unx) {
            constexpr bool pend = decltype(PEND)::value;
            PPF_FRESH("+v"(bcol), "+v"(bocol), "+v"(brcol), "+v"(bo2col));
            const int zc = ghost_any<BC>(zs, nz), zn = ghost_any<BC>(zs + 1, nz);
            const int bsel = DB ? ((zs - z0 + 1) & 1) : 0;
            V *const A1 = v1 + bsel * (NG * S1);
            V *const A2 = v2 + bsel * (NG * S2);
            const V *const A2p = v2 + (DB ? bsel ^ 1 : 0) * (NG * S2);
            tnx.s = ld_plane_b(PLANE(s_in, zn), bcol);
            tnx.l = ld_plane_b(PLANE(l_in, zn), bcol);
            tnx.c = ld_plane_b(PLANE(c_in, zn), bcol);
            if (ring) {
                qnx.s = ld_plane_b(PLANE(s_in, zn), brcol);
                qnx.l = ld_plane_b(PLANE(l_in, zn), brcol);
                qnx.c = ld_plane_b(PLANE(c_in, zn), brcol);
            }
            // (read only by the outer threads, which load them; OEARLY: loaded
            // by the previous plane)
            T os = __builtin_nondeterministic_value(os), ol = __builtin_nondeterministic_value(ol),
              oc = __builtin_nondeterministic_value(oc);
            if constexpr (OEARLY) {
                // this plane's outer site (loaded by the previous plane), and
                // the next plane's loads issued before the pending stores, so
                // that the wait for this plane's loads leaves the stores in
                // flight
                os = eos;
                ol = eol;
                oc = eoc;
                if (outer) {
                    eos = ld_plane_b(PLANE(s_in, zn), bo2col);
                    eol = ld_plane_b(PLANE(l_in, zn), bo2col);
                    eoc = ld_plane_b(PLANE(c_in, zn), bo2col);
                }
            } else if (outer) {
                os = ld_plane_b(PLANE(s_in, zc), bo2col);
                ol = ld_plane_b(PLANE(l_in, zc), bo2col);
                oc = ld_plane_b(PLANE(c_in, zc), bo2col);
            }
            if constexpr (pend) {
                __builtin_nontemporal_store(ps, at_plane_b(PLANE(s_out, zs - 2), bocol));
                __builtin_nontemporal_store(pl, at_plane_b(PLANE(l_out, zs - 2), bocol));
                __builtin_nontemporal_store(pc, at_plane_b(PLANE(c_out, zs - 2), bocol));
            }
            // A1: step t of plane zs
            lds_put4<S1>(A1, i1, tc.s, tc.l, tc.mu, tc.M);
            if (ring)
                lds_put4<S1>(A1, ir1, qc.s, qc.l, qc.mu, qc.M);
            if (outer) {
                T oh, omu, oM;
                P::point(os, ol, oc, k, oh, omu, oM);
                lds_put4<S1>(A1, io1, os, ol, omu, oM);
            }
            if constexpr (OEARLY) {
                // the z+1 columns' point values after the barrier: no wave
                // waits for this plane's loads before it
                __syncthreads();
                P::point(tnx.s, tnx.l, tnx.c, k, tnx.h, tnx.mu, tnx.M);
                if (ring)
                    P::point(qnx.s, qnx.l, qnx.c, k, qnx.h, qnx.mu, qnx.M);
            } else {
                P::point(tnx.s, tnx.l, tnx.c, k, tnx.h, tnx.mu, tnx.M);
                if (ring)
                    P::point(qnx.s, qnx.l, qnx.c, k, qnx.h, qnx.mu, qnx.M);
                __syncthreads();
            }

            // stage 1: step t+1 of plane zs at the tile site (mirror
            // boundaries: a ghost plane is not stepped, it keeps the step
            // t+1 of the plane before it)
            const bool zghost = BC == 1 && (zs < 0 || zs >= nz);
            if constexpr (BC == 1)
                unx = uc;
            if (!zghost) {
                T sn[6] = {tnx.s, tp.s}, ln[6] = {tnx.l, tp.l}, mun[6] = {tnx.mu, tp.mu},
                  Mn[6] = {tnx.M, tp.M};
                lds_nbrs<S1, W1>(A1, i1, sn, ln, mun, Mn);
                double ds, dl, dc;
                P::update(tc.s, tc.l, tc.c, tc.h, tc.mu, tc.M, sn, ln, mun, Mn, k, unx.s, unx.l,
                          unx.c, ds, dl, dc);
                P::point(unx.s, unx.l, unx.c, k, unx.h, unx.mu, unx.M);
            }
            // ... and at the ring site
            T qs1 = 0, ql1 = 0, qmu1 = 0, qM1 = 0;
            if (ring && ring_in && !zghost) {
                T sn[6] = {qnx.s, qp.s}, ln[6] = {qnx.l, qp.l}, mun[6] = {qnx.mu, qp.mu},
                  Mn[6] = {qnx.M, qp.M};
                lds_nbrs<S1, W1>(A1, ir1, sn, ln, mun, Mn);
                T qc1, qh1;
                double ds, dl, dc;
                P::update(qc.s, qc.l, qc.c, qc.h, qc.mu, qc.M, sn, ln, mun, Mn, k, qs1, ql1, qc1, ds,
                          dl, dc);
                P::point(qs1, ql1, qc1, k, qh1, qmu1, qM1);
            }
            // A2: step t+1 of plane zs (plain stores, see tile_march2's loop)
            auto store_a2 = [&]() {
                if (BC == 0 || active)
                    lds_put4<S2>(A2, i2, unx.s, unx.l, unx.mu, unx.M);
                if constexpr (BC == 1) {
                    if (active) {
                        if (x0 + tx == 0)
                            lds_put4<S2>(A2, i2 - 1, unx.s, unx.l, unx.mu, unx.M);
                        if (x0 + tx == nx - 1)
                            lds_put4<S2>(A2, i2 + 1, unx.s, unx.l, unx.mu, unx.M);
                        if (y0 + ty == 0)
                            lds_put4<S2>(A2, i2 - W2, unx.s, unx.l, unx.mu, unx.M);
                        if (y0 + ty == ny - 1)
                            lds_put4<S2>(A2, i2 + W2, unx.s, unx.l, unx.mu, unx.M);
                    }
                }
                if (ring && ring_in)
                    lds_put4<S2>(A2, ir2, qs1, ql1, qmu1, qM1);
            };
            if (DB)
                store_a2();
            // stage 2: step t+2 of plane zs-1 at the tile site
            if ((pend || zs >= z0 + 1) && (DEF || active)) {
                T sn[6] = {unx.s, up.s}, ln[6] = {unx.l, up.l}, mun[6] = {unx.mu, up.mu},
                  Mn[6] = {unx.M, up.M};
                lds_nbrs<S2, W2>(A2p, i2, sn, ln, mun, Mn);
                T so, lo, co;
                double ds, dl, dc;
                P::update(uc.s, uc.l, uc.c, uc.h, uc.mu, uc.M, sn, ln, mun, Mn, k, so, lo, co, ds, dl,
                          dc);
                if constexpr (DEF) {
                    ps = so;
                    pl = lo;
                    pc = co;
                } else {
                    __builtin_nontemporal_store(so, at_plane_b(PLANE(s_out, zs - 1), bocol));
                    __builtin_nontemporal_store(lo, at_plane_b(PLANE(l_out, zs - 1), bocol));
                    __builtin_nontemporal_store(co, at_plane_b(PLANE(c_out, zs - 1), bocol));
                }
            }
            if (!DB) {
                __syncthreads();
                store_a2();
            }
            // mirror boundary below plane 0: ghost plane -1 holds plane 0's
            // step t+1 (uc becomes the next plane's z-1 set)
            if (BC == 1 && z0 == 0 && zs == 0) {
                uc.s = unx.s;
                uc.l = unx.l;
                uc.mu = unx.mu;
                uc.M = unx.M;
            }
        };
        using NoPend = std::integral_constant<bool, false>;
        using Pend = std::integral_constant<bool, DEF>;
        int zs = z0 - 1;
        if constexpr (OEARLY) {
            if (outer) {  // the first plane's outer site
                const int zo = ghost_any<BC>(zs, nz);
                eos = ld_plane_b(PLANE(s_in, zo), bo2col);
                eol = ld_plane_b(PLANE(l_in, zo), bo2col);
                eoc = ld_plane_b(PLANE(c_in, zo), bo2col);
            }
        }
        if constexpr (DEF) {
            // planes z0-1 .. z0+1 (zend >= z0 + 1): nothing pending yet
            plane_step(NoPend{}, zs, ta, tb, tn, qa, qb, qn, ua, ub, un);
            plane_step(NoPend{}, zs + 1, tb, tn, ta, qb, qn, qa, ub, un, ua);
            plane_step(NoPend{}, zs + 2, tn, ta, tb, qn, qa, qb, un, ua, ub);
            zs += 3;
        }
        for (; zs + 2 <= zend; zs += 3) {
            plane_step(Pend{}, zs, ta, tb, tn, qa, qb, qn, ua, ub, un);
            plane_step(Pend{}, zs + 1, tb, tn, ta, qb, qn, qa, ub, un, ua);
            plane_step(Pend{}, zs + 2, tn, ta, tb, qn, qa, qb, un, ua, ub);
        }
        // the last one or two planes, rotating by copies.  (Policies whose
        // registers would not hold the third set at three waves per SIMD,
        // the bitwise FP64 one, take the loop below instead.)
        for (; zs <= zend; ++zs) {
            plane_step(Pend{}, zs, ta, tb, tn, qa, qb, qn, ua, ub, un);
            ta = tb;
            tb = tn;
            qa = qb;
            qb = qn;
            ua = ub;
            ub = un;
        }
        if constexpr (DEF) {
            // plane zend-1, stepped by the last plane_step
            __builtin_nontemporal_store(ps, at_plane_b(PLANE(s_out, zend - 1), bocol));
            __builtin_nontemporal_store(pl, at_plane_b(PLANE(l_out, zend - 1), bocol));
            __builtin_nontemporal_store(pc, at_plane_b(PLANE(c_out, zend - 1), bocol));
        }
    } else {
    // LDS as 16-byte site vectors, as in the rotating march above
    using V = typename LdsVec<T>::V;
    constexpr int NG = LdsVec<T>::NG;
    constexpr int W1 = TX + 4, W2 = TX + 2, S1 = (TY + 4) * W1, S2 = (TY + 2) * W2;
    V *const v1 = reinterpret_cast<V *>(&a1[0][0][0]);
    V *const v2 = reinterpret_cast<V *>(&a2[0][0][0]);
    const int i1 = (ty + 2) * W1 + tx + 2, ir1 = (ry + 1) * W1 + rx + 1, io1 = oy * W1 + ox;
    const int i2 = (ty + 1) * W2 + tx + 1, ir2 = ry * W2 + rx;
    // without PF: the outer site's loads are issued a plane ahead, and the
    // z+1 columns' point values are computed after the plane's barrier (as
    // in the rotating march: no wave waits for the plane's loads before it)
    if (!PF && outer) {
        const int zo = ghost_any<BC>(z0 - 1, nz);
        os = ld_plane(PLANE(s_in, zo), o2col);
        ol = ld_plane(PLANE(l_in, zo), o2col);
        oc = ld_plane(PLANE(c_in, zo), o2col);
    }
    for (int zs = z0 - 1; zs <= zend; ++zs) {
        PPF_FRESH("+v"(col), "+v"(ocol), "+v"(rcol), "+v"(o2col));
        const int zc = ghost_any<BC>(zs, nz), zn = ghost_any<BC>(zs + 1, nz);
        // this plane's A1 and A2 buffers, and the A2 buffer of plane zs-1
        const int bsel = DB ? ((zs - z0 + 1) & 1) : 0;
        V *const A1 = v1 + bsel * (NG * S1);
        V *const A2 = v2 + bsel * (NG * S2);
        const V *const A2p = v2 + (DB ? bsel ^ 1 : 0) * (NG * S2);
        // (without PF: this plane's outer site, loaded by the previous plane.
        // OPRE: the next plane's is loaded with this plane's columns, so that
        // it is not still in flight when the barrier has passed (in FP32: 249
        // to 261 GLUPS for the 256-member ensemble); FP64 loads it after
        // writing A1, which holds three fewer values live across the loads
        // and measured 1% faster)
        constexpr bool OPRE = !PF && sizeof(T) == 4;
        const T cos_ = os, col_ = ol, coc_ = oc;
        if (!PF) {
            rs_n = ld_plane(PLANE(s_in, zn), col);
            rl_n = ld_plane(PLANE(l_in, zn), col);
            rc_n = ld_plane(PLANE(c_in, zn), col);
            if (ring) {
                qs_n = ld_plane(PLANE(s_in, zn), rcol);
                ql_n = ld_plane(PLANE(l_in, zn), rcol);
                qc_n = ld_plane(PLANE(c_in, zn), rcol);
            }
            if (OPRE && outer) {
                os = ld_plane(PLANE(s_in, zn), o2col);
                ol = ld_plane(PLANE(l_in, zn), o2col);
                oc = ld_plane(PLANE(c_in, zn), o2col);
            }
        }
        // PF: the next iteration's values (plane zs+2 of the columns, zs+1
        // of the outer site)
        T fs = 0, fl = 0, fc = 0, fqs = 0, fql = 0, fqc = 0, fos = 0, fol = 0, foc = 0;
        if (PF && zs < zend) {
            const int zf = ghost_any<BC>(zs + 2, nz);
            fs = ld_plane(PLANE(s_in, zf), col);
            fl = ld_plane(PLANE(l_in, zf), col);
            fc = ld_plane(PLANE(c_in, zf), col);
            if (ring) {
                fqs = ld_plane(PLANE(s_in, zf), rcol);
                fql = ld_plane(PLANE(l_in, zf), rcol);
                fqc = ld_plane(PLANE(c_in, zf), rcol);
            }
            if (outer) {
                fos = ld_plane(PLANE(s_in, zn), o2col);
                fol = ld_plane(PLANE(l_in, zn), o2col);
                foc = ld_plane(PLANE(c_in, zn), o2col);
            }
        }
        // A1: step t of plane zs
        lds_put4<S1>(A1, i1, s_c, l_c, mu_c, M_c);
        if (ring)
            lds_put4<S1>(A1, ir1, qs_c, ql_c, qmu_c, qM_c);
        if (outer) {
            const T xs = OPRE ? cos_ : os, xl = OPRE ? col_ : ol, xc = OPRE ? coc_ : oc;
            T oh, omu, oM;
            P::point(xs, xl, xc, k, oh, omu, oM);
            lds_put4<S1>(A1, io1, xs, xl, omu, oM);
            if (!PF && !OPRE) {
                os = ld_plane(PLANE(s_in, zn), o2col);
                ol = ld_plane(PLANE(l_in, zn), o2col);
                oc = ld_plane(PLANE(c_in, zn), o2col);
            }
        }
        T h_n, mu_n, M_n, qh_n = 0, qmu_n = 0, qM_n = 0;
        if (PF) {
            P::point(rs_n, rl_n, rc_n, k, h_n, mu_n, M_n);
            if (ring)
                P::point(qs_n, ql_n, qc_n, k, qh_n, qmu_n, qM_n);
        }
        __syncthreads();
        if (!PF) {
            P::point(rs_n, rl_n, rc_n, k, h_n, mu_n, M_n);
            if (ring)
                P::point(qs_n, ql_n, qc_n, k, qh_n, qmu_n, qM_n);
        }

        // stage 1: step t+1 of plane zs at the tile site.  Mirror
        // boundaries: a ghost plane (zs = -1 or nz) is not stepped; plane
        // nz's step t+1 is plane nz-1's (plane -1's is set after plane 0).
        const bool zghost = BC == 1 && (zs < 0 || zs >= nz);
        T s1, l1, c1, h1, mu1, M1;
        if constexpr (BC == 1) {
            s1 = us_c;
            l1 = ul_c;
            c1 = uc_c;
            h1 = uh_c;
            mu1 = umu_c;
            M1 = uM_c;
        }
        if (!zghost) {
            T sn[6] = {rs_n, s_p}, ln[6] = {rl_n, l_p}, mun[6] = {mu_n, mu_p}, Mn[6] = {M_n, M_p};
            lds_nbrs<S1, W1>(A1, i1, sn, ln, mun, Mn);
            double ds, dl, dc;
            P::update(s_c, l_c, c_c, h_c, mu_c, M_c, sn, ln, mun, Mn, k, s1, l1, c1, ds, dl, dc);
            P::point(s1, l1, c1, k, h1, mu1, M1);
        }
        // ... and at the ring site
        T qs1 = 0, ql1 = 0, qmu1 = 0, qM1 = 0;
        if (ring && ring_in && !zghost) {
            T sn[6] = {qs_n, qs_p}, ln[6] = {ql_n, ql_p}, mun[6] = {qmu_n, qmu_p},
              Mn[6] = {qM_n, qM_p};
            lds_nbrs<S1, W1>(A1, ir1, sn, ln, mun, Mn);
            T qc1, qh1;
            double ds, dl, dc;
            P::update(qs_c, ql_c, qc_c, qh_c, qmu_c, qM_c, sn, ln, mun, Mn, k, qs1, ql1, qc1, ds, dl, dc);
            P::point(qs1, ql1, qc1, k, qh1, qmu1, qM1);
        }
        // A2: step t+1 of plane zs (a buffer of its own with DB; otherwise
        // after stage 2 and a barrier, below)
        auto store_a2 = [&]() {
            if (BC == 0 || active)
                lds_put4<S2>(A2, i2, s1, l1, mu1, M1);
            if constexpr (BC == 1) {
                // the ghost slots this site mirrors
                if (active) {
                    if (x0 + tx == 0)
                        lds_put4<S2>(A2, i2 - 1, s1, l1, mu1, M1);
                    if (x0 + tx == nx - 1)
                        lds_put4<S2>(A2, i2 + 1, s1, l1, mu1, M1);
                    if (y0 + ty == 0)
                        lds_put4<S2>(A2, i2 - W2, s1, l1, mu1, M1);
                    if (y0 + ty == ny - 1)
                        lds_put4<S2>(A2, i2 + W2, s1, l1, mu1, M1);
                }
            }
            if (ring && ring_in)
                lds_put4<S2>(A2, ir2, qs1, ql1, qmu1, qM1);
        };
        if (DB)
            store_a2();
        // stage 2: step t+2 of plane zs-1 at the tile site (A2p holds step
        // t+1 of plane zs-1)
        if (zs >= z0 + 1 && active) {
            T sn[6] = {s1, us_p}, ln[6] = {l1, ul_p}, mun[6] = {mu1, umu_p}, Mn[6] = {M1, uM_p};
            lds_nbrs<S2, W2>(A2p, i2, sn, ln, mun, Mn);
            T so, lo, co;
            double ds, dl, dc;
            P::update(us_c, ul_c, uc_c, uh_c, umu_c, uM_c, sn, ln, mun, Mn, k, so, lo, co, ds, dl, dc);
            __builtin_nontemporal_store(so, at_plane(PLANE(s_out, zs - 1), ocol));
            __builtin_nontemporal_store(lo, at_plane(PLANE(l_out, zs - 1), ocol));
            __builtin_nontemporal_store(co, at_plane(PLANE(c_out, zs - 1), ocol));
        }
        if (!DB) {
            __syncthreads();
            store_a2();
        }
        // shift the columns
        s_p = s_c;
        l_p = l_c;
        mu_p = mu_c;
        M_p = M_c;
        s_c = rs_n;
        l_c = rl_n;
        c_c = rc_n;
        h_c = h_n;
        mu_c = mu_n;
        M_c = M_n;
        qs_p = qs_c;
        ql_p = ql_c;
        qmu_p = qmu_c;
        qM_p = qM_c;
        qs_c = qs_n;
        ql_c = ql_n;
        qc_c = qc_n;
        qh_c = qh_n;
        qmu_c = qmu_n;
        qM_c = qM_n;
        us_p = us_c;
        ul_p = ul_c;
        umu_p = umu_c;
        uM_p = uM_c;
        us_c = s1;
        ul_c = l1;
        uc_c = c1;
        uh_c = h1;
        umu_c = mu1;
        uM_c = M1;
        if (BC == 1 && z0 == 0 && zs == 0) {  // (compiled out for periodic grids)
            // mirror boundary below plane 0: ghost plane -1 holds plane 0
            us_p = us_c;
            ul_p = ul_c;
            umu_p = umu_c;
            uM_p = uM_c;
        }
        if (PF) {
            rs_n = fs;
            rl_n = fl;
            rc_n = fc;
            qs_n = fqs;
            ql_n = fql;
            qc_n = fqc;
            os = fos;
            ol = fol;
            oc = foc;
        }
    }
    }
#undef PLANE
}

// Two steps per launch with the ring in waves of its own (periodic grids,
// double-buffered LDS).  The same march as tile_march2<P, TX, TY, false,
// true>, split by wave: the TX*TY tile threads (whole waves) update their
// tile site in stage 1 and stage 2; the ring waves that follow (NRL lanes,
// NR1 of them with a one-cell ring site) load the two-cell ring and corners
// (the outer sites) and update the ring sites in stage 1.  The two kinds
// of wave run separate loops over the same planes with one barrier per
// plane in each, so a wave holds only its own column state (fewer
// registers) and no wave does the tile's and a ring site's update in the
// same plane.  Every site sees the same operands in the same order as in
// tile_march2, so the results are bit for bit the same.
template <class P, int TX, int TY>
__device__ __forceinline__ void tile_march2s(
    const typename P::T *__restrict__ s_in, const typename P::T *__restrict__ l_in,
    const typename P::T *__restrict__ c_in, typename P::T *__restrict__ s_out,
    typename P::T *__restrict__ l_out, typename P::T *__restrict__ c_out, int nx, int ny, int nz,
    int x0, int y0, int z0, int zend, const typename P::K &k,
    typename P::T (*__restrict__ a1)[TY + 4][TX + 4], typename P::T (*__restrict__ a2)[TY + 2][TX + 2])
{
    constexpr int NT = TX * TY, NR1 = 2 * TX + 2 * TY, NO2 = 2 * TX + 2 * TY + 4;
    constexpr int NRL = (NR1 + 63) / 64 * 64;
    static_assert(NT % 64 == 0, "tile threads must fill whole waves");
    static_assert(NO2 <= 2 * NRL, "two outer sites per ring lane at most");
    using T = typename P::T;
    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const size_t plane = (size_t)nx * ny;
#define PLANE(p, zz) uniform_ptr((p) + (size_t)(zz) * plane)
    auto colof = [&](int ax, int ay) {
        return (unsigned)ghost_any<0>(y0 + ay, ny) * (unsigned)nx + (unsigned)ghost_any<0>(x0 + ax, nx);
    };
    if (wave < NT / 64) {
        // ---- tile waves ----
        const int tx = tid % TX, ty = tid / TX;
        const bool active = x0 + tx < nx && y0 + ty < ny;
        unsigned col = colof(tx, ty);
        unsigned ocol = (unsigned)(y0 + ty) * (unsigned)nx + (unsigned)(x0 + tx);
        T s_p, l_p, mu_p, M_p, s_c, l_c, c_c, h_c, mu_c, M_c;
        {
            const int zp = ghost_any<0>(z0 - 2, nz), zc = ghost_any<0>(z0 - 1, nz);
            T hp;
            s_p = ld_plane(PLANE(s_in, zp), col);
            l_p = ld_plane(PLANE(l_in, zp), col);
            P::point(s_p, l_p, ld_plane(PLANE(c_in, zp), col), k, hp, mu_p, M_p);
            s_c = ld_plane(PLANE(s_in, zc), col);
            l_c = ld_plane(PLANE(l_in, zc), col);
            c_c = ld_plane(PLANE(c_in, zc), col);
            P::point(s_c, l_c, c_c, k, h_c, mu_c, M_c);
        }
        T us_p = 0, ul_p = 0, umu_p = 0, uM_p = 0, us_c = 0, ul_c = 0, uc_c = 0, uh_c = 0,
          umu_c = 0, uM_c = 0;
        for (int zs = z0 - 1; zs <= zend; ++zs) {
            PPF_FRESH("+v"(col), "+v"(ocol));
            const int zn = ghost_any<0>(zs + 1, nz);
            const int bsel = (zs - z0 + 1) & 1;
            T(*__restrict__ A1)[TY + 4][TX + 4] = a1 + 4 * bsel;
            T(*__restrict__ A2)[TY + 2][TX + 2] = a2 + 4 * bsel;
            const T(*__restrict__ A2p)[TY + 2][TX + 2] = a2 + 4 * (bsel ^ 1);
            const T rs_n = ld_plane(PLANE(s_in, zn), col);
            const T rl_n = ld_plane(PLANE(l_in, zn), col);
            const T rc_n = ld_plane(PLANE(c_in, zn), col);
            A1[0][ty + 2][tx + 2] = s_c;
            A1[1][ty + 2][tx + 2] = l_c;
            A1[2][ty + 2][tx + 2] = mu_c;
            A1[3][ty + 2][tx + 2] = M_c;
            T h_n, mu_n, M_n;
            P::point(rs_n, rl_n, rc_n, k, h_n, mu_n, M_n);
            __syncthreads();
            // stage 1: step t+1 of plane zs at the tile site
            T s1, l1, c1, h1, mu1, M1;
            {
                const int ax = tx + 2, ay = ty + 2;
                const T sn[6] = {rs_n, s_p, A1[0][ay + 1][ax], A1[0][ay - 1][ax], A1[0][ay][ax + 1], A1[0][ay][ax - 1]};
                const T ln[6] = {rl_n, l_p, A1[1][ay + 1][ax], A1[1][ay - 1][ax], A1[1][ay][ax + 1], A1[1][ay][ax - 1]};
                const T mun[6] = {mu_n, mu_p, A1[2][ay + 1][ax], A1[2][ay - 1][ax], A1[2][ay][ax + 1], A1[2][ay][ax - 1]};
                const T Mn[6] = {M_n, M_p, A1[3][ay + 1][ax], A1[3][ay - 1][ax], A1[3][ay][ax + 1], A1[3][ay][ax - 1]};
                double ds, dl, dc;
                P::update(s_c, l_c, c_c, h_c, mu_c, M_c, sn, ln, mun, Mn, k, s1, l1, c1, ds, dl, dc);
                P::point(s1, l1, c1, k, h1, mu1, M1);
            }
            A2[0][ty + 1][tx + 1] = s1;
            A2[1][ty + 1][tx + 1] = l1;
            A2[2][ty + 1][tx + 1] = mu1;
            A2[3][ty + 1][tx + 1] = M1;
            // stage 2: step t+2 of plane zs-1 at the tile site
            if (zs >= z0 + 1 && active) {
                const int bx = tx + 1, by = ty + 1;
                const T sn[6] = {s1, us_p, A2p[0][by + 1][bx], A2p[0][by - 1][bx], A2p[0][by][bx + 1], A2p[0][by][bx - 1]};
                const T ln[6] = {l1, ul_p, A2p[1][by + 1][bx], A2p[1][by - 1][bx], A2p[1][by][bx + 1], A2p[1][by][bx - 1]};
                const T mun[6] = {mu1, umu_p, A2p[2][by + 1][bx], A2p[2][by - 1][bx], A2p[2][by][bx + 1], A2p[2][by][bx - 1]};
                const T Mn[6] = {M1, uM_p, A2p[3][by + 1][bx], A2p[3][by - 1][bx], A2p[3][by][bx + 1], A2p[3][by][bx - 1]};
                T so, lo, co;
                double ds, dl, dc;
                P::update(us_c, ul_c, uc_c, uh_c, umu_c, uM_c, sn, ln, mun, Mn, k, so, lo, co, ds, dl, dc);
                __builtin_nontemporal_store(so, at_plane(PLANE(s_out, zs - 1), ocol));
                __builtin_nontemporal_store(lo, at_plane(PLANE(l_out, zs - 1), ocol));
                __builtin_nontemporal_store(co, at_plane(PLANE(c_out, zs - 1), ocol));
            }
            s_p = s_c;
            l_p = l_c;
            mu_p = mu_c;
            M_p = M_c;
            s_c = rs_n;
            l_c = rl_n;
            c_c = rc_n;
            h_c = h_n;
            mu_c = mu_n;
            M_c = M_n;
            us_p = us_c;
            ul_p = ul_c;
            umu_p = umu_c;
            uM_p = uM_c;
            us_c = s1;
            ul_c = l1;
            uc_c = c1;
            uh_c = h1;
            umu_c = mu1;
            uM_c = M1;
        }
    } else {
        // ---- ring waves ----
        const int r = tid - NT;
        const bool ring = r < NR1;
        // one-cell ring site (A2 frame rx, ry; tile frame rx-1, ry-1)
        int rx = 0, ry = 0;
        if (r < TY) {
            rx = 0;
            ry = r + 1;
        } else if (r < 2 * TY) {
            rx = TX + 1;
            ry = r - TY + 1;
        } else if (r < 2 * TY + TX) {
            rx = r - 2 * TY + 1;
            ry = 0;
        } else if (r < NR1) {
            rx = r - 2 * TY - TX + 1;
            ry = TY + 1;
        }
        unsigned rcol = colof(rx - 1, ry - 1);
        // outer sites j = r and r + NRL (A1 frame ox, oy; tile frame ox-2, oy-2)
        int ox[2] = {0, 0}, oy[2] = {0, 0};
        bool outer[2];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            const int j = r + u * NRL;
            outer[u] = j < NO2;
            if (j < TY) {
                ox[u] = 0;
                oy[u] = j + 2;
            } else if (j < 2 * TY) {
                ox[u] = TX + 3;
                oy[u] = j - TY + 2;
            } else if (j < 2 * TY + TX) {
                ox[u] = j - 2 * TY + 2;
                oy[u] = 0;
            } else if (j < 2 * TY + 2 * TX) {
                ox[u] = j - 2 * TY - TX + 2;
                oy[u] = TY + 3;
            } else if (j < NO2) {
                const int c = j - 2 * TY - 2 * TX;
                ox[u] = (c & 1) ? TX + 2 : 1;
                oy[u] = (c & 2) ? TY + 2 : 1;
            }
        }
        unsigned o2col[2] = {colof(ox[0] - 2, oy[0] - 2), colof(ox[1] - 2, oy[1] - 2)};
        T qs_p = 0, ql_p = 0, qmu_p = 0, qM_p = 0, qs_c = 0, ql_c = 0, qc_c = 0, qh_c = 0, qmu_c = 0,
          qM_c = 0;
        if (ring) {
            const int zp = ghost_any<0>(z0 - 2, nz), zc = ghost_any<0>(z0 - 1, nz);
            qs_p = ld_plane(PLANE(s_in, zp), rcol);
            ql_p = ld_plane(PLANE(l_in, zp), rcol);
            T qh;
            P::point(qs_p, ql_p, ld_plane(PLANE(c_in, zp), rcol), k, qh, qmu_p, qM_p);
            qs_c = ld_plane(PLANE(s_in, zc), rcol);
            ql_c = ld_plane(PLANE(l_in, zc), rcol);
            qc_c = ld_plane(PLANE(c_in, zc), rcol);
            P::point(qs_c, ql_c, qc_c, k, qh_c, qmu_c, qM_c);
        }
        for (int zs = z0 - 1; zs <= zend; ++zs) {
            PPF_FRESH("+v"(rcol), "+v"(o2col[0]), "+v"(o2col[1]));
            const int zc = ghost_any<0>(zs, nz), zn = ghost_any<0>(zs + 1, nz);
            const int bsel = (zs - z0 + 1) & 1;
            T(*__restrict__ A1)[TY + 4][TX + 4] = a1 + 4 * bsel;
            T(*__restrict__ A2)[TY + 2][TX + 2] = a2 + 4 * bsel;
            T qs_n = 0, ql_n = 0, qc_n = 0;
            if (ring) {
                qs_n = ld_plane(PLANE(s_in, zn), rcol);
                ql_n = ld_plane(PLANE(l_in, zn), rcol);
                qc_n = ld_plane(PLANE(c_in, zn), rcol);
            }
            T os[2], ol[2], oc[2];
#pragma unroll
            for (int u = 0; u < 2; u++) {
                os[u] = ol[u] = oc[u] = 0;
                if (outer[u]) {
                    os[u] = ld_plane(PLANE(s_in, zc), o2col[u]);
                    ol[u] = ld_plane(PLANE(l_in, zc), o2col[u]);
                    oc[u] = ld_plane(PLANE(c_in, zc), o2col[u]);
                }
            }
            if (ring) {
                A1[0][ry + 1][rx + 1] = qs_c;
                A1[1][ry + 1][rx + 1] = ql_c;
                A1[2][ry + 1][rx + 1] = qmu_c;
                A1[3][ry + 1][rx + 1] = qM_c;
            }
#pragma unroll
            for (int u = 0; u < 2; u++) {
                if (outer[u]) {
                    T oh, omu, oM;
                    P::point(os[u], ol[u], oc[u], k, oh, omu, oM);
                    A1[0][oy[u]][ox[u]] = os[u];
                    A1[1][oy[u]][ox[u]] = ol[u];
                    A1[2][oy[u]][ox[u]] = omu;
                    A1[3][oy[u]][ox[u]] = oM;
                }
            }
            T qh_n = 0, qmu_n = 0, qM_n = 0;
            if (ring)
                P::point(qs_n, ql_n, qc_n, k, qh_n, qmu_n, qM_n);
            __syncthreads();
            // stage 1: step t+1 of plane zs at the ring site
            if (ring) {
                const int ax = rx + 1, ay = ry + 1;
                const T sn[6] = {qs_n, qs_p, A1[0][ay + 1][ax], A1[0][ay - 1][ax], A1[0][ay][ax + 1], A1[0][ay][ax - 1]};
                const T ln[6] = {ql_n, ql_p, A1[1][ay + 1][ax], A1[1][ay - 1][ax], A1[1][ay][ax + 1], A1[1][ay][ax - 1]};
                const T mun[6] = {qmu_n, qmu_p, A1[2][ay + 1][ax], A1[2][ay - 1][ax], A1[2][ay][ax + 1], A1[2][ay][ax - 1]};
                const T Mn[6] = {qM_n, qM_p, A1[3][ay + 1][ax], A1[3][ay - 1][ax], A1[3][ay][ax + 1], A1[3][ay][ax - 1]};
                T qs1, ql1, qc1, qh1, qmu1, qM1;
                double ds, dl, dc;
                P::update(qs_c, ql_c, qc_c, qh_c, qmu_c, qM_c, sn, ln, mun, Mn, k, qs1, ql1, qc1, ds, dl, dc);
                P::point(qs1, ql1, qc1, k, qh1, qmu1, qM1);
                A2[0][ry][rx] = qs1;
                A2[1][ry][rx] = ql1;
                A2[2][ry][rx] = qmu1;
                A2[3][ry][rx] = qM1;
            }
            qs_p = qs_c;
            ql_p = ql_c;
            qmu_p = qmu_c;
            qM_p = qM_c;
            qs_c = qs_n;
            ql_c = ql_n;
            qc_c = qc_n;
            qh_c = qh_n;
            qmu_c = qmu_n;
            qM_c = qM_n;
        }
    }
#undef PLANE
}

// Kernel of tile_march2s: TX*TY tile threads and the ring waves after them.
template <class P, int TX, int TY, int WPE>
__global__ __launch_bounds__(TX *TY + (2 * TX + 2 * TY + 63) / 64 * 64)
__attribute__((amdgpu_waves_per_eu(WPE, WPE))) void k_tiled2s(
    const typename P::T *__restrict__ s_in, const typename P::T *__restrict__ l_in,
    const typename P::T *__restrict__ c_in, typename P::T *__restrict__ s_out,
    typename P::T *__restrict__ l_out, typename P::T *__restrict__ c_out, int nx, int ny, int nz,
    int tz, int ntx, int nty, int order, typename P::K k)
{
    using T = typename P::T;
    __shared__ T a1[8][TY + 4][TX + 4];
    __shared__ T a2[8][TY + 2][TX + 2];
    const int ntile = ntx * nty;
    const int nblk = gridDim.x;
    int b = blockIdx.x;
    if ((order & 2) && (nblk & 7) == 0)
        b = (b & 7) * (nblk >> 3) + (b >> 3);
    const int tile = (order & 1) ? b / ((nblk + ntile - 1) / ntile) : b % ntile;
    const int zc = (order & 1) ? b % ((nblk + ntile - 1) / ntile) : b / ntile;
    const int x0 = (tile % ntx) * TX, y0 = (tile / ntx) * TY, z0 = zc * tz;
    const int zend = min(z0 + tz, nz);
    tile_march2s<P, TX, TY>(s_in, l_in, c_in, s_out, l_out, c_out, nx, ny, nz, x0, y0, z0, zend, k,
                            a1, a2);
}

// WPE: the waves per SIMD the register allocation aims for (FP64 needs a
// cap: left alone it takes ~185 VGPRs, two waves per SIMD).  PF: prefetch
// one plane ahead, DB: double-buffered LDS, one barrier per plane, FULL:
// tiles that divide the plane, deferred output stores (tile_march2).
template <class P, int TX, int TY, int WPE, bool PF = false, bool DB = false, int BC = 0,
          bool FULL = false>
__global__ __launch_bounds__(TX *TY) __attribute__((amdgpu_waves_per_eu(WPE, WPE))) void k_tiled2(
    const typename P::T *__restrict__ s_in, const typename P::T *__restrict__ l_in,
    const typename P::T *__restrict__ c_in, typename P::T *__restrict__ s_out,
    typename P::T *__restrict__ l_out, typename P::T *__restrict__ c_out, int nx, int ny, int nz,
    int tz, int ntx, int nty, int order, typename P::K k)
{
    using T = typename P::T;
    __shared__ __align__(16) T a1[(DB ? 8 : 4)][TY + 4][TX + 4];
    __shared__ __align__(16) T a2[(DB ? 8 : 4)][TY + 2][TX + 2];
    const int ntile = ntx * nty;
    const int nblk = gridDim.x;
    int b = blockIdx.x;
    if ((order & 2) && (nblk & 7) == 0)
        b = (b & 7) * (nblk >> 3) + (b >> 3);
    const int tile = (order & 1) ? b / ((nblk + ntile - 1) / ntile) : b % ntile;
    const int zc = (order & 1) ? b % ((nblk + ntile - 1) / ntile) : b / ntile;
    const int x0 = (tile % ntx) * TX, y0 = (tile / ntx) * TY, z0 = zc * tz;
    const int zend = min(z0 + tz, nz);
    tile_march2<P, TX, TY, PF, DB, BC, P::kRotate, FULL>(s_in, l_in, c_in, s_out, l_out, c_out, nx,
                                                         ny, nz, x0, y0,
                                       z0, zend, k, a1, a2);
}
