// k_step.h -- step kernels: the naive FP64 form and the tiled form for any
// physics policy (PhysF64: bitwise equal to the reference; PhysF32).
//
// One launch = one sweep of phasefield.c:301-403 over the whole grid, with
// the ghost fill of phasefield.c:1112-1189 replaced by the index map
// ghost<BC>.  Input and output are separate buffers (ping-pong); the
// reference updates in place after computing every increment from the old
// state (pass C, phasefield.c:393-401), which is the same thing.
#pragma once

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
// flight while it computes (more registers, fewer waves per SIMD).  Compulsory HBM traffic is one read and one write of
// each field per site; ring reads and the extra planes of each tz chunk
// come mostly from L2.
// ---------------------------------------------------------------------------
// One workgroup's march over the column tile [x0, x0+TX) x [y0, y0+TY) and
// the planes [z0, zend) of one grid.  sh is the workgroup's LDS,
// 2 x 4 x (TY+2) x (TX+2) values; acc gets the REPORT sums of squared
// increments.  Every thread of the workgroup must call it.
template <class P, int TX, int TY, int BC, bool CONLY, bool REPORT, bool PIPE>
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
    const unsigned col = (unsigned)ghost_any<BC>(y, ny) * (unsigned)nx + (unsigned)ghost_any<BC>(x, nx);
    const unsigned ocol = (unsigned)y * (unsigned)nx + (unsigned)x;

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
    const unsigned rcol = (unsigned)ghost_any<BC>(y0 + ry - 1, ny) * (unsigned)nx +
                          (unsigned)ghost_any<BC>(x0 + rx - 1, nx);

    // Prologue: planes z0-1 and z0 computed, plane z0+1 and the ring of z0
    // loaded raw.
    T s_p, l_p, mu_p, M_p;
    T s_c, l_c, c_c, h_c, mu_c, M_c;
    T rs_n, rl_n, rc_n;                // raw column values one plane ahead
    T rs_r = 0, rl_r = 0, rc_r = 0;    // raw ring values of the current plane
    {
        const int zp = ghost<BC>(z0 - 1, nz), zn = ghost_any<BC>(z0 + 1, nz);
        s_p = PLANE(s_in, zp)[col];
        l_p = PLANE(l_in, zp)[col];
        const T c_p = PLANE(c_in, zp)[col];
        s_c = PLANE(s_in, z0)[col];
        l_c = PLANE(l_in, z0)[col];
        c_c = PLANE(c_in, z0)[col];
        rs_n = PLANE(s_in, zn)[col];
        rl_n = PLANE(l_in, zn)[col];
        rc_n = PLANE(c_in, zn)[col];
        if (ring) {
            rs_r = PLANE(s_in, z0)[rcol];
            rl_r = PLANE(l_in, z0)[rcol];
            rc_r = PLANE(c_in, z0)[rcol];
        }
        T hp;
        P::point(s_p, l_p, c_p, k, hp, mu_p, M_p);
        P::point(s_c, l_c, c_c, k, h_c, mu_c, M_c);
    }

    for (int z = z0; z < zend; ++z) {
        const int buf = (z - z0) & 1;
        const bool more = z + 1 < zend;
        if (!PIPE && z > z0) {
            // unpipelined form: this plane's ring and the next plane now
            const int zn = ghost_any<BC>(z + 1, nz);
            rs_n = PLANE(s_in, zn)[col];
            rl_n = PLANE(l_in, zn)[col];
            rc_n = PLANE(c_in, zn)[col];
            if (ring) {
                rs_r = PLANE(s_in, z)[rcol];
                rl_r = PLANE(l_in, z)[rcol];
                rc_r = PLANE(c_in, z)[rcol];
            }
        }
        // prefetch: column two planes ahead, ring one plane ahead
        T ps = 0, pl = 0, pc = 0, prs = 0, prl = 0, prc = 0;
        if (PIPE && more) {
            const int zf = ghost_any<BC>(z + 2, nz);
            ps = PLANE(s_in, zf)[col];
            pl = PLANE(l_in, zf)[col];
            pc = PLANE(c_in, zf)[col];
            if (ring) {
                prs = PLANE(s_in, z + 1)[rcol];
                prl = PLANE(l_in, z + 1)[rcol];
                prc = PLANE(c_in, z + 1)[rcol];
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
                __builtin_nontemporal_store((T)(c_c + dc), PLANE(c_out, z) + ocol);
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
                __builtin_nontemporal_store(so, PLANE(s_out, z) + ocol);
                __builtin_nontemporal_store(lo, PLANE(l_out, z) + ocol);
                __builtin_nontemporal_store(co, PLANE(c_out, z) + ocol);
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
            rs_n = ps;
            rl_n = pl;
            rc_n = pc;
            rs_r = prs;
            rl_r = prl;
            rc_r = prc;
        }
    }
#undef PLANE
}

template <class P, int TX, int TY, int BC, bool CONLY, bool REPORT, bool PIPE>
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
    tile_march<P, TX, TY, BC, CONLY, REPORT, PIPE>(s_in, l_in, c_in, s_out, l_out, c_out, nx, ny,
                                                   nz, x0, y0, z0, zend, k, sh, acc);
    if (REPORT)
        block_sum3<NT>(acc, partials + 3 * (size_t)blockIdx.x);
}
